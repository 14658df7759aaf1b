"""GPU test of the least-squares migration workload at its own size: the config-1 geometry (nx 256,
nz 200, nt 1000, 2-60 Hz, v(z) = 1500 + 0.5 z) with 32 shots at sx = 40 + 80 k, 10 iterations of
ConjugateGradients (src/Solvers/ConjugateGradients.jl:1-39, mu = 0) with the GPU operator pair and
model, data and residual resident on the device, for the two GPU precision variants (FP32
propagation with FP64 kz tables and compensated image sums; naive FP32).  Reference: the FP64
end-to-end CG of tests/golden/cg_fp64_config5.npz (written by tests/golden/make_cg_fp64.py with the
FP64 oracle as the operator).  Checked: the image within 1e-4 rel-L2, the normalised misfit curve
within 1e-4 relative (1e-3 for naive FP32), and the dot test of the operator pair within 1e-5."""
import os

import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import solvers, synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_fp64_config5.npz")


@pytest.fixture(scope="module")
def workload():
    ref = np.load(GOLDEN)
    g, vel, refl, wav, _ = synth.config1()
    sx = synth.config5_shots()
    assert np.array_equal(sx, ref["sx"]) and int(ref["niter"]) == 10
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax, padt=g.padt,
                    padx=g.padx, pspi=False)
    d_obs = O.shots(p, vel, wav, sx, m=refl, adj=False, f64=True)
    # the records the golden CG was run on (fingerprint: sum and sum of squares)
    assert abs((d_obs ** 2).sum() / ref["d_sumsq"] - 1) <= 1e-9
    assert abs(d_obs.sum() - ref["d_sum"]) <= 1e-9 * np.sqrt(ref["d_sumsq"] * d_obs.size)
    return g, vel, wav, sx, d_obs.astype(np.float32), ref


def run_variant(workload, prec):
    g, vel, wav, sx, d_obs, ref = workload
    plan = seiswem.Plan(g, vel, wav, precision=prec)
    m, cost = solvers.ConjugateGradientsDevice(plan, seiswem.DeviceArray(d_obs.shape, d_obs), sx, Niter=10)
    img = m.download().astype(np.float64)
    err = rel_l2(img, ref["image"])
    cost_err = float(np.max(np.abs(np.array(cost) / ref["cost"] - 1)))
    rng = np.random.default_rng(5)
    mr = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dr = rng.standard_normal(d_obs.shape).astype(np.float32)
    lm = plan.model(seiswem.DeviceArray(mr.shape, mr), sx).download()
    lhd = plan.migrate(seiswem.DeviceArray(dr.shape, dr), sx).download()
    a, b = float(np.sum(dr.astype(np.float64) * lm)), float(np.sum(mr.astype(np.float64) * lhd))
    plan.close()
    assert len(cost) == 11
    return err, cost_err, abs(a - b) / abs(b)


def test_config5_cg_matches_fp64_reference(gpu, workload):
    """The default variant (FP32 propagation, FP64 kz tables, compensated sums) meets all three
    bars.  The naive FP32 variant meets the image and dot-test bars; its misfit curve departs
    further from the FP64 one (1.35e-4 measured), which is what the FP64 tables buy: it is held
    to 1e-3 and reported."""
    res = {}
    for name, prec in (("fp32_prop_fp64_kz", seiswem.PREC_KZ64), ("naive_fp32", seiswem.PREC_FP32)):
        res[name] = run_variant(workload, prec)
        err, cost_err, dot = res[name]
        print(f"{name}: image rel-L2 {err:.3e}, misfit curve {cost_err:.3e}, dot test {dot:.3e}", flush=True)
    err, cost_err, dot = res["fp32_prop_fp64_kz"]
    assert err <= 1e-4 and cost_err <= 1e-4 and dot <= 1e-5
    err, cost_err, dot = res["naive_fp32"]
    assert err <= 1e-4 and cost_err <= 1e-3 and dot <= 1e-5
