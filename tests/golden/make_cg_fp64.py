"""Writes tests/golden/cg_fp64_config5.npz: the FP64 end-to-end reference of the least-squares
migration workload, computed with the FP64 oracle (oracle/, a C restatement of wem.c).

Workload: the config-1 geometry (nx=256, dx=10; nz=200, dz=10; nt=1000, dt=4 ms; 2-60 Hz;
v(z)=1500+0.5 z; 20 Hz Ricker) with 32 shots at sx = 40 + 80 k.  Observed records: FP64 oracle
demigration of the config-1 reflectivity.  10 iterations of ConjugateGradients
(src/Solvers/ConjugateGradients.jl:1-39, mu = 0) with the FP64 oracle as the operator pair and
double-precision inner products.  Stored: the final image (float64), the cost curve, and a
fingerprint of the observed records (sum and sum of squares) so a consumer can check it
rebuilt the same data.  Run from the repository root: python tests/golden/make_cg_fp64.py
(about half an hour on 8 cores)."""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "seismic.jl_amd")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from seiswem import synth  # noqa: E402
from seiswem.solvers import ConjugateGradients, InnerProduct64  # noqa: E402


def params(g):
    return O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax,
                       padt=g.padt, padx=g.padx, pspi=False)


def main():
    g, vel, refl, wav, _ = synth.config1()
    sx = synth.config5_shots()
    p = params(g)
    t0 = time.time()
    d_obs = O.shots(p, vel, wav, sx, m=refl, adj=False, f64=True)
    print(f"records {time.time() - t0:.1f} s", flush=True)

    calls = [0]

    def op(x, adj):
        calls[0] += 1
        t = time.time()
        y = O.shots(p, vel, wav, sx, d=x, adj=True, f64=True) if adj else O.shots(p, vel, wav, sx, m=x, adj=False,
                                                                                f64=True)
        print(f"operator call {calls[0]} ({'adj' if adj else 'fwd'}) {time.time() - t:.1f} s", flush=True)
        return y

    m, cost = ConjugateGradients(d_obs, [op], [{}], Niter=10, mu=0.0, inner=InnerProduct64)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_fp64_config5.npz")
    np.savez_compressed(out, image=m, cost=np.array(cost), d_sum=d_obs.sum(), d_sumsq=(d_obs ** 2).sum(),
                        sx=sx, niter=10)
    print("cost", cost, flush=True)
    print(f"wrote {out} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
