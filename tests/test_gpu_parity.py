"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the same seeded inputs.

Tolerance: relative L2 <= 1e-4 of the image / record against the FP64 oracle (the workload
bound), dot-product tests <= 1e-5.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4
DOT_TOL = 1e-5


def small_case(pspi: bool, seed: int = 1, nx: int = 64, nz: int = 40, nt: int = 256, lateral: bool = True):
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2.0, fmax=40.0, nref=4, pspi=pspi)
    rng = np.random.default_rng(seed)
    if lateral:
        vel = synth.blob_velocity(nx, nz, g.dmx, g.dz, rng, nblobs=4, v0=2000.0)
    else:
        vel = np.repeat((1500.0 + 0.5 * np.arange(nz) * g.dz)[None, :], nx, 0).astype(np.float32)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    return g, vel, wav, rng


def oparams(g: seiswem.WemGeometry, pspi_mode: int = 0) -> O.WemParams:
    return O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, nmy=g.nmy, dmy=g.dmy, ot=g.ot,
                       omx=g.omx, oz=g.oz, gz=g.gz, sz=g.sz, nref=g.nref, fmin=g.fmin, fmax=g.fmax, padt=g.padt,
                       padx=g.padx, pspi=g.pspi, pspi_mode=pspi_mode)


VARIANTS = [seiswem.VARIANT_SUBGROUP, seiswem.VARIANT_WAVE, seiswem.VARIANT_SUBGROUP16, seiswem.VARIANT_PAIR,
            seiswem.VARIANT_PAIR8, seiswem.VARIANT_PACKED]
CASES = [(False, 0), (True, 0), (True, 1)]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("nx", [64, 256, 512])
@pytest.mark.parametrize("pspi,mode", CASES)
def test_migrate_matches_oracle(gpu, pspi, mode, nx, variant):
    """nx=64 pads to nkx=128 with no taper (nmx <= 100); nx=256 / 512 pad to nkx=512 / 1024 with
    the taper (the radix-16 variant needs nkx >= 1024 and falls back below it)."""
    g, vel, wav, rng = small_case(pspi, nx=nx, nz=30 if nx > 64 else 40, nt=256 if nx < 512 else 128)
    sx = 320.0
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g, mode), vel, wav, sx, d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode, variant=variant)
    print(f"migrate pspi={pspi} mode={mode} nx={nx} variant={variant}: rel-L2 {rel_l2(m_gpu, m_ref):.3e}")
    assert rel_l2(m_gpu, m_ref) <= TOL


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("pspi,mode", CASES)
def test_model_matches_oracle(gpu, pspi, mode, variant):
    g, vel, wav, rng = small_case(pspi, seed=3)
    sx = 200.0
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d_ref = O.wem(oparams(g, mode), vel, wav, sx, m=m, adj=False)
    d_gpu = seiswem.ShotProfileWEM(m, False, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode, variant=variant)[0]
    print(f"model pspi={pspi} mode={mode} variant={variant}: rel-L2 {rel_l2(d_gpu, d_ref):.3e}")
    assert rel_l2(d_gpu, d_ref) <= TOL


@pytest.mark.parametrize("pspi,mode", [(False, 0), (True, 1)])
def test_dot_product(gpu, pspi, mode):
    """<d, L m> == <L^H d, m> for the GPU forward/adjoint pair (SSF and linear PSPI)."""
    g, vel, wav, rng = small_case(pspi, seed=5)
    sx = 300.0
    plan = seiswem.Plan(g, vel, wav, pspi_mode=mode)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    dd = seiswem.DeviceArray(d.shape, d)
    md = seiswem.DeviceArray(m.shape, m)
    lm = plan.model(md, [sx]).download()
    lhd = plan.migrate(dd, [sx]).download()
    a = float(np.sum(d.astype(np.float64) * lm))
    b = float(np.sum(m.astype(np.float64) * lhd))
    print(f"dot pspi={pspi} mode={mode}: {a:.9e} {b:.9e} rel {abs(a - b) / abs(b):.3e}")
    assert abs(a - b) / abs(b) <= DOT_TOL


def test_pspi_compat_is_not_adjoint_on_gpu(gpu):
    """The reference's PSPI pair (real-part cut, select after propagation in both directions,
    wem.c:427-460) is not an adjoint pair; the GPU reproduces that, as the oracle does."""
    g, vel, wav, rng = small_case(True, seed=5)
    plan = seiswem.Plan(g, vel, wav, pspi_mode=seiswem.PSPI_COMPAT)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    lm = plan.model(seiswem.DeviceArray(m.shape, m), [300.0]).download()
    lhd = plan.migrate(seiswem.DeviceArray(d.shape, d), [300.0]).download()
    a, b = float(np.sum(d * lm)), float(np.sum(m * lhd))
    assert abs(a - b) / abs(b) > 1e-3


def test_config1_full_size(gpu):
    """Config 1 at full size: GPU demigration and migration against the FP64 oracle."""
    g, vel, refl, wav, sx = synth.config1()
    p = oparams(g)
    d_ref = O.wem(p, vel, wav, float(sx[0]), m=refl, adj=False)
    d_gpu = seiswem.ShotProfileWEM(refl, False, vel=vel, wav=wav, geom=g, sx=sx)[0]
    e_d = rel_l2(d_gpu, d_ref)
    m_ref = O.wem(p, vel, wav, float(sx[0]), d=d_ref, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d_ref[None].astype(np.float32), True, vel=vel, wav=wav, geom=g, sx=sx)
    e_m = rel_l2(m_gpu, m_ref)
    print(f"config1: record rel-L2 {e_d:.3e}, image rel-L2 {e_m:.3e}")
    assert e_d <= TOL and e_m <= TOL
    # kinematics: the three reflectors image at their depth indices with their polarities
    prof = m_gpu[g.nmx // 2]
    for iz, sgn in ((50, 1), (100, -1), (150, 1)):
        win = prof[iz - 3: iz + 4]
        assert np.argmax(np.abs(win)) == 3 and np.sign(win[3]) == sgn


def test_shots_stack_equals_sum_and_each_matches_alone(gpu):
    """Batched shots: the stacked image equals the sum of single-shot images, and the per-shot
    images of a batch are bit-identical to the same shots migrated alone."""
    g, vel, wav, rng = small_case(True, seed=7)
    sxs = np.array([100.0, 250.0, 330.0, 500.0], np.float32)
    d = rng.standard_normal((sxs.size, g.nm, g.nt)).astype(np.float32)
    plan = seiswem.Plan(g, vel, wav)
    dd = seiswem.DeviceArray(d.shape, d)
    each = plan.migrate_each(dd, sxs).download()
    stack = plan.migrate(dd, sxs).download()
    for i, s in enumerate(sxs):
        alone = plan.migrate(seiswem.DeviceArray((1,) + d.shape[1:], d[i:i + 1]), [s]).download()
        assert np.array_equal(alone, each[i])
    assert rel_l2(stack, each.astype(np.float64).sum(0)) <= 1e-6


@pytest.mark.parametrize("variant,tol", [(seiswem.VARIANT_PAIR, 1e-5), (seiswem.VARIANT_PACKED, 2e-5)])
def test_linearity_full_config2_shape(gpu, variant, tol):
    """Size-independent check at the full config-2 shape (nkx=2048, nz=500, 641 frequencies):
    migration is linear, image(a d1 + d2) = a image(d1) + image(d2), to the FP32 rounding of a
    500-step march.  The packed kernel rounds both wavefields in one transform (rounding relative
    to their combined size): its bound is twice PAIR's (measured 1.0e-5)."""
    g, vel, refl, wav, sx = synth.config2(nshot=1)
    plan = seiswem.Plan(g, vel, wav, variant=variant)
    rng = np.random.default_rng(11)
    d1 = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    d2 = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    a = 0.75
    m1 = plan.migrate(seiswem.DeviceArray(d1.shape, d1), sx[:1]).download()
    m2 = plan.migrate(seiswem.DeviceArray(d2.shape, d2), sx[:1]).download()
    m12 = plan.migrate(seiswem.DeviceArray(d1.shape, a * d1 + d2), sx[:1]).download()
    err = rel_l2(m12, a * m1.astype(np.float64) + m2)
    print(f"config2 linearity (variant {variant}) rel-L2 {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("pspi,mode", CASES)
def test_frequency_chunks_match_single_chunk(gpu, pspi, mode):
    """A phase-table cap smaller than the band splits it into frequency chunks whose tables are
    rebuilt per call; migration and modelling agree with the single-chunk plan and the oracle."""
    g, vel, wav, rng = small_case(pspi, seed=9)
    sx = [150.0, 410.0]
    s = g.sampling()
    nf, nkh = s["ifmax"] - s["ifmin"], s["nkx"] // 2 + 1
    per_freq = 8 * g.nz * ((g.nref if pspi else 1) * nkh + (0 if pspi else g.nmx))
    whole = seiswem.Plan(g, vel, wav, pspi_mode=mode)
    chunked = seiswem.Plan(g, vel, wav, pspi_mode=mode, table_bytes=per_freq * (nf // 3 + 1))
    d = rng.standard_normal((2, g.nm, g.nt)).astype(np.float32)
    dd = seiswem.DeviceArray(d.shape, d)
    m1 = whole.migrate(dd, sx).download()
    m2 = chunked.migrate(dd, sx).download()
    e2 = chunked.migrate_each(dd, sx).download()
    assert rel_l2(m2, m1) <= 1e-6
    assert rel_l2(m2, e2.astype(np.float64).sum(0)) <= 1e-6
    alone = chunked.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[1:2]), sx[1:]).download()
    assert np.array_equal(alone, e2[1])
    m_ref = O.wem(oparams(g, mode), vel, wav, sx[0], d=d[0], adj=True)
    assert rel_l2(e2[0], m_ref) <= TOL
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    md = seiswem.DeviceArray(m.shape, m)
    r1 = whole.model(md, sx).download()
    r2 = chunked.model(md, sx).download()
    assert rel_l2(r2, r1) <= 1e-6


@pytest.mark.parametrize("case", ["plateau", "vz", "source_below_surface", "nkx2048", "three_uncut", "five_uncut"])
def test_packed_compat_migration(gpu, case):
    """The packed PSPI compat kernel (both real wavefields in one transform) against the oracle
    and the PAIR kernel: a velocity clipped to a plateau at its maximum (many uncut points per
    depth), a v(z) model (vmax <= vmin + 10 at every depth: every point uncut), a source below
    the surface (not every depth steps the source: the host falls back to PAIR), nkx = 2048
    (the 128-thread subgroup kernel) and exactly 3 / 5 uncut points per depth (the direct sums
    for at most 4, the transforms above)."""
    nx, nz, nt = (1024, 12, 128) if case == "nkx2048" else (256, 30, 256)
    g, vel, wav, rng = small_case(True, seed=5, nx=nx, nz=nz, nt=nt, lateral=case != "vz")
    if case == "plateau":
        vel = np.minimum(vel, np.float32(np.percentile(vel, 90)))
    if case in ("three_uncut", "five_uncut"):  # the depth maximum at 3 / 5 points: sparse / dense paths
        top = vel.max(axis=0) + np.float32(100.0)
        for x in rng.choice(nx, size=3 if case == "three_uncut" else 5, replace=False):
            vel[x] = top
    if case == "source_below_surface":
        g.sz = 55.0
    sx = 640.0
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g, 0), vel, wav, sx, d=d, adj=True)
    m_packed = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=0,
                                      variant=seiswem.VARIANT_PACKED)
    m_pair = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=0,
                                    variant=seiswem.VARIANT_PAIR)
    m_auto = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=0)
    print(f"packed {case}: rel-L2 vs oracle {rel_l2(m_packed, m_ref):.3e}, vs PAIR {rel_l2(m_packed, m_pair):.3e}")
    assert rel_l2(m_packed, m_ref) <= TOL
    assert rel_l2(m_packed, m_pair) <= 1e-5
    assert np.array_equal(m_auto, m_packed)



def test_autotune_shapes_2d_times_group_sizes(gpu):
    """swem_plan_autotune_ex on a 2-D SSF plan: variants with several frequency group sizes are
    timed, the fastest is kept, and migration after tuning still matches the oracle."""
    g, vel, wav, rng = small_case(False, nx=128, nz=30)
    plan = seiswem.Plan(g, vel, wav)
    r = plan.autotune_shapes(nshot=2)
    fpbs = {c["freqs_per_block"] for c in r["candidates"]}
    print(f"2-D shapes: best variant {r['variant']} x {r['freqs_per_block']} in {r['best_ms']:.3f} ms; "
          f"{len(r['candidates'])} candidates, group sizes {sorted(fpbs)}")
    assert len(r["candidates"]) >= 3 and len(fpbs) >= 2
    assert r["best_ms"] == min(c["ms"] for c in r["candidates"])
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[None]), [300.0]).download()
    m_ref = O.wem(oparams(g, 0), vel, wav, 300.0, d=d, adj=True)
    assert rel_l2(m, m_ref) <= TOL


def test_autotune_shapes_3d_and_parity(gpu):
    """3-D: frequencies per y-pass workgroup x step state per launch, then parity with the oracle."""
    g, vel, refl, wav, sx, sy = synth.config3(seed=2, nx=64, ny=64, nz=12, nt=128, fmax=25.0)
    plan = seiswem.Plan(g, vel, wav)
    r = plan.autotune_shapes(nshot=1)
    print(f"3-D shapes: best F={r['freqs_per_block']} chunk {r['chunk_bytes'] >> 20} MiB in {r['best_ms']:.3f} ms; "
          + ", ".join(f"{c['freqs_per_block']}/{c['chunk_bytes'] >> 20}:{c['ms']:.2f}" for c in r["candidates"]))
    assert len(r["candidates"]) == 4 * 3 and r["chunk_bytes"] in (64 << 20, 128 << 20, 256 << 20)
    d = np.random.default_rng(4).standard_normal((g.nm, g.nt)).astype(np.float32)
    m = plan.migrate(seiswem.DeviceArray((1, g.nm, g.nt), d[None]), sx, sy).download()
    po = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nmy=g.nmy, dmy=g.dmy, nz=g.nz, dz=g.dz, fmin=g.fmin,
                     fmax=g.fmax, padt=g.padt, padx=g.padx, pspi=False)
    m_ref = O.wem(po, vel, wav, float(sx[0]), float(sy[0]), d=d, adj=True)
    assert rel_l2(m, m_ref) <= TOL


@pytest.mark.parametrize("pspi,mode", [(True, 0), (False, 0)])
def test_autotune_picks_the_fastest_measured_variant_and_keeps_parity(gpu, pspi, mode):
    """swem_plan_autotune times the distinct kernel variants for the plan's nkx and keeps the
    fastest; migrations after tuning still match the oracle."""
    g, vel, wav, rng = small_case(pspi, nx=256, nz=30)
    plan = seiswem.Plan(g, vel, wav, pspi_mode=mode)
    r = plan.autotune(nshot=2)
    print(f"autotune pspi={pspi}: best {r['best']}, ms {r['ms']}")
    assert r["best"] in r["ms"] and r["ms"][r["best"]] == min(r["ms"].values()) and len(r["ms"]) >= 3
    sx = 320.0
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    m = plan.migrate(seiswem.DeviceArray(d.shape, d), [sx]).download()
    m_ref = O.wem(oparams(g, mode), vel, wav, sx, d=d[0], adj=True)
    assert rel_l2(m, m_ref) <= TOL
