"""GPU parity for transform lengths nkx = padx * nmx that are not powers of two 128..4096.

The reference transforms any nkx with FFTW (wem.c:100-104).  The 2-D kernels take lengths
2^a 3^b 5^c up to 4096 through the mixed-radix transform on the next power-of-two array, and the
other lengths up to 2048 through Bluestein's chirp-z identity on the power-of-two array length
M >= 2 nkx - 1.  Tolerances are the power-of-two ones: relative L2 <= 1e-4 against the FP64 oracle,
dot test <= 1e-5.  The oracle's own transform handles every length (including the prime factors
37, 101, 151, 509 and 1013 below).
"""
import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import synth
from conftest import rel_l2
from test_gpu_parity import oparams

pytestmark = pytest.mark.gpu
TOL = 1e-4
DOT_TOL = 1e-5
CASES = [(False, 0), (True, 0), (True, 1)]


def case(pspi: bool, nx: int, padx: int, nz: int, nt: int, seed: int = 1):
    g = seiswem.WemGeometry(nt=nt, dt=0.004, nmx=nx, dmx=10.0, nz=nz, dz=10.0, fmin=2.0, fmax=40.0, nref=4,
                            pspi=pspi, padx=padx)
    rng = np.random.default_rng(seed)
    vel = synth.blob_velocity(nx, nz, g.dmx, g.dz, rng, nblobs=4, v0=2000.0)
    wav = synth.wavelet_trace(nt, g.dt, 15.0)
    return g, vel, wav, rng


# (nmx, padx) -> nkx, array length M.  Bluestein: nkx with a prime factor above 5.
SIZES_BLUE = [
    (37, 2),    # nkx 74 = 2 * 37, M 256
    (101, 2),   # nkx 202 = 2 * 101, M 512
    (151, 2),   # nkx 302 = 2 * 151, M 1024
    (509, 2),   # nkx 1018 = 2 * 509, M 2048
    (1013, 2),  # nkx 2026 = 2 * 1013, M 4096
]
# mixed radix: nkx = 2^a 3^b 5^c, M the next power of two
SIZES_MIXED = [
    (64, 1),    # nkx 64 (a power of two below the power-of-two kernels), M 128
    (150, 2),   # nkx 300 = 4 3 25, M 512
    (500, 2),   # nkx 1000 = 8 125, M 1024
    (1000, 2),  # nkx 2000 = 16 125, M 2048
    (1500, 2),  # nkx 3000 = 8 3 125, M 4096 (beyond Bluestein's reach)
]
SIZES = [pytest.param(*c, id=f"blue{c[0] * c[1]}") for c in SIZES_BLUE] + \
    [pytest.param(*c, id=f"mixed{c[0] * c[1]}") for c in SIZES_MIXED]


def _dims(nx):
    # oracle cost grows with nx * nz * nt: keep the largest cases short
    return (24, 128) if nx <= 151 else (10, 64)


@pytest.mark.parametrize("nx,padx", SIZES)
@pytest.mark.parametrize("pspi,mode", CASES)
def test_migrate_bluestein(gpu, pspi, mode, nx, padx):
    nz, nt = _dims(nx)
    g, vel, wav, rng = case(pspi, nx, padx, nz, nt)
    sx = g.dmx * (nx // 3)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    m_ref = O.wem(oparams(g, mode), vel, wav, sx, d=d, adj=True)
    m_gpu = seiswem.ShotProfileWEM(d[None], True, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode)
    err = rel_l2(m_gpu, m_ref)
    print(f"bluestein migrate pspi={pspi} mode={mode} nx={nx} padx={padx}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,padx", SIZES)
@pytest.mark.parametrize("pspi,mode", CASES)
def test_model_bluestein(gpu, pspi, mode, nx, padx):
    nz, nt = _dims(nx)
    g, vel, wav, rng = case(pspi, nx, padx, nz, nt, seed=3)
    sx = g.dmx * (nx // 2)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d_ref = O.wem(oparams(g, mode), vel, wav, sx, m=m, adj=False)
    d_gpu = seiswem.ShotProfileWEM(m, False, vel=vel, wav=wav, geom=g, sx=[sx], pspi_mode=mode)[0]
    err = rel_l2(d_gpu, d_ref)
    print(f"bluestein model pspi={pspi} mode={mode} nx={nx} padx={padx}: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("nx,padx", [(37, 2), (150, 2), (1500, 2)])
@pytest.mark.parametrize("pspi,mode", [(False, 0), (True, 1)])
def test_dot_product_bluestein(gpu, pspi, mode, nx, padx):
    """<d, L m> == <L^H d, m>: the Bluestein (nkx 74) and mixed-radix (300, 3000) steps keep SSF
    and linear PSPI an adjoint pair."""
    g, vel, wav, rng = case(pspi, nx, padx, 20, 128, seed=5)
    sx = g.dmx * (nx // 2)
    plan = seiswem.Plan(g, vel, wav, pspi_mode=mode)
    m = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    d = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
    lm = plan.model(seiswem.DeviceArray(m.shape, m), [sx]).download()
    lhd = plan.migrate(seiswem.DeviceArray(d.shape, d), [sx]).download()
    a = float(np.sum(d.astype(np.float64) * lm))
    b = float(np.sum(m.astype(np.float64) * lhd))
    print(f"bluestein dot pspi={pspi} mode={mode} nx={nx}: rel {abs(a - b) / abs(b):.3e}", flush=True)
    assert abs(a - b) / abs(b) <= DOT_TOL


@pytest.mark.parametrize("nx", [101, 150], ids=["blue202", "mixed300"])
def test_bluestein_shots_and_chunks(gpu, nx):
    """Several shots in one call at a Bluestein and a mixed-radix length, with the band split in
    frequency chunks by a phase-table cap: the stack equals the single-chunk plan and the sum of
    the oracle's single-shot images."""
    g, vel, wav, rng = case(True, nx, 2, 20, 128, seed=7)
    sx = [200.0, 500.0, 800.0]
    s = g.sampling()
    nf, nkh = s["ifmax"] - s["ifmin"], s["nkx"] // 2 + 1
    per_freq = 8 * g.nz * g.nref * nkh
    d = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    dd = seiswem.DeviceArray(d.shape, d)
    whole = seiswem.Plan(g, vel, wav)
    chunked = seiswem.Plan(g, vel, wav, table_bytes=per_freq * (nf // 3 + 1))
    m1 = whole.migrate(dd, sx).download()
    m2 = chunked.migrate(dd, sx).download()
    assert rel_l2(m2, m1) <= 1e-6
    ref = sum(O.wem(oparams(g, seiswem.PSPI_COMPAT), vel, wav, x, d=d[i], adj=True) for i, x in enumerate(sx))
    err = rel_l2(m2, ref)
    print(f"nkx {2 * nx}: 3 shots, chunked: rel-L2 {err:.3e}", flush=True)
    assert err <= TOL
