"""GPU test of the file-based drop-in: shotwem.so's `main` with the reference's argv protocol
and .seis files (shotwem.c:12-140), migrating and modelling one shot, checked against the FP64
oracle and the reference's header rewrites (shotwem.c:86-98, 113-117)."""
import ctypes

import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import _lib, synth
from conftest import rel_l2
from test_capi import header_array, write_seis

pytestmark = pytest.mark.gpu


def call_shotwem(args):
    so = ctypes.CDLL(seiswem.SHOTWEM_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    argv = [b"0"] + [a.encode() for a in args]
    return so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv))


def read_seis(name, n1, ntr):
    lib = seiswem.load()
    fh = _lib.SeisExtent()
    data = np.zeros((ntr, n1), np.float32)
    h = header_array(ntr)
    _lib.check(lib.swem_seis_read(name.encode(), data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), h.ctypes.data,
                                  ctypes.byref(fh)))
    return data, h, fh


@pytest.mark.parametrize("pspi", [False, True])
def test_shotwem_main_migrates_and_models_like_the_oracle(gpu, tmp_path, monkeypatch, pspi):
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    lib = seiswem.load()
    nx, nz, nt, dx, dz, dt, ox = 64, 40, 256, 10.0, 10.0, 0.004, 0.0
    rng = np.random.default_rng(2)
    vel = synth.blob_velocity(nx, nz, dx, dz, rng, nblobs=3, v0=2000.0)
    wav = synth.wavelet_trace(nt, dt, 15.0)
    d = rng.standard_normal((nx, nt)).astype(np.float32)
    mx = ox + dx * np.arange(nx)
    write_seis(lib, "vel.seis", vel, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nz, mx=mx, imx=np.arange(nx)),
               nz, nx, o1=0.0, d1=dz, o2=ox, d2=dx)
    write_seis(lib, "wav.seis", wav[None], header_array(1, n1=nt), nt, 1, o1=0.0, d1=dt)
    write_seis(lib, "d.seis", d, header_array(nx, tracenum=np.arange(1, nx + 1), n1=nt, gx=mx, sx=200.0), nt, nx,
               o1=0.0, d1=dt, o2=ox, d2=dx)
    common = [f"pspi={'y' if pspi else 'n'}", "vel=vel.seis", "wav=wav.seis", "sx=200", "fmax=40", "nref=4",
              "padt=2", "padx=2"]
    assert call_shotwem(["adj=y", "d=d.seis", "m=m.seis"] + common) == 0
    m, hm, fh = read_seis("m.seis", nz, nx)
    assert (fh.n1, fh.n2) == (nz, nx) and abs(fh.d1 - dz) < 1e-6
    assert np.allclose(hm["mx"], mx) and np.array_equal(hm["imx"], np.arange(nx))  # shotwem.c:113-117
    p = O.WemParams(nt=nt, dt=dt, nmx=nx, dmx=dx, nz=nz, dz=dz, fmin=0.0, fmax=40.0, nref=4, padt=2, padx=2,
                    pspi=pspi, pspi_mode=0)
    m_ref = O.wem(p, vel, wav, 200.0, d=d, adj=True)
    assert rel_l2(m, m_ref) <= 1e-4
    # modelling of the image just written
    assert call_shotwem(["adj=n", "d=dm.seis", "m=m.seis"] + common) == 0
    dm, hd, fh = read_seis("dm.seis", nt, nx)
    assert (fh.n1, fh.n2) == (nt, nx) and abs(fh.d1 - dt) < 1e-7
    assert np.allclose(hd["gx"], mx) and np.allclose(hd["hx"], mx - 200.0)  # shotwem.c:86-98
    d_ref = O.wem(p, vel, wav, 200.0, m=m, adj=False)
    assert rel_l2(dm, d_ref) <= 1e-4
