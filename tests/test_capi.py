"""C ABI checks that need no GPU: the library loads, exports every symbol include/seiswem.h
declares, evaluates the reference's discrete decisions bit-exactly like the oracle, reads and
writes the reference's .seis layout, and reports errors instead of exiting (unlike
shotwem.c:30-33, which calls exit(1) inside the caller's process)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import seiswem
from seiswem import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "seiswem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(swem_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = seiswem.load()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", seiswem.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_shotwem_alias_exports_main():
    out = subprocess.run(["nm", "-D", "--defined-only", seiswem.SHOTWEM_PATH], capture_output=True, text=True).stdout
    assert re.search(r"\bT main$", out, re.M)


def test_shotewem_alias_exports_main_and_refuses_missing_names():
    out = subprocess.run(["nm", "-D", "--defined-only", seiswem.SHOTEWEM_PATH], capture_output=True, text=True).stdout
    assert re.search(r"\bT main$", out, re.M)
    so = ctypes.CDLL(seiswem.SHOTEWEM_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    argv = [b"0", b"adj=y", b"ux=a", b"uy=b", b"uz=c", b"mpp=d", b"mps1=e", b"mps2=f", b"vp=g", b"vs=h"]
    assert so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv)) == _lib.SWEM_EINVAL  # no wav=


def test_compute_angles_alias_exports_main():
    out = subprocess.run(["nm", "-D", "--defined-only", seiswem.ANGLES_PATH], capture_output=True, text=True).stdout
    assert re.search(r"\bT main$", out, re.M)


def test_compute_angles_main_missing_argument_returns_error():
    """compute_angles.c:29-34 calls exit(1) in the caller's process; the replacement returns EINVAL."""
    so = ctypes.CDLL(seiswem.ANGLES_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    for args in ([b"angx=a", b"angy=b", b"vel=v"], [b"angx=a", b"angy=b", b"vel=v", b"wav=w", b"dip_flag=y"]):
        argv = [b"0"] + args
        assert so.main(len(argv), (ctypes.c_char_p * len(argv))(*argv)) == _lib.SWEM_EINVAL
    assert "required" in seiswem.load().swem_last_error().decode()


def geoms(rng, n=200):
    for _ in range(n):
        yield seiswem.WemGeometry(
            nt=int(rng.integers(10, 3000)), dt=float(rng.choice([0.001, 0.002, 0.004, 0.0025, 0.003])),
            nmx=int(rng.integers(1, 600)), dmx=float(rng.choice([5.0, 10.0, 12.5, 7.3])),
            nz=int(rng.integers(1, 50)), dz=float(rng.choice([5.0, 10.0])),
            fmin=float(rng.uniform(0, 30)), fmax=float(rng.uniform(10, 200)),
            padt=int(rng.integers(1, 4)), padx=int(rng.integers(1, 3)), nref=int(rng.integers(2, 8)))


def to_oracle(g):
    return O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, nmy=g.nmy, dmy=g.dmy, omx=g.omx,
                       fmin=g.fmin, fmax=g.fmax, padt=g.padt, padx=g.padx, nref=g.nref, pspi=g.pspi)


def test_sampling_matches_oracle_bit_exact():
    rng = np.random.default_rng(0)
    checked = 0
    for g in geoms(rng):
        p = to_oracle(g)
        s = O.sampling(p) if (2 * (g.padt * g.nt // 2) >= g.nt or g.padt > 1) else None
        try:
            got = g.sampling()
        except seiswem.SwemError:
            with pytest.raises(ValueError):
                O.sampling(p)
            continue
        s = O.sampling(p)
        assert (got["ntfft"], got["nw"], got["nkx"], got["nky"], got["ifmin"], got["ifmax"], got["lmx"]) == \
            (s.ntfft, s.nw, s.nkx, s.nky, s.ifmin, s.ifmax, s.lmx)
        checked += 1
    assert checked > 150


def test_source_index_matches_oracle():
    rng = np.random.default_rng(1)
    lib = seiswem.load()
    for _ in range(300):
        g = seiswem.WemGeometry(nt=10, dt=0.004, nmx=int(rng.integers(2, 500)), dmx=float(rng.uniform(2, 25)),
                                nz=5, dz=10, omx=float(rng.uniform(-50, 50)))
        sx = float(rng.uniform(-100, g.nmx * g.dmx + 100))
        ix, iy = ctypes.c_int(), ctypes.c_int()
        c = g.c()
        rc = lib.swem_source_index(ctypes.byref(c), sx, 0.0, ctypes.byref(ix), ctypes.byref(iy))
        orc, oix, _ = O.source_index(to_oracle(g), sx)
        assert (rc == 0) == (orc == 0)
        if rc == 0:
            assert ix.value == oix


@pytest.mark.parametrize("seed", range(6))
def test_pspi_tables_match_oracle(seed):
    rng = np.random.default_rng(seed)
    nmx, nmy, nz, nref = int(rng.integers(2, 80)), int(rng.choice([1, 1, 3])), int(rng.integers(1, 30)), int(rng.integers(2, 9))
    g = seiswem.WemGeometry(nt=10, dt=0.004, nmx=nmx, nmy=nmy, dmx=10, dmy=10, nz=nz, dz=10, nref=nref)
    vel = rng.uniform(1500, 4000, size=(nmx * nmy, nz)).astype(np.float32)
    vel[:, nz // 2] = 2000.0  # one laterally invariant layer
    vref = np.zeros((nref, nz), np.float32)
    i1 = np.zeros((nmx * nmy, nz), np.int32)
    i2 = np.zeros_like(i1)
    cut = np.zeros((nmx * nmy, nz), np.uint8)
    c = g.c()
    P = ctypes.POINTER
    seiswem._lib.check(seiswem.load().swem_pspi_tables(
        ctypes.byref(c), vel.ctypes.data_as(P(ctypes.c_float)), vref.ctypes.data_as(P(ctypes.c_float)),
        i1.ctypes.data_as(P(ctypes.c_int)), i2.ctypes.data_as(P(ctypes.c_int)), cut.ctypes.data_as(P(ctypes.c_ubyte))))
    ovref, oi1, oi2, ocut = O.pspi_tables(to_oracle(g), vel)
    assert np.array_equal(vref, ovref) and np.array_equal(i1, oi1) and np.array_equal(i2, oi2)
    assert np.array_equal(cut, ocut)


@pytest.mark.parametrize("n,l", [(256, 30), (101, 30), (1024, 30), (64, 0)])
def test_taper_matches_oracle_bit_exact(n, l):
    t = np.zeros(n, np.float32)
    seiswem._lib.check(seiswem.load().swem_taper(n, l, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    assert np.array_equal(t, O.taper(n, l, f64=False))


def header_array(n, **fields):
    names = ["tracenum", "o1", "n1", "d1", "sx", "sy", "gx", "gy", "mx", "my", "hx", "hy", "h", "az", "ang", "isx",
             "isy", "igx", "igy", "imx", "imy", "ihx", "ihy", "ih", "iaz", "iang", "selev", "gelev", "sstat", "gstat",
             "trid"]
    ints = {"tracenum", "n1", "isx", "isy", "igx", "igy", "imx", "imy", "ihx", "ihy", "ih", "iaz", "iang", "trid"}
    dt = np.dtype([(k, "<i4" if k in ints else "<f4") for k in names])
    h = np.zeros(n, dt)
    for k, v in fields.items():
        h[k] = v
    return h


def write_seis(lib, name, data, h, n1, n2, o1=0.0, d1=1.0, o2=0.0, d2=1.0):
    fh = _lib.SeisExtent()
    lib.swem_seis_extent_init(ctypes.byref(fh))
    fh.n1, fh.n2, fh.o1, fh.d1, fh.o2, fh.d2 = n1, n2, o1, d1, o2, d2
    data = np.ascontiguousarray(data, np.float32)
    _lib.check(lib.swem_seis_write(name.encode(), data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   h.ctypes.data, ctypes.byref(fh)))


def test_seis_roundtrip_and_text_layout(tmp_path, monkeypatch):
    lib = seiswem.load()
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    d = np.arange(12, dtype=np.float32).reshape(3, 4)
    h = header_array(3, tracenum=[1, 2, 3], n1=4, imx=[0, 1, 2], mx=[0.0, 10.0, 20.0])
    write_seis(lib, "tmp_a.seis", d, h, 4, 3, o1=0.0, d1=0.004, d2=10.0)
    txt = open("tmp_a.seis").read().splitlines()
    # WriteFileHeader (seismic.c:427-457): n1..n5, o1..o5 and d1..d5 with %f, labels, unit5 twice
    assert txt[0] == "\tn1=4" and txt[5] == "\to1=0.000000" and txt[10] == "\td1=0.004000"
    assert sum(1 for t in txt if t.startswith("\tunit5=")) == 2
    assert txt[-2] == '\tin="tmp_a.seis@data@"' and txt[-1] == '\theaders="tmp_a.seis@headers@"'
    assert os.path.getsize("tmp_a.seis@headers@") == 3 * 124
    assert np.array_equal(np.fromfile("tmp_a.seis@data@", np.float32).reshape(3, 4), d)
    fh = _lib.SeisExtent()
    out = np.zeros((3, 4), np.float32)
    hb = np.zeros(3, h.dtype)
    _lib.check(lib.swem_seis_read(b"tmp_a.seis", out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), hb.ctypes.data,
                                  ctypes.byref(fh)))
    assert np.array_equal(out, d) and np.array_equal(hb, h)
    assert (fh.n1, fh.n2, fh.n3) == (4, 3, 1) and abs(fh.d1 - 0.004) < 1e-7


def test_seis_datapath_prefix(tmp_path, monkeypatch):
    lib = seiswem.load()
    monkeypatch.chdir(tmp_path)
    (tmp_path / "dp").mkdir()
    monkeypatch.setenv("DATAPATH", str(tmp_path / "dp") + "/")
    write_seis(lib, "x.seis", np.ones((1, 2)), header_array(1), 2, 1)
    assert (tmp_path / "dp" / "x.seis@data@").exists() and (tmp_path / "x.seis").exists()


@pytest.mark.parametrize("extra", [0, 1, 2])
def test_seis_write_rejects_header_name_past_its_field(tmp_path, monkeypatch, extra):
    """The trace-header file name (base + "@headers@") is 3 bytes longer than the data file name
    (base + "@data@"): a base of 91..93 bytes fits the 100-byte data field but not the header
    field, and is refused instead of overrunning it."""
    lib = seiswem.load()
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    name = "n" * (91 + extra)
    assert len(name + "@data@") < 100 <= len(name + "@headers@")
    fh = _lib.SeisExtent()
    lib.swem_seis_extent_init(ctypes.byref(fh))
    fh.n1, fh.n2 = 2, 1
    data = np.ones(2, np.float32)
    rc = lib.swem_seis_write(name.encode(), data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None,
                             ctypes.byref(fh))
    assert rc == _lib.SWEM_EINVAL and "too long" in lib.swem_last_error().decode()
    assert not os.path.exists(name)


@pytest.mark.parametrize("key,val", [("n2", "0"), ("n1", "-3"), ("n3", "0")])
def test_seis_header_with_empty_or_negative_extent_is_refused(tmp_path, monkeypatch, key, val):
    """A velocity header with n2 = 0 (no traces) or a negative extent is refused with EINVAL by the
    header reader, before any buffer is sized from it or any trace header is indexed."""
    lib = seiswem.load()
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    write_seis(lib, "vel.seis", np.full((2, 4), 2000.0), header_array(2, imx=[0, 1], mx=[0.0, 10.0]), 4, 2)
    write_seis(lib, "wav.seis", np.zeros((1, 8)), header_array(1), 8, 1, d1=0.004)
    txt = open("vel.seis").read().replace(f"\t{key}=", f"\t{key}={val}\n\told_{key}=")
    open("vel.seis", "w").write(txt)
    fh = _lib.SeisExtent()
    assert lib.swem_seis_read_header(b"vel.seis", ctypes.byref(fh)) == _lib.SWEM_EINVAL
    assert call_main(["adj=y", "d=d.seis", "m=m.seis", "vel=vel.seis", "wav=wav.seis"]) == _lib.SWEM_EINVAL
    assert "extents must be >= 1" in lib.swem_last_error().decode()


def call_main(args):
    lib = seiswem.load()
    argv = [b"0"] + [a.encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return lib.swem_shot_main(len(argv), arr)


def test_shot_main_missing_argument_returns_error():
    """shotwem.c:30-33 calls exit(1) in the caller's process; the replacement returns EINVAL."""
    assert call_main(["adj=y", "m=img", "vel=vel", "wav=wav"]) == _lib.SWEM_EINVAL
    assert "required" in seiswem.load().swem_last_error().decode()


def test_shot_main_missing_file_is_io_error(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    assert call_main(["d=nod", "m=nom", "vel=novel", "wav=nowav"]) == _lib.SWEM_EIO


def test_shotwem_so_main_callable_in_process(tmp_path, monkeypatch):
    """The zero-change Julia binding: ccall((:main, "shotwem"), Int32, (Int32, Ptr{Ptr{UInt8}}), ...)."""
    monkeypatch.chdir(tmp_path)
    so = ctypes.CDLL(seiswem.SHOTWEM_PATH)
    so.main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    so.main.restype = ctypes.c_int
    argv = [b"0", b"adj=y"]
    assert so.main(len(argv), (ctypes.c_char_p * 2)(*argv)) == _lib.SWEM_EINVAL


def test_seis_module_counts_traces_from_the_files(tmp_path, monkeypatch):
    """seiswem.seis writes every trace it is given and reads the trace count from the headers file
    (SeisWrite.jl:31-40, Header.jl:223-230), also when n2*n3*n4*n5 differs from it."""
    from seiswem import seis
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("DATAPATH", raising=False)
    assert seis.HEADER_DTYPE.itemsize == 124
    d = np.arange(30, dtype=np.float32).reshape(6, 5)
    h = seis.headers(6, tracenum=np.arange(1, 7), sx=[0, 0, 0, 10, 10, 10], gx=np.arange(6) * 2.5)
    fh = seis.extent((5, 3, 1, 2, 2), (0.0, 0.0, 0.0, 0.0, 0.0), (0.004, 2.5, 2.5, 10.0, 1.0),
                     ("Time", "gx", "gy", "sx", "sy"), ("s", "m", "m", "m", "m"))
    seis.write("c.seis", d, h, fh)
    d2, h2, fh2 = seis.read("c.seis")
    assert np.array_equal(d2, d) and np.array_equal(h2, h)
    assert (fh2.n4, fh2.n5, fh2.label5, fh2.unit1) == (2, 2, b"sy", b"s")
    # a consistent extent reads back identically through the C reader
    fh3 = seis.extent((5, 6))
    seis.write("e.seis", d, h, fh3)
    d3, h3, _ = seis.read("e.seis")
    data = np.zeros((6, 5), np.float32)
    hc = header_array(6)
    fh4 = _lib.SeisExtent()
    _lib.check(seiswem.load().swem_seis_read(b"e.seis", data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             hc.ctypes.data, ctypes.byref(fh4)))
    assert np.array_equal(data, d3) and hc.tobytes() == h3.tobytes()
