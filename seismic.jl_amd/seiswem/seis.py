"""The .seis file format through the C ABI (swem_seis_*): a text header, a data file of native
floats (n1 fastest) and 31 four-byte trace-header words per trace (src/API/c/seismic.c:270-461,
src/ReadWrite/Header.jl:1-33)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, load

HEADER_FIELDS = ("tracenum", "o1", "n1", "d1", "sx", "sy", "gx", "gy", "mx", "my", "hx", "hy", "h", "az", "ang", "isx",
                 "isy", "igx", "igy", "imx", "imy", "ihx", "ihy", "ih", "iaz", "iang", "selev", "gelev", "sstat",
                 "gstat", "trid")
_INTS = {"tracenum", "n1", "isx", "isy", "igx", "igy", "imx", "imy", "ihx", "ihy", "ih", "iaz", "iang", "trid"}
HEADER_DTYPE = np.dtype([(k, "<i4" if k in _INTS else "<f4") for k in HEADER_FIELDS])


def headers(n: int, **fields) -> np.ndarray:
    """n zero trace headers (InitSeisHeader) with the given fields set."""
    h = np.zeros(n, HEADER_DTYPE)
    for k, v in fields.items():
        h[k] = v
    return h


def read(name: str):
    """Returns (data [ntrace, n1] float32, headers [ntrace], extent).  The trace count is the
    headers file's size over 124 bytes (GetNumTraces, Header.jl:223-230), so files whose extent
    does not multiply out to their trace count (ShotProfileWEM.jl:290-296 writes n4 * n5 =
    nsx * nsy for nsx shots) read whole."""
    lib = load()
    fh = _lib.SeisExtent()
    check(lib.swem_seis_read_header(name.encode(), ctypes.byref(fh)))
    hname, dname = fh.hname.decode(), fh.dname.decode()
    if hname != "NULL" and os.path.exists(hname):
        h = np.fromfile(hname, HEADER_DTYPE)
    else:
        h = np.zeros(0, HEADER_DTYPE)
    data = np.fromfile(dname, np.float32)
    n1 = max(fh.n1, 1)
    ntr = h.size if h.size else data.size // n1
    if data.size < ntr * n1:
        raise _lib.SwemError(_lib.SWEM_EIO, f"SeisRead: {dname} holds {data.size} of {ntr * n1} samples")
    return data[: ntr * n1].reshape(ntr, n1), h, fh
def extent(n=(), o=(), d=(), labels=(), units=(), title=""):
    """A file header (InitFileHeader, seismic.c:346) with the first axes given."""
    fh = _lib.SeisExtent()
    load().swem_seis_extent_init(ctypes.byref(fh))
    for i, v in enumerate(n):
        setattr(fh, f"n{i + 1}", int(v))
    for i, v in enumerate(o):
        setattr(fh, f"o{i + 1}", float(v))
    for i, v in enumerate(d):
        setattr(fh, f"d{i + 1}", float(v))
    for i, v in enumerate(labels):
        setattr(fh, f"label{i + 1}", v.encode())
    for i, v in enumerate(units):
        setattr(fh, f"unit{i + 1}", v.encode())
    fh.title = title.encode()
    return fh


def write(name: str, data: np.ndarray, h: np.ndarray, fh) -> None:
    """SeisWrite (SeisWrite.jl:19-41): the text header from fh, then every trace of data and h
    (the count comes from the arrays, not from the extent)."""
    data = np.ascontiguousarray(data, np.float32)
    h = np.ascontiguousarray(h, HEADER_DTYPE)
    if data.ndim != 2 or data.shape[0] != h.size:
        raise ValueError("data [ntrace, n1] and one header per trace")
    check(load().swem_seis_write(name.encode(), None, None, ctypes.byref(fh)))
    h.tofile(fh.hname.decode())
    data.tofile(fh.dname.decode())
