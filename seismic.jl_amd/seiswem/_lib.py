"""ctypes binding of libseiswem.so (include/seiswem.h).

The library is the product path: every function here calls the HIP library, and loading fails
with an error when it has not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libseiswem.so")
SHOTWEM_PATH = os.path.join(LIB_DIR, "shotwem.so")
ANGLES_PATH = os.path.join(LIB_DIR, "compute_angles.so")
SHOTEWEM_PATH = os.path.join(LIB_DIR, "shotewem.so")

SWEM_OK, SWEM_EINVAL, SWEM_ENOMEM, SWEM_EDEVICE, SWEM_EIO, SWEM_EUNSUPPORTED, SWEM_EBUSY = range(7)
PSPI_COMPAT, PSPI_LINEAR = 0, 1
PREC_KZ64, PREC_FP32 = 0, 1
VARIANT_AUTO, VARIANT_SUBGROUP, VARIANT_WAVE, VARIANT_SUBGROUP16, VARIANT_PAIR, VARIANT_PAIR8, VARIANT_PACKED = range(7)

# Every symbol declared in include/seiswem.h.
EXPORTS = (
    "swem_version", "swem_last_error", "swem_options_init", "swem_device_count",
    "swem_shot_main", "swem_angles_main", "swem_eshot_main", "swem_wem", "swem_wem_ex", "swem_shots", "swem_compute_angles", "swem_compute_angles_ex", "swem_ewem", "swem_eshots",
    "swem_plan_create", "swem_plan_destroy", "swem_plan_migrate", "swem_plan_migrate_each", "swem_plan_migrate_noalloc",
    "swem_plan_model", "swem_plan_timing", "swem_plan_autotune", "swem_plan_autotune_ex", "swem_plan_set_graphs", "swem_plan_bytes",
    "swem_malloc", "swem_free", "swem_memcpy_htod", "swem_memcpy_dtoh", "swem_memcpy_dtod", "swem_memset", "swem_synchronize",
    "swem_sampling", "swem_source_index", "swem_pspi_tables", "swem_taper",
    "swem_seis_extent_init", "swem_seis_read_header", "swem_seis_read", "swem_seis_write",
    "swem_server_config_init", "swem_server_create", "swem_server_destroy", "swem_batch_submit",
    "swem_batch_poll", "swem_batch_wait", "swem_server_get_stats", "swem_vec_cgstep", "swem_vec_dot", "swem_vec_mul",
)
REQ_PENDING, REQ_RUNNING, REQ_DONE, REQ_FAILED = range(4)


class SwemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"seiswem error {code}: {msg}")
        self.code = code


class Geometry(ctypes.Structure):
    """swem_geometry: the scalar arguments of wem() (wem.c:3-11)."""

    _fields_ = [
        ("nt", ctypes.c_int), ("ot", ctypes.c_float), ("dt", ctypes.c_float),
        ("nmx", ctypes.c_int), ("omx", ctypes.c_float), ("dmx", ctypes.c_float),
        ("nmy", ctypes.c_int), ("omy", ctypes.c_float), ("dmy", ctypes.c_float),
        ("nz", ctypes.c_int), ("oz", ctypes.c_float), ("dz", ctypes.c_float),
        ("gz", ctypes.c_float), ("sz", ctypes.c_float),
        ("nref", ctypes.c_int), ("fmin", ctypes.c_float), ("fmax", ctypes.c_float),
        ("padt", ctypes.c_int), ("padx", ctypes.c_int),
        ("pspi", ctypes.c_int),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("pspi_mode", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int),
        ("freqs_per_block", ctypes.c_int), ("verbose", ctypes.c_int), ("scratch_bytes", ctypes.c_size_t),
        ("variant", ctypes.c_int), ("table_bytes", ctypes.c_size_t),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double), ("spectra_ms", ctypes.c_double), ("march_ms", ctypes.c_double),
        ("reduce_ms", ctypes.c_double), ("march_launches", ctypes.c_int), ("march_flops", ctypes.c_double),
    ]


TUNE_MAX = 48


class TuneResult(ctypes.Structure):
    """swem_tune_result: measured launch shapes of a plan."""

    _fields_ = [
        ("variant", ctypes.c_int), ("freqs_per_block", ctypes.c_int), ("chunk_bytes", ctypes.c_size_t),
        ("best_ms", ctypes.c_float), ("ncand", ctypes.c_int),
        ("cand_variant", ctypes.c_int * TUNE_MAX), ("cand_fpb", ctypes.c_int * TUNE_MAX),
        ("cand_chunk_bytes", ctypes.c_size_t * TUNE_MAX), ("cand_ms", ctypes.c_float * TUNE_MAX),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int), ("device_bytes", ctypes.c_size_t), ("pinned_bytes", ctypes.c_size_t),
        ("streams", ctypes.c_int), ("max_inflight", ctypes.c_int), ("max_pending", ctypes.c_int),
        ("skip_window_ms", ctypes.c_double), ("max_fuse", ctypes.c_int),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("g", Geometry), ("pspi_mode", ctypes.c_int), ("precision", ctypes.c_int),
        ("vel", ctypes.POINTER(ctypes.c_float)), ("wav", ctypes.POINTER(ctypes.c_float)),
        ("d", ctypes.POINTER(ctypes.c_float)), ("sx", ctypes.POINTER(ctypes.c_float)),
        ("sy", ctypes.POINTER(ctypes.c_float)), ("nshot", ctypes.c_int),
    ]


class RequestStatus(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int), ("error", ctypes.c_int), ("submit_ms", ctypes.c_double),
        ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double), ("message", ctypes.c_char * 256),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
        ("pending", ctypes.c_int), ("inflight", ctypes.c_int), ("device_bytes", ctypes.c_size_t),
        ("device_bytes_used", ctypes.c_size_t), ("device_bytes_peak", ctypes.c_size_t),
        ("basis_bytes", ctypes.c_size_t), ("busy_ms", ctypes.c_double), ("fused_launches", ctypes.c_longlong),
        ("fused_requests", ctypes.c_longlong),
    ]


class SeisExtent(ctypes.Structure):
    """swem_seis_extent == struct SeisFileHeader (seismic.h:54-86)."""

    _fields_ = (
        [(f"n{i}", ctypes.c_int) for i in range(1, 6)]
        + [(f"o{i}", ctypes.c_float) for i in range(1, 6)]
        + [(f"d{i}", ctypes.c_float) for i in range(1, 6)]
        + [(f"label{i}", ctypes.c_char * 100) for i in range(1, 6)]
        + [(f"unit{i}", ctypes.c_char * 100) for i in range(1, 6)]
        + [("title", ctypes.c_char * 100), ("data_format", ctypes.c_char * 100), ("esize", ctypes.c_int),
           ("dname", ctypes.c_char * 100), ("hname", ctypes.c_char * 100)]
    )


_lib = None


def load():
    """Load libseiswem.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C seismic.jl_amd/csrc` "
                          "or __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, f32p, i32p = ctypes.c_void_p, P(ctypes.c_float), P(ctypes.c_int)
    gp, op_ = P(Geometry), P(Options)
    sig = {
        "swem_version": ([], ctypes.c_int),
        "swem_last_error": ([], ctypes.c_char_p),
        "swem_options_init": ([op_], None),
        "swem_device_count": ([i32p], ctypes.c_int),
        "swem_shot_main": ([ctypes.c_int, P(ctypes.c_char_p)], ctypes.c_int),
        "swem_angles_main": ([ctypes.c_int, P(ctypes.c_char_p)], ctypes.c_int),
        "swem_eshot_main": ([ctypes.c_int, P(ctypes.c_char_p)], ctypes.c_int),
        "swem_wem": ([f32p, f32p, f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                      ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "swem_wem_ex": ([f32p, f32p, f32p, gp, ctypes.c_float, ctypes.c_float, f32p, ctypes.c_int, op_], ctypes.c_int),
        "swem_ewem": ([f32p] * 7 + [ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, f32p, f32p, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "swem_compute_angles_ex": ([f32p, f32p, f32p, gp, ctypes.c_float, ctypes.c_float, f32p, op_], ctypes.c_int),
        "swem_eshots": ([gp, f32p, f32p, f32p, f32p, f32p, ctypes.c_int] + [f32p] * 6 + [ctypes.c_int, op_],
                        ctypes.c_int),
        "swem_compute_angles": ([f32p, f32p, f32p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, f32p, ctypes.c_float, ctypes.c_float, ctypes.c_int], ctypes.c_int),
        "swem_shots": ([gp, f32p, f32p, f32p, f32p, ctypes.c_int, f32p, f32p, ctypes.c_int, op_], ctypes.c_int),
        "swem_plan_create": ([gp, f32p, f32p, op_, P(vp)], ctypes.c_int),
        "swem_plan_destroy": ([vp], ctypes.c_int),
        "swem_plan_migrate": ([vp, vp, f32p, f32p, ctypes.c_int, vp, ctypes.c_int, vp], ctypes.c_int),
        "swem_plan_migrate_noalloc": ([vp, vp, f32p, f32p, ctypes.c_int, vp, ctypes.c_int, vp], ctypes.c_int),
        "swem_plan_migrate_each": ([vp, vp, f32p, f32p, ctypes.c_int, vp, vp], ctypes.c_int),
        "swem_plan_model": ([vp, vp, f32p, f32p, ctypes.c_int, vp, vp], ctypes.c_int),
        "swem_plan_timing": ([vp, P(Timing)], ctypes.c_int),
        "swem_plan_autotune": ([vp, ctypes.c_int, i32p, f32p], ctypes.c_int),
        "swem_plan_autotune_ex": ([vp, ctypes.c_int, P(TuneResult)], ctypes.c_int),
        "swem_plan_set_graphs": ([vp, ctypes.c_int], ctypes.c_int),
        "swem_plan_bytes": ([vp, P(ctypes.c_size_t), P(ctypes.c_size_t)], ctypes.c_int),
        "swem_malloc": ([P(vp), ctypes.c_size_t], ctypes.c_int),
        "swem_free": ([vp], ctypes.c_int),
        "swem_memcpy_htod": ([vp, vp, ctypes.c_size_t], ctypes.c_int),
        "swem_memcpy_dtoh": ([vp, vp, ctypes.c_size_t], ctypes.c_int),
        "swem_memcpy_dtod": ([vp, vp, ctypes.c_size_t], ctypes.c_int),
        "swem_memset": ([vp, ctypes.c_int, ctypes.c_size_t], ctypes.c_int),
        "swem_synchronize": ([], ctypes.c_int),
        "swem_sampling": ([gp, i32p], ctypes.c_int),
        "swem_source_index": ([gp, ctypes.c_float, ctypes.c_float, i32p, i32p], ctypes.c_int),
        "swem_pspi_tables": ([gp, f32p, f32p, i32p, i32p, P(ctypes.c_ubyte)], ctypes.c_int),
        "swem_taper": ([ctypes.c_int, ctypes.c_int, f32p], ctypes.c_int),
        "swem_seis_extent_init": ([P(SeisExtent)], None),
        "swem_seis_read_header": ([ctypes.c_char_p, P(SeisExtent)], ctypes.c_int),
        "swem_seis_read": ([ctypes.c_char_p, f32p, vp, P(SeisExtent)], ctypes.c_int),
        "swem_seis_write": ([ctypes.c_char_p, f32p, vp, P(SeisExtent)], ctypes.c_int),
        "swem_server_config_init": ([P(ServerConfig)], None),
        "swem_server_create": ([P(ServerConfig), P(vp)], ctypes.c_int),
        "swem_server_destroy": ([vp], ctypes.c_int),
        "swem_batch_submit": ([vp, P(Request), P(ctypes.c_longlong)], ctypes.c_int),
        "swem_batch_poll": ([vp, ctypes.c_longlong, P(RequestStatus)], ctypes.c_int),
        "swem_batch_wait": ([vp, ctypes.c_longlong, f32p, P(RequestStatus)], ctypes.c_int),
        "swem_server_get_stats": ([vp, P(ServerStats)], ctypes.c_int),
        "swem_vec_cgstep": ([ctypes.c_float, vp, ctypes.c_float, vp, ctypes.c_size_t, vp], ctypes.c_int),
        "swem_vec_dot": ([vp, vp, ctypes.c_size_t, P(ctypes.c_double), vp], ctypes.c_int),
        "swem_vec_mul": ([vp, vp, vp, ctypes.c_size_t, vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int):
    if rc != SWEM_OK:
        raise SwemError(rc, load().swem_last_error().decode(errors="replace"))


def options(pspi_mode: int = PSPI_COMPAT, precision: int = PREC_KZ64, device: int = 0, freqs_per_block: int = 0,
            verbose: bool = False, scratch_bytes: int = 0, variant: int = VARIANT_AUTO,
            table_bytes: int = 0) -> Options:
    o = Options()
    load().swem_options_init(ctypes.byref(o))
    o.pspi_mode, o.precision, o.device = pspi_mode, precision, device
    o.freqs_per_block, o.verbose, o.scratch_bytes = freqs_per_block, int(verbose), scratch_bytes
    o.variant, o.table_bytes = variant, table_bytes
    return o


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = load().swem_device_count(ctypes.byref(n))
    return n.value if rc == SWEM_OK else 0
