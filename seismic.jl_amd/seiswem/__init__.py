"""seiswem -- MI355X shot-profile wave-equation migration (host-side mirror of Seismic.jl's
ShotProfileWEM path over the C ABI of libseiswem.so)."""
from ._lib import (PREC_FP32, PREC_KZ64, PSPI_COMPAT, PSPI_LINEAR, VARIANT_AUTO, VARIANT_SUBGROUP,  # noqa: F401
                   VARIANT_WAVE, VARIANT_SUBGROUP16, VARIANT_PAIR, VARIANT_PAIR8, VARIANT_PACKED, SwemError, device_count, load, LIB_PATH, SHOTWEM_PATH, ANGLES_PATH, SHOTEWEM_PATH,
                   REQ_PENDING, REQ_RUNNING, REQ_DONE, REQ_FAILED, SWEM_OK, SWEM_EINVAL, SWEM_ENOMEM, SWEM_EDEVICE,
                   SWEM_EIO, SWEM_EUNSUPPORTED, SWEM_EBUSY)
from .wem import DeviceArray, Plan, ShotProfileWEM, WemGeometry, copy_device, synchronize, wem  # noqa: F401
from .serve import Server  # noqa: F401
from .files import ShotProfileWEMFiles  # noqa: F401
from . import seis, solvers  # noqa: F401
from .angles import ShotProfileWEMAngles, compute_angles  # noqa: F401
from .elastic import ShotProfileEWEM, ewem  # noqa: F401
from . import synth  # noqa: F401
