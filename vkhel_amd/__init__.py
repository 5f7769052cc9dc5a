"""vkhel_amd: MI355X-native implementation of lolzballs/vkhel's hot path.

The product is the C-ABI shared library ``vkhel_amd/lib/libvkhel.so``
(include/vkhel/vkhel.h, the reference API unchanged, plus
include/vkhel/vkhel_ext.h).  This package is a thin ctypes mirror of that
interface for Python callers and tests; it never falls back to CPU code --
if the library is missing or no GPU is present, calls raise.
"""
from .lib import LIB_PATH, VkhelError, load, symbols  # noqa: F401
from .vkhel import (Batch, Context, HostBuffer, NttTables, Server, Vector,  # noqa: F401
                    JOB_ADD, JOB_ADD_REF, JOB_CENTER, JOB_GTADD, JOB_GTSUB, JOB_INTT_CENTER, JOB_RESCALE,
                    JOB_RESCALE_NTT,
                    JOB_MUL, JOB_NTT_FWD, JOB_NTT_INV, JOB_POLYMUL,
                    NTT_AUTO, NTT_FUSED, NTT_RADIX2)
