"""orkan_amd -- MI355X-native state-update path of the Orkan quantum-state simulator.

The product is the C-ABI shared library orkan_amd/lib/liborkan.so (built by
orkan_amd.build): it exports the Orkan C API (include/orkan.h) with states
resident in HBM and every operation executed by hand-written gfx950 kernels.
This module locates and binds that library with ctypes; it never substitutes
a host implementation -- `load()` raises when the library is missing.
"""
import ctypes as C
import os

from . import _ffi
from ._ffi import MIXED_PACKED, MIXED_TILED, PURE, Cplx, KrausT, StateT, SuperopT  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
ORKAN_FP32 = 0x100

LIBS = {5: "liborkan.so", 3: "liborkan_t3.so"}

# additions of this library (include/orkan_amd.h)
EXTRA = {
    "orkan_state_upload": (C.c_int, [_ffi.PS, C.c_void_p]),
    "orkan_state_download": (C.c_int, [_ffi.PS, C.c_void_p]),
    "orkan_state_convert": (_ffi.StateT, [_ffi.PS, C.c_int]),
    "orkan_state_relayout": (C.c_int, [_ffi.PS, _ffi.PS, C.c_int]),
    "orkan_layer_1q": (C.c_int, [_ffi.PS, C.c_void_p]),
    "orkan_state_save": (C.c_int, [_ffi.PS, C.c_char_p]),
    "orkan_state_load": (C.c_int, [_ffi.PS, C.c_char_p]),
    "orkan_dev_alloc": (C.c_void_p, [C.c_uint64]),
    "orkan_dev_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "orkan_dev_free": (None, [C.c_void_p]),
    "orkan_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "orkan_copy_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "orkan_stream": (C.c_void_p, []),
    "orkan_sync": (C.c_int, []),
    "orkan_event_record": (C.c_int, [C.c_int]),
    "orkan_event_ms": (C.c_double, [C.c_int, C.c_int]),
    "orkan_set_launch_shape": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "orkan_exp_diag_noalloc": (C.c_int, [_ffi.PS, C.c_void_p, C.c_double]),
}


def lib_path(log_tile=5):
    return os.path.join(LIB_DIR, LIBS[log_tile])


_cache = {}


def load(log_tile=5):
    """Bind the HIP library; raises FileNotFoundError when it has not been built."""
    if log_tile not in _cache:
        path = lib_path(log_tile)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: build it with `python -m orkan_amd.build`")
        extra = dict(EXTRA)
        from . import batch as _batch
        from . import surrogate as _sur
        extra.update(_batch.EXTRA)
        extra.update(_sur.EXTRA)
        _cache[log_tile] = _ffi.bind(path, extra)
    return _cache[log_tile]
