"""ctypes binding of libtlbm.so (the C ABI declared in include/tlbm.h).

The library is loaded from this directory.  There is no fallback: if it is missing or does not
load, every entry point raises, so nothing silently computes the step elsewhere.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtlbm.so")

TLBM_OK = 0
TLBM_NSUMS = 14
TLBM_NOALLOC_MAX_BLOCKS = 131072

BGK, TRT, LES = 0, 1, 2
CYLINDER, SPHERE, BOX, FIELD, HEATSINK = 0, 1, 2, 3, 4
QUIET, REPORT = 0, 1
PERIODIC_Y, SLIP_Y, LID_Y = 0, 1, 2
F32, F16, BF16 = 0, 1, 2


class Spec(ctypes.Structure):
    _fields_ = [("collision", ctypes.c_int32), ("obstacle", ctypes.c_int32),
                ("diag", ctypes.c_int32), ("side", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("heat", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("block_x", ctypes.c_int32), ("qstride", ctypes.c_int64),
                ("tile_rows", ctypes.c_int32), ("xcd_group", ctypes.c_int32),
                ("omega", ctypes.c_float), ("tau0", ctypes.c_float), ("cs", ctypes.c_float),
                ("magic", ctypes.c_float), ("obs", ctypes.c_float * 8), ("u0", ctypes.c_float),
                ("sponge_start", ctypes.c_float), ("sponge_strength", ctypes.c_float),
                ("lid", ctypes.c_float * 3), ("omega_rot", ctypes.c_void_p),
                ("phi", ctypes.c_void_p)]


FIELD_NAMES = ("ux", "uy", "uz", "p", "vmag", "r_rho", "r_ux", "r_uy", "r_uz", "r_ke",
               "fx", "fy", "fz")


class Fields(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in FIELD_NAMES]


class Heat(ctypes.Structure):
    _fields_ = [("t_solid", ctypes.c_void_p), ("q_solid", ctypes.c_void_p),
                ("temp", ctypes.c_void_p), ("alpha", ctypes.c_float), ("nu", ctypes.c_float),
                ("pr_t", ctypes.c_float)]


# name -> (restype, argtypes); the exported surface of include/tlbm.h
_c = ctypes
_vp = _c.c_void_p
SIGNATURES = {
    "tlbm_last_error": (_c.c_char_p, []),
    "tlbm_version": (_c.c_int, []),
    "tlbm_device_count": (_c.c_int, [_c.POINTER(_c.c_int)]),
    "tlbm_set_device": (_c.c_int, [_c.c_int]),
    "tlbm_alloc_managed": (_c.c_int, [_c.c_size_t, _c.POINTER(_vp)]),
    "tlbm_alloc_device": (_c.c_int, [_c.c_size_t, _c.POINTER(_vp)]),
    "tlbm_alloc_host": (_c.c_int, [_c.c_size_t, _c.POINTER(_vp)]),
    "tlbm_free": (_c.c_int, [_vp]),
    "tlbm_free_host": (_c.c_int, [_vp]),
    "tlbm_memcpy": (_c.c_int, [_vp, _vp, _c.c_size_t]),
    "tlbm_memcpy_async": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp]),
    "tlbm_memset": (_c.c_int, [_vp, _c.c_int, _c.c_size_t, _vp]),
    "tlbm_stream_create": (_c.c_int, [_c.POINTER(_vp)]),
    "tlbm_stream_destroy": (_c.c_int, [_vp]),
    "tlbm_stream_sync": (_c.c_int, [_vp]),
    "tlbm_device_sync": (_c.c_int, []),
    "tlbm_event_create": (_c.c_int, [_c.POINTER(_vp)]),
    "tlbm_event_destroy": (_c.c_int, [_vp]),
    "tlbm_event_record": (_c.c_int, [_vp, _vp]),
    "tlbm_event_elapsed_ms": (_c.c_int, [_vp, _vp, _c.POINTER(_c.c_float)]),
    "tlbm_kernel_lookup": (_c.c_int, [_c.POINTER(Spec), _c.POINTER(_c.c_int)]),
    "tlbm_set_launch": (_c.c_int, [_c.c_int, _c.c_int]),
    "tlbm_population_elems": (_c.c_int, [_c.POINTER(Params), _c.POINTER(_c.c_uint64)]),
    "tlbm_init": (_c.c_int, [_c.c_int, _vp, _c.POINTER(Params), _c.c_float, _c.c_float,
                             _c.c_uint64, _vp]),
    "tlbm_step": (_c.c_int, [_c.c_int, _vp, _vp, _c.POINTER(Params), _c.POINTER(Fields),
                             _vp, _vp]),
    "tlbm_step_noalloc": (_c.c_int, [_c.c_int, _vp, _vp, _c.POINTER(Params), _c.POINTER(Fields),
                                     _vp, _vp]),
    "tlbm_copy": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp]),
    "tlbm_pull_copy": (_c.c_int, [_vp, _vp, _c.POINTER(Params), _c.c_int, _vp]),
    "tlbm_batch_create": (_c.c_int, [_c.c_int, _c.POINTER(Params), _c.c_int, _c.POINTER(_vp)]),
    "tlbm_batch_destroy": (_c.c_int, [_vp]),
    "tlbm_batch_layout": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_uint64),
                                     _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64),
                                     _c.POINTER(_c.c_uint64)]),
    "tlbm_batch_set_launch": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "tlbm_batch_set_order": (_c.c_int, [_vp, _c.c_int]),
    "tlbm_batch_init": (_c.c_int, [_vp, _vp, _c.c_float, _c.c_float, _c.c_uint64, _vp]),
    "tlbm_batch_step": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.POINTER(Fields), _vp, _vp]),
    "tlbm_batch_graph": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.c_int, _c.POINTER(Fields), _vp,
                                    _vp, _c.POINTER(_vp)]),
    "tlbm_graph_launch": (_c.c_int, [_vp, _vp]),
    "tlbm_graph_destroy": (_c.c_int, [_vp]),
    "tlbm_capture_begin": (_c.c_int, [_vp]),
    "tlbm_capture_end": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "tlbm_device_cus": (_c.c_int, [_c.POINTER(_c.c_int)]),
    "tlbm_step_heat": (_c.c_int, [_c.c_int, _vp, _vp, _vp, _vp, _c.POINTER(Params),
                                  _c.POINTER(Heat), _c.POINTER(Fields), _vp, _vp]),
    "tlbm_velocity_gradient": (_c.c_int, [_vp, _vp, _vp, _vp, _c.c_int, _c.c_int, _c.c_int,
                                          _c.c_float, _vp]),
    "tlbm_vortex_fields": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_int, _c.c_int, _c.c_int,
                                      _c.c_float, _c.c_int, _vp]),
    "tlbm_adjoint_step": (_c.c_int, [_c.POINTER(Params), _vp, _vp, _vp, _vp, _c.c_float,
                                     _c.c_float, _c.POINTER(_c.c_int32),
                                     _c.POINTER(_c.c_int32), _vp, _c.c_int, _vp]),
    "tlbm_device_name": (_c.c_int, [_c.c_char_p, _c.c_int]),
    "tlbm_coarsen": (_c.c_int, [_vp, _vp, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
                                _c.c_int, _c.c_int, _vp]),
}

_lib = None


def lib():
    """The loaded library (loaded on first use); raises RuntimeError if it cannot be loaded."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(make -C tack-lbm_amd/csrc)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error():
    return lib().tlbm_last_error().decode(errors="replace")


def check(rc, what):
    """Raise RuntimeError (as lbm.py:533-534 does for an allocation failure) on a nonzero status."""
    if rc != TLBM_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def device_count():
    n = ctypes.c_int(0)
    rc = lib().tlbm_device_count(ctypes.byref(n))
    return n.value if rc == TLBM_OK else 0
