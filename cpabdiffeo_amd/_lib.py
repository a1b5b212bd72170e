"""ctypes binding of libcpab_hip.so (the C ABI declared in include/cpab_hip.h).

The library is loaded from this package's _lib/ directory; there is no
fallback: if it is missing or no GPU is present, calls raise.
"""
import ctypes
import os
import threading

from .build import SO

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p

CPAB_OK, CPAB_ERR_ARG, CPAB_ERR_UNSUPPORTED, CPAB_ERR_ALIAS, CPAB_ERR_HIP, CPAB_ERR_NOMEM = range(6)

# name -> argtypes (every function returns int status unless listed in _RESTYPE)
_GEO = [_i, _i, _i, _d, _d, _d]
SIGNATURES = {
    "cpab_create": [_i, ctypes.POINTER(_p)],
    "cpab_destroy": [_p],
    "cpab_last_error": [_p],
    "cpab_set_stream": [_p, _p],
    "cpab_synchronize": [_p],
    "cpab_version": [],
    "cpab_prep_trels": [_p, _i, _i, _p, _i, _d, _p, _p],
    "cpab_expm": [_p, _i, _i, _p, _p, _i],
    "cpab_calc_T": [_p, _i, _i, _i, _p, _p, _p, _p, _d, _i, _i, _i64] + _GEO + [_p],
    "cpab_calc_T_prep": [_p, _i, _i, _i, _p, _p, _p, _i, _d, _i, _i, _i64] + _GEO + [_p, _p, _p],
    "cpab_calc_T_simple": [_p, _i, _i, _i, _p, _p, _p, _d, _i, _i, _i64] + _GEO,
    "cpab_calc_trajectory": [_p, _i, _i, _i, _p, _p, _p, _p, _d, _i, _i, _i64] + _GEO,
    "cpab_calc_grad_theta": [_p, _i, _i, _i, _p, _p, _p, _p, _p, _i, _d, _i, _i, _i64] + _GEO,
    "cpab_calc_T_simple_vjp": [_p, _i, _i, _i, _p, _p, _p, _d, _i, _i, _i64, _i, _i, _d, _d, _p, _p],
    "cpab_ssd_grad2d": [_p, _i, _i, _i, _p, _p, _p, _i, _i, _d, _i, _i, _i, _i, _d, _d, _p, _p, _p],
    "cpab_calc_T_vjp": [_p, _i, _i, _i, _p, _p, _p, _p, _d, _i, _i, _i64, _i, _i, _d, _d, _p, _p],
    "cpab_ssd_grad2d_calc_T": [_p, _i, _i, _i, _p, _p, _p, _p, _i, _i, _d, _i, _i, _i, _i, _d, _d, _p, _p, _p, _p,
                               _p],
    "cpab_batch_create": [_p, _i, _i, _i, _p, ctypes.POINTER(_p)],
    "cpab_batch_sizes": [_p, _p, _p, _p, _p],
    "cpab_batch_prep": [_p, _p, _i, _d, _p],
    "cpab_batch_calc_T": [_p, _p, _p, _p, _d, _i, _i, _p, _i, _p],
    "cpab_batch_destroy": [_p],
    "cpab_calc_T_items": [_p, _i, _i, _i, _p, _d, _i, _i, _p],
    "cpab_calc_T_items_noalloc": [_p, _i, _i, _i, _p, _d, _i, _i, _p],
    "cpab_theta2avees": [_p, _i, _i, _i, _p, _p, _p],
    "cpab_calc_v": [_p, _i, _i, _i, _p, _p, _p, _i64] + _GEO,
    "cpab_calc_cell_idx": [_p, _i, _i, _p, _p, _i64] + _GEO,
    "cpab_remap2d": [_p, _p, _p, _p, _i, _i],
    "cpab_remap3d": [_p, _p, _p, _p, _i, _i, _i],
    "cpab_ssd_ll": [_p, _p, _p, _i64, _d, _p],
    "cpab_reg2d_create": [_p, _p, ctypes.POINTER(_p)],
    "cpab_reg2d_init": [_p],
    "cpab_reg2d_run": [_p, _i, _i],
    "cpab_reg2d_last_proposal": [_p, _p, _p],
    "cpab_reg2d_spec_depth": [_p, ctypes.POINTER(_i)],
    "cpab_reg2d_set_spec_depth": [_p, _i],
    "cpab_reg2d_launch_counts": [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "cpab_reg2d_destroy": [_p],
    "cpab_prof_enable": [_p, _i],
    "cpab_prof_read": [_p, ctypes.POINTER(_i64), ctypes.POINTER(_d)],
}
_RESTYPE = {"cpab_last_error": ctypes.c_char_p, "cpab_version": ctypes.c_char_p}

_lib = None
_lock = threading.Lock()


class CpabError(RuntimeError):
    pass


def lib():
    """Load libcpab_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(SO):
                    raise CpabError("libcpab_hip.so is not built (run python -m cpabdiffeo_amd.build): "
                                    "the HIP path has no CPU fallback")
                L = ctypes.CDLL(SO)
                for name, args in SIGNATURES.items():
                    f = getattr(L, name)
                    f.argtypes = args
                    f.restype = _RESTYPE.get(name, _i)
                _lib = L
    return _lib


_EXC = {CPAB_ERR_ARG: ValueError, CPAB_ERR_UNSUPPORTED: NotImplementedError,
        CPAB_ERR_ALIAS: ValueError, CPAB_ERR_NOMEM: MemoryError}


def check(status, ctx=None):
    """Map a C status to the reference's exception types."""
    if status == CPAB_OK:
        return
    msg = lib().cpab_last_error(ctx).decode() if ctx else "status %d" % status
    raise _EXC.get(status, CpabError)(msg)


class Context:
    """One cpab_ctx per device; launches go to torch's current stream."""

    _by_device = {}

    def __init__(self, device):
        h = _p()
        st = lib().cpab_create(int(device), ctypes.byref(h))
        if st != CPAB_OK:
            raise CpabError("cpab_create(%d) failed with status %d (no usable GPU?)" % (device, st))
        self.handle = h
        self.device = device

    @classmethod
    def get(cls, device=0):
        c = cls._by_device.get(device)
        if c is None:
            c = cls._by_device[device] = Context(device)
        return c

    def bind_stream(self, stream_ptr):
        check(lib().cpab_set_stream(self.handle, stream_ptr), self.handle)

    def call(self, name, *args):
        check(getattr(lib(), name)(self.handle, *args), self.handle)

    def synchronize(self):
        self.call("cpab_synchronize")

    def prof_enable(self, on=True):
        self.call("cpab_prof_enable", 1 if on else 0)

    def prof_read(self):
        n = _i64()
        ms = _d()
        self.call("cpab_prof_read", ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value
