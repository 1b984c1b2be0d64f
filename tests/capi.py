"""ctypes declarations of the C-ABI (include/block_sigma.h) for tests that
call libblock_sigma.so directly, and a device-buffer helper."""
import ctypes

import numpy as np

import block_amd

vp = ctypes.c_void_p
lib = block_amd.sigma_lib
_i64, _i32, _d = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
for name, args in [
    ("bsg_init", [ctypes.POINTER(vp), ctypes.c_int]),
    ("bsg_free", [vp]),
    ("bsg_synchronize", [vp]),
    ("bsg_malloc", [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
    ("bsg_mfree", [vp, vp]),
    ("bsg_memcpy_h2d", [vp, vp, vp, ctypes.c_size_t]),
    ("bsg_memcpy_d2h", [vp, vp, vp, ctypes.c_size_t]),
    ("bsg_memset_zero", [vp, vp, ctypes.c_size_t]),
    ("bsg_dot", [vp, vp, vp, _i64, ctypes.POINTER(_d)]),
    ("bsg_precondition", [vp, vp, vp, _d, _d, _i64]),
    ("bsg_olsen_precondition", [vp, vp, vp, vp, _d, _d, vp, _i64]),
    ("bsg_gram", [vp, vp, vp, _i32, _i64, _i64, vp]),
    ("bsg_lincomb", [vp, vp, _i32, _i64, _i64, vp, _i32, vp, _i64]),
    ("bsg_gs_orthonormalise", [vp, vp, vp, _i32, _i64, _i64, _i32]),
    ("bsg_sigma_apply", [vp, vp, vp, vp, vp, _i32, _i32]),
    ("bsg_sigma_apply_noalloc", [vp, vp, vp, vp, vp, _i32, _i32]),
]:
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    f.argtypes = args
lib.bsg_last_error.restype = ctypes.c_char_p
lib.bsg_plan_partial_doubles.restype = ctypes.c_int64
lib.bsg_plan_partial_doubles.argtypes = [vp]


def check(rc):
    if rc != 0:
        raise RuntimeError(lib.bsg_last_error().decode())


class Device:
    """A context plus device buffers freed on close."""

    def __init__(self, device=0):
        self.ctx = vp()
        check(lib.bsg_init(ctypes.byref(self.ctx), device))
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = vp()
        check(lib.bsg_malloc(self.ctx, max(8, a.nbytes), ctypes.byref(p)))
        self.bufs.append(p)
        check(lib.bsg_memcpy_h2d(self.ctx, p, a.ctypes.data, a.nbytes))
        return p

    def zeros(self, n):
        return self.upload(np.zeros(n))

    def download(self, p, shape):
        out = np.empty(shape)
        check(lib.bsg_memcpy_d2h(self.ctx, out.ctypes.data, p, out.nbytes))
        return out

    def close(self):
        for p in self.bufs:
            lib.bsg_mfree(self.ctx, p)
        self.bufs = []
        lib.bsg_free(self.ctx)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
