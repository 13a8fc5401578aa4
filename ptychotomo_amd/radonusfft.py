"""Drop-in for the reference SWIG class ``radonusfft``
(src/include/radonusfft.cuh:3-37, src/cuda/radonusfft.i:15-46).

Same constructor and methods; buffers are raw device addresses (ints, e.g.
``tensor.data_ptr()``), output first, unnormalised.  Work is enqueued on
PyTorch's current HIP stream.
"""
import ctypes

from . import _lib
from ._dev import stream


class radonusfft:
    def __init__(self, Ntheta, Nz, N):
        h = ctypes.c_void_p()
        _lib.call("pt_tomo_create", int(Ntheta), int(Nz), int(N), ctypes.byref(h))
        self._h = h
        self.Ntheta, self.Nz, self.N = int(Ntheta), int(Nz), int(N)

    def setobj(self, theta):
        _lib.call("pt_tomo_set_theta", self._h, ctypes.c_void_p(int(theta)), stream())

    def fwd(self, g, f):
        _lib.call("pt_tomo_fwd", self._h, ctypes.c_void_p(int(g)), ctypes.c_void_p(int(f)), stream())

    def adj(self, f, g):
        _lib.call("pt_tomo_adj", self._h, ctypes.c_void_p(int(f)), ctypes.c_void_p(int(g)), stream())

    def setfilter(self, kind):
        """Filter of later adjoints: "ramp" (applyfilter, kernels.cuh:174-187,
        i.e. filtered backprojection) or None (radonusfft::adj)."""
        _lib.call("pt_tomo_set_filter", self._h, _lib.PT_FILTER[kind])

    # scaled variants used by Solver (coeftomo folded into the last kernel)
    def fwd_scaled(self, g, f, scale):
        _lib.call("pt_tomo_fwd_scaled", self._h, ctypes.c_void_p(int(g)), ctypes.c_void_p(int(f)),
                  float(scale), stream())

    def adj_scaled(self, f, g, scale):
        _lib.call("pt_tomo_adj_scaled", self._h, ctypes.c_void_p(int(f)), ctypes.c_void_p(int(g)),
                  float(scale), stream())

    def free(self):
        if getattr(self, "_h", None):
            _lib.lib.pt_tomo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
