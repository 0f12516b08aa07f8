"""Python handle on the CPU oracle (oracle/mvdr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py to time the CPU baseline.  The product path
(multiviewdynrecon_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# record widths (doubles) of the exported arrays
WIDTH = {"dense_terms": 16, "density_terms": 16, "foreground_terms": 16, "feature_terms": 20,
         "cross_terms": 20, "smooth_terms": 9, "jtj_diag": 36, "jtj_diag_inv": 36, "jtr": 6,
         "jtj_offdiag": 36, "twist": 6, "node_se3": 8}
INDEX_ARRAYS = {"node2term_offset", "node2term_index", "nodepair_key", "nodepair_range",
                "nodepair_term_index", "term_offsets"}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmvdr_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        L = C.CDLL(path)
        L.mvdr_oracle_create.restype = C.c_void_p
        L.mvdr_oracle_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.mvdr_oracle_destroy.argtypes = [C.c_void_p]
        for name in ("build_index", "eval_terms", "assemble", "apply_update", "gn_iteration", "solve"):
            fn = getattr(L, "mvdr_oracle_" + name)
            fn.argtypes = [C.c_void_p]
            fn.restype = C.c_int
        L.mvdr_oracle_set_iteration.argtypes = [C.c_void_p, C.c_uint32]
        L.mvdr_oracle_pcg.argtypes = [C.c_void_p, C.c_uint32]
        L.mvdr_oracle_total_residual.argtypes = [C.c_void_p]
        L.mvdr_oracle_total_residual.restype = C.c_double
        L.mvdr_oracle_apply_jtj.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.mvdr_oracle_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
        L.mvdr_oracle_get.restype = C.c_longlong
        L.mvdr_oracle_set.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
        L.mvdr_oracle_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


class Oracle:
    """One frame's solve on the CPU.  precision: 'fp32' or 'fp64'."""

    def __init__(self, config, host_frame, precision="fp64"):
        self.frame = host_frame  # keeps host arrays alive
        self.config = config
        self.h = lib().mvdr_oracle_create(C.byref(config), C.byref(host_frame.struct),
                                          0 if precision == "fp32" else 1)
        if not self.h:
            raise RuntimeError(lib().mvdr_oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().mvdr_oracle_destroy(self.h)
            self.h = None

    def _call(self, name, *args):
        rc = getattr(lib(), "mvdr_oracle_" + name)(self.h, *args)
        if rc != 0:
            raise RuntimeError(f"oracle {name} failed ({rc})")

    def build_index(self):
        self._call("build_index")

    def set_iteration(self, k):
        self._call("set_iteration", k)

    def eval_terms(self):
        self._call("eval_terms")

    def assemble(self):
        self._call("assemble")

    def pcg(self, iters):
        self._call("pcg", iters)

    def apply_update(self):
        self._call("apply_update")

    def gn_iteration(self):
        self._call("gn_iteration")

    def solve(self):
        self._call("solve")

    def total_residual(self):
        return lib().mvdr_oracle_total_residual(self.h)

    def apply_jtj(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().mvdr_oracle_apply_jtj(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def get(self, name):
        L = lib()
        size = L.mvdr_oracle_get(self.h, name.encode(), None, 0)
        if size < 0:
            raise KeyError(name)
        dtype = np.uint32 if name in INDEX_ARRAYS else np.float64
        out = np.empty(size // np.dtype(dtype).itemsize, dtype=dtype)
        if size:
            L.mvdr_oracle_get(self.h, name.encode(), out.ctypes.data, size)
        w = WIDTH.get(name)
        return out.reshape(-1, w) if w else out

    def set(self, name, values):
        values = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if lib().mvdr_oracle_set(self.h, name.encode(), values.ctypes.data, values.size) != 0:
            raise ValueError(f"oracle set {name}: size mismatch")
