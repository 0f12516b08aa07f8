"""Python mirror of SparseSurfelFusion::NonRigidSolver over the C ABI.

Method names follow the reference class (core/NonRigidSolver/NonRigidSolver.h):
SetSolverInput, SolveNonRigidAlignment, SolvedNodeSE3,
ComputeTotalResidualSynced, JtResidualValue, JtJDiagonalBlockValue, plus the
two halves of the streamed solve (BuildSolverIndex, SolverIteration) used by
stage-wise parity checks.  Every call goes to libmvdr_solver.so; there is no
other implementation behind it.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from .config import default_config

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# tap -> (dtype, record width in elements of dtype)
TAP_FORMAT = {
    "node2term_offset": (np.uint32, 1), "node2term_index": (np.uint32, 1), "nodepair_key": (np.uint32, 1),
    "nodepair_range": (np.uint32, 2), "nodepair_term_index": (np.uint32, 1), "dense_terms": (np.uint32, 16),
    "smooth_terms": (np.uint32, 8), "feature_terms": (np.uint32, 16), "cross_terms": (np.uint32, 16),
    "foreground_terms": (np.uint32, 16), "dense_pixels": (np.uint16, 4), "jtj_diag": (np.float32, 36),
    "jtj_diag_inv": (np.float32, 36), "jtr": (np.float32, 6), "jtj_offdiag": (np.float32, 36),
    "twist": (np.float32, 6), "node_se3": (np.float32, 8), "penalty": (np.float32, 1),
    "density_terms": (np.uint32, 16), "spmv_input": (np.float32, 6), "spmv_output": (np.float32, 6),
    "foreground_pixels": (np.uint16, 4),
}


class SolverError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_abi.STATUS.get(status, status)}: {message}")
        self.status = status


def lib():
    """Load libmvdr_solver.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmvdr_solver.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: the HIP solver library must be built (__graft_entry__.build())")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 /
        # libhsa-runtime64.so.1 with the same sonames as /opt/rocm's.  Whichever
        # loads first serves the whole process, and torch finds no GPU when
        # /opt/rocm's copy came first, so torch's runtime is loaded before ours.
        import torch  # noqa: F401

        L = C.CDLL(path)
        L.mvdr_solver_create.argtypes = [C.POINTER(_abi.SolverConfig), C.POINTER(C.c_void_p)]
        L.mvdr_solver_destroy.argtypes = [C.c_void_p]
        L.mvdr_solver_last_error.argtypes = [C.c_void_p]
        L.mvdr_solver_last_error.restype = C.c_char_p
        L.mvdr_solver_set_input.argtypes = [C.c_void_p, C.POINTER(_abi.FrameInput)]
        for name in ("solve", "build_index", "gn_iteration"):
            getattr(L, "mvdr_solver_" + name).argtypes = [C.c_void_p, C.c_void_p]
        L.mvdr_solver_get_node_se3.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]
        L.mvdr_solver_total_residual.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.mvdr_solver_get_sizes.argtypes = [C.c_void_p, C.POINTER(_abi.SolverSizes)]
        L.mvdr_solver_debug_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.mvdr_solver_default_config.argtypes = [C.POINTER(_abi.SolverConfig)]
        L.mvdr_status_string.restype = C.c_char_p
        L.mvdr_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_void_p)]
        L.mvdr_batch_destroy.argtypes = [C.c_void_p]
        L.mvdr_batch_last_error.argtypes = [C.c_void_p]
        L.mvdr_batch_last_error.restype = C.c_char_p
        L.mvdr_batch_solve.argtypes = [C.c_void_p, C.c_void_p]
        L.mvdr_batch_bench_stage.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_void_p, C.POINTER(C.c_float)]
        L.mvdr_solver_bench_stage.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_void_p, C.POINTER(C.c_float)]
        L.mvdr_solve_noalloc.argtypes = [C.POINTER(_abi.SolverConfig), C.POINTER(_abi.FrameInput), C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        L.mvdr_solver_debug_jacobian.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        LP = C.POINTER(_abi.LaunchParams)
        L.mvdr_solver_autotune.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, LP]
        L.mvdr_batch_autotune.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, LP]
        L.mvdr_batch_set_lanes.argtypes = [C.c_void_p, C.c_uint32]
        L.mvdr_batch_lanes.argtypes = [C.c_void_p]
        L.mvdr_batch_lanes.restype = C.c_uint32
        L.mvdr_launch_params_get.argtypes = [C.c_uint32, C.c_uint32, LP]
        L.mvdr_launch_params_set.argtypes = [C.c_uint32, C.c_uint32, LP]
        L.mvdr_launch_table_save.argtypes = [C.c_char_p]
        L.mvdr_launch_table_load.argtypes = [C.c_char_p]
        _LIB = L
    return _LIB


_HIP = None


def _memcpy_d2h(dst, src, nbytes):
    """hipMemcpy device -> host through the process's HIP runtime (the one torch
    loaded, found by its soname)."""
    global _HIP
    if _HIP is None:
        import torch  # noqa: F401  (loads the runtime)

        _HIP = C.CDLL("libamdhip64.so.7")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    rc = _HIP.hipMemcpy(dst, src, nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise SolverError(4, f"hipMemcpy failed ({rc})")


def _stream_handle(stream):
    if stream is None:
        import torch

        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


class NonRigidSolver:
    """One solver instance (owns all device scratch memory, like the reference)."""

    def __init__(self, config=None, **overrides):
        self.config = config if config is not None else default_config(**overrides)
        self.h = C.c_void_p()
        rc = lib().mvdr_solver_create(C.byref(self.config), C.byref(self.h))
        if rc != 0:
            raise SolverError(rc, "mvdr_solver_create failed")
        self._frame = None

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mvdr_solver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if lib is not None:  # module globals are cleared at interpreter exit
            self.close()

    def _check(self, rc):
        if rc != 0:
            raise SolverError(rc, lib().mvdr_solver_last_error(self.h).decode())

    # -- NonRigidSolver::SetSolverInput
    def SetSolverInput(self, frame):
        """frame: a DeviceFrame (or anything with a `.struct` mvdr_frame_input)."""
        self._frame = frame  # keeps device buffers alive
        self._check(lib().mvdr_solver_set_input(self.h, C.byref(frame.struct)))

    # -- NonRigidSolver::SolveNonRigidAlignment
    def SolveNonRigidAlignment(self, stream=None):
        self._check(lib().mvdr_solver_solve(self.h, _stream_handle(stream)))

    # -- buildSolverIndexStreamed / solverIteration*Streamed
    def BuildSolverIndex(self, stream=None):
        self._check(lib().mvdr_solver_build_index(self.h, _stream_handle(stream)))

    def SolverIteration(self, stream=None):
        self._check(lib().mvdr_solver_gn_iteration(self.h, _stream_handle(stream)))

    # -- NonRigidSolver::SolvedNodeSE3: (device pointer, node count)
    def SolvedNodeSE3(self):
        p = C.c_void_p()
        n = C.c_uint32()
        self._check(lib().mvdr_solver_get_node_se3(self.h, C.byref(p), C.byref(n)))
        return p.value, n.value

    def ComputeTotalResidualSynced(self, stream=None):
        out = C.c_double()
        self._check(lib().mvdr_solver_total_residual(self.h, _stream_handle(stream), C.byref(out)))
        return out.value

    def sizes(self):
        s = _abi.SolverSizes()
        self._check(lib().mvdr_solver_get_sizes(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _abi.SolverSizes._fields_}

    def autotune(self, reps=10, stream=None):
        """Times the tunable launch shapes on this solver's last solve and keeps
        the fastest in the process-wide launch table; returns the chosen
        workgroup counts (no result changes)."""
        out = _abi.LaunchParams()
        self._check(lib().mvdr_solver_autotune(self.h, reps, _stream_handle(stream), C.byref(out)))
        return launch_params_dict(out)

    def debug_jacobian(self, values, transpose=False):
        """The term Jacobian at the current warp, re-evaluated from the maps with
        unit term weights (MVDR_JTJ_RECOMPUTE): J x for x of shape (nodes, 6)
        (reference twists) -> (terms, 4) rows, or J^T y for y of shape
        (terms, 4) -> (nodes, 6) with transpose=True."""
        import torch

        sz = self.sizes()
        terms = sum(sz[k] for k in ("num_dense_terms", "num_smooth_terms", "num_foreground_terms",
                                    "num_feature_terms", "num_cross_terms"))
        shape_in, shape_out = ((terms, 4), (sz["num_nodes"], 6)) if transpose else ((sz["num_nodes"], 6), (terms, 4))
        host = np.ascontiguousarray(values, dtype=np.float32)
        if host.shape != shape_in:
            raise ValueError(f"expected shape {shape_in}, got {host.shape}")
        src = torch.from_numpy(host).to("cuda")
        out = torch.empty(shape_out, dtype=torch.float32, device="cuda")
        self._check(lib().mvdr_solver_debug_jacobian(self.h, int(bool(transpose)), C.c_void_p(src.data_ptr()),
                                                      C.c_void_p(out.data_ptr()), _stream_handle(None)))
        return out.cpu().numpy()

    def debug_get(self, tap):
        """Copy a stage buffer to a host numpy array (synchronises)."""
        code = _abi.TAP[tap]
        need = C.c_size_t()
        rc = lib().mvdr_solver_debug_get(self.h, code, None, 0, C.byref(need))
        if rc not in (0, 3):
            self._check(rc)
        dtype, width = TAP_FORMAT[tap]
        out = np.empty(need.value // np.dtype(dtype).itemsize, dtype=dtype)
        self._check(lib().mvdr_solver_debug_get(self.h, code, out.ctypes.data if out.size else None, need.value,
                                                C.byref(need)))
        return out.reshape(-1, width) if width > 1 else out

    # reference accessors (NonRigidSolver.h:329-330)
    def JtResidualValue(self):
        return self.debug_get("jtr").reshape(-1)

    def JtJDiagonalBlockValue(self):
        return self.debug_get("jtj_diag").reshape(-1)

    def solved_node_se3_host(self):
        return self.debug_get("node_se3")


def decode_scalar_terms(raw):
    """Split MVDR_TAP_{DENSE,FOREGROUND,DENSITY}_TERMS records (16 x u32 each)."""
    raw = np.asarray(raw).reshape(-1, 16)
    f = raw.view(np.float32)
    knn = np.stack([raw[:, 0] & 0xFFFF, raw[:, 0] >> 16, raw[:, 1] & 0xFFFF, raw[:, 1] >> 16], 1).astype(np.uint16)
    return {"knn": knn, "r": f[:, 2].copy(), "w": f[:, 4:8].copy(), "J": f[:, 8:14].copy()}


def decode_smooth_terms(raw):
    """MVDR_TAP_SMOOTH_TERMS records (8 x u32): i | j << 16, valid, T_i x_j, and the
    residual r = T_i x_j - T_j x_j; T_j x_j is returned in float64 as T_i x_j - r."""
    raw = np.asarray(raw).reshape(-1, 8)
    f = raw.view(np.float32)
    ti = f[:, 2:5].astype(np.float64)
    r = f[:, 5:8].astype(np.float64)
    return {"i": (raw[:, 0] & 0xFFFF).astype(np.uint16), "j": (raw[:, 0] >> 16).astype(np.uint16),
            "valid": raw[:, 1].astype(np.uint8), "Ti_xj": ti, "r": r, "Tj_xj": ti - r}


def decode_p2p_terms(raw):
    raw = np.asarray(raw).reshape(-1, 16)
    f = raw.view(np.float32)
    knn = np.stack([raw[:, 0] & 0xFFFF, raw[:, 0] >> 16, raw[:, 1] & 0xFFFF, raw[:, 1] >> 16], 1).astype(np.uint16)
    return {"knn": knn, "w": f[:, 4:8].copy(), "target": f[:, 8:11].copy(), "warped": f[:, 12:15].copy()}


class NonRigidBatch:
    """Solves the current frames of several NonRigidSolver instances with one
    launch per stage (no counterpart in the reference, which solves one frame at
    a time).  The members must share their iteration schedule and precision and
    must outlive the batch; each member reads as after its own solve."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        self.h = C.c_void_p()
        rc = lib().mvdr_batch_create(arr, len(self.solvers), C.byref(self.h))
        if rc != 0:
            raise SolverError(rc, "mvdr_batch_create failed")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mvdr_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if lib is not None:  # module globals are cleared at interpreter exit
            self.close()

    def _check(self, rc):
        if rc != 0:
            raise SolverError(rc, lib().mvdr_batch_last_error(self.h).decode())

    def SolveNonRigidAlignment(self, stream=None):
        self._check(lib().mvdr_batch_solve(self.h, _stream_handle(stream)))

    def autotune(self, reps=10, stream=None):
        """NonRigidSolver.autotune for the batch's shape."""
        out = _abi.LaunchParams()
        self._check(lib().mvdr_batch_autotune(self.h, reps, _stream_handle(stream), C.byref(out)))
        return launch_params_dict(out)

    def bench_stage(self, stage, reps, stream=None):
        """Mean ms per launch of one stage over all members (hipEvents on the stream)."""
        out = C.c_float()
        self._check(lib().mvdr_batch_bench_stage(self.h, stage, reps, _stream_handle(stream), C.byref(out)))
        return out.value

    @property
    def lanes(self):
        """Solve lanes: runs of members solved concurrently on their own streams."""
        return lib().mvdr_batch_lanes(self.h)

    def set_lanes(self, lanes):
        self._check(lib().mvdr_batch_set_lanes(self.h, lanes))


# ------------------------------------------------------------- launch table
def launch_params_dict(p):
    return {name: getattr(p, name) for name, _ in _abi.LaunchParams._fields_}


def launch_params(nodes, frames):
    """Workgroup counts the launches for this shape use (table entry or rule)."""
    out = _abi.LaunchParams()
    rc = lib().mvdr_launch_params_get(nodes, frames, C.byref(out))
    if rc != 0:
        raise SolverError(rc, "mvdr_launch_params_get")
    return launch_params_dict(out)


def set_launch_params(nodes, frames, cg_workgroups=0, eval_workgroups=0, transform_workgroups=0, gram_rows=0):
    p = _abi.LaunchParams(cg_workgroups, eval_workgroups, transform_workgroups, gram_rows)
    rc = lib().mvdr_launch_params_set(nodes, frames, C.byref(p))
    if rc != 0:
        raise SolverError(rc, "mvdr_launch_params_set")


def clear_launch_table():
    lib().mvdr_launch_table_clear()


def save_launch_table(path):
    rc = lib().mvdr_launch_table_save(str(path).encode())
    if rc != 0:
        raise SolverError(rc, f"cannot write {path}")


def load_launch_table(path):
    rc = lib().mvdr_launch_table_load(str(path).encode())
    if rc != 0:
        raise SolverError(rc, f"cannot load {path}")
