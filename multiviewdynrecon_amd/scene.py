"""Seeded synthetic scenes for the non-rigid solve (wraps libmvdr_scene.so).

`make_scene(workload)` returns a `Scene` holding every array a solve needs as
numpy arrays on the host.  Workload ids follow the benchmark list: 0 = one-view
sphere, 1 = three-view body at ~4k nodes, 3 = the same body scaled to ~8k
nodes; any field of the C parameter struct can be overridden by keyword.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _abi

_LIB = None
_HERE = os.path.dirname(os.path.abspath(__file__))

PER_VIEW = {
    "vertex_confid": (np.float32, 4), "normal_radius": (np.float32, 4), "foreground_mask": (np.uint8, 1),
    "filtered_foreground": (np.float32, 1), "foreground_gradient": (np.float32, 2), "gray": (np.float32, 1),
    "gray_gradient": (np.float32, 2), "edge_mask": (np.uint8, 1), "reference_vertex": (np.float32, 4),
    "reference_normal": (np.float32, 4), "warp_vertex": (np.float32, 4), "index_map": (np.uint32, 1),
    "normalized_rgb": (np.float32, 4),
}
GLOBAL = {
    "surfel_pos": (np.float32, 4), "surfel_normal": (np.float32, 4), "surfel_knn": (np.uint16, 4),
    "surfel_knn_weight": (np.float32, 4), "nodes": (np.float32, 4), "node_graph": (np.uint16, 2),
    "node_se3_init": (np.float32, 8), "node_se3_gt": (np.float32, 8), "cross_pairs": (np.uint16, 6),
}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmvdr_scene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        L = C.CDLL(path)
        L.mvdr_scene_default_params.argtypes = [C.c_int, C.POINTER(_abi.SceneParams)]
        L.mvdr_scene_create.argtypes = [C.POINTER(_abi.SceneParams), C.POINTER(C.c_void_p)]
        L.mvdr_scene_destroy.argtypes = [C.c_void_p]
        L.mvdr_scene_get_info.argtypes = [C.c_void_p, C.POINTER(_abi.SceneInfo)]
        L.mvdr_scene_get.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32, C.c_void_p, C.c_size_t]
        L.mvdr_scene_get.restype = C.c_longlong
        L.mvdr_scene_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


@dataclass
class View:
    intrinsic: np.ndarray            # (4,) fx, fy, cx, cy
    world2camera: np.ndarray         # (12,) mvdr_mat34 floats
    initial_camera_se3: np.ndarray   # (12,)
    maps: dict = field(default_factory=dict)
    feature_pairs: np.ndarray = None  # (F, 4) uint16


@dataclass
class Scene:
    params: dict
    width: int
    height: int
    num_views: int
    voxel_size: float
    arrays: dict
    views: list

    @property
    def num_nodes(self):
        return int(self.arrays["nodes"].shape[0])

    @property
    def num_surfels(self):
        return int(self.arrays["surfel_pos"].shape[0])


def default_params(workload=1, **overrides):
    p = _abi.SceneParams()
    if lib().mvdr_scene_default_params(int(workload), C.byref(p)) != 0:
        raise ValueError(f"unknown scene workload {workload}")
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _get(handle, name, view, dtype, comps):
    L = lib()
    size = L.mvdr_scene_get(handle, name.encode(), view, None, 0)
    if size < 0:
        raise KeyError(name)
    arr = np.empty(size // np.dtype(dtype).itemsize, dtype=dtype)
    if size:
        L.mvdr_scene_get(handle, name.encode(), view, arr.ctypes.data, size)
    return arr.reshape(-1, comps) if comps > 1 else arr


def make_scene(workload=1, **overrides):
    """Generate a seeded synthetic scene (deterministic for given parameters)."""
    L = lib()
    p = default_params(workload, **overrides)
    h = C.c_void_p()
    if L.mvdr_scene_create(C.byref(p), C.byref(h)) != 0:
        raise RuntimeError("scene generation failed: " + L.mvdr_scene_last_error().decode())
    try:
        info = _abi.SceneInfo()
        L.mvdr_scene_get_info(h, C.byref(info))
        arrays = {k: _get(h, k, 0, dt, c) for k, (dt, c) in GLOBAL.items()}
        views = []
        for v in range(info.num_views):
            intr = _get(h, "intrinsic", v, np.float32, 1)
            w2c = _get(h, "world2camera", v, np.float32, 1)
            init = _get(h, "initial_camera_se3", v, np.float32, 1)
            vw = View(intr, w2c, init)
            for k, (dt, c) in PER_VIEW.items():
                a = _get(h, k, v, dt, c)
                vw.maps[k] = a.reshape(info.height, info.width, c) if c > 1 else a.reshape(info.height, info.width)
            vw.feature_pairs = _get(h, "feature_pairs", v, np.uint16, 4)
            vw.maps["corr_map"] = _get(h, "corr_map", v, np.uint16, 4).reshape(info.height, info.width, 4)
            views.append(vw)
        params = {name: getattr(p, name) for name, _ in _abi.SceneParams._fields_}
        return Scene(params, info.width, info.height, info.num_views, float(info.voxel_size), arrays, views)
    finally:
        L.mvdr_scene_destroy(h)


UNEVEN_SEED = 0x4D564452 + 100


def uneven_scene_params(count=16):
    """Sizes of the uneven-batch workload: scene s has N ~ U[1000, 8000] nodes
    (about 37.5 surfels per node) and 1-3 views, drawn from one seeded stream."""
    rng = np.random.default_rng(UNEVEN_SEED)
    nodes = rng.integers(1000, 8001, count)
    views = rng.integers(1, 4, count)
    return [dict(seed=UNEVEN_SEED + s, target_nodes=int(nodes[s]), target_surfels=int(37.5 * nodes[s]),
                 num_views=int(views[s]), cross_pairs=2000 if views[s] > 1 else 0) for s in range(count)]


def make_uneven_scenes(count=16, which=None):
    """The scenes of the uneven-batch workload (all, or the indices in `which`)."""
    params = uneven_scene_params(count)
    idx = range(count) if which is None else which
    return [make_scene(4, **params[s]) for s in idx]
