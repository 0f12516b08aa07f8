"""Build `mvdr_frame_input` structs from a synthetic `Scene`.

`host_frame(scene)` points the struct at numpy arrays (host memory, used by the
CPU oracle); `DeviceFrame(scene)` uploads every array to the GPU with torch
and points the struct at device memory (used by the HIP solver).  Both keep
references to the backing arrays for as long as the returned object lives.
"""
import ctypes as C

import numpy as np

from . import _abi

# (scene map name, frame field, dtype)
MAP_FIELDS = [
    ("vertex_confid", "vertex_confid"), ("normal_radius", "normal_radius"),
    ("filtered_foreground", "filtered_foreground"), ("foreground_gradient", "foreground_gradient"),
    ("gray", "gray"), ("gray_gradient", "gray_gradient"), ("edge_mask", "edge_mask"),
    ("reference_vertex", "reference_vertex"), ("reference_normal", "reference_normal"),
    ("warp_vertex", "warp_vertex"), ("index_map", "index_map"), ("normalized_rgb", "normalized_rgb"),
]


def _fill_common(fi, scene):
    fi.num_views = scene.num_views
    fi.width = scene.width
    fi.height = scene.height
    for v, view in enumerate(scene.views):
        fi.intrinsic[v] = _abi.Intrinsic(*[float(x) for x in view.intrinsic])
        m = _abi.Mat34()
        m.rot[:] = [float(x) for x in view.world2camera[:9]]
        m.trans[:] = [float(x) for x in view.world2camera[9:12]]
        fi.world2camera[v] = m
        m = _abi.Mat34()
        m.rot[:] = [float(x) for x in view.initial_camera_se3[:9]]
        m.trans[:] = [float(x) for x in view.initial_camera_se3[9:12]]
        fi.initial_camera_se3[v] = m
    fi.num_surfels = scene.num_surfels
    fi.num_nodes = scene.num_nodes
    fi.num_graph_edges = int(scene.arrays["node_graph"].shape[0])
    fi.num_cross_pairs = int(scene.arrays["cross_pairs"].shape[0])
    for v, view in enumerate(scene.views):
        fi.num_feature_pairs[v] = int(view.feature_pairs.shape[0])


class HostFrame:
    """mvdr_frame_input over host (numpy) memory.  `corr_map` attaches each
    view's correspondence map (scene map "corr_map") to the frame."""

    def __init__(self, scene, node_se3=None, corr_map=False):
        self.scene = scene
        self.keep = []
        fi = _abi.FrameInput()
        _fill_common(fi, scene)

        def ptr(a):
            a = np.ascontiguousarray(a)
            self.keep.append(a)
            return a.ctypes.data if a.size else None

        for v, view in enumerate(scene.views):
            for name, fld in MAP_FIELDS:
                getattr(fi, fld)[v] = ptr(view.maps[name])
            fi.feature_pairs[v] = ptr(view.feature_pairs)
            if corr_map:
                fi.corr_map[v] = ptr(view.maps["corr_map"])
        A = scene.arrays
        fi.surfel_knn = ptr(A["surfel_knn"])
        fi.surfel_knn_weight = ptr(A["surfel_knn_weight"])
        se3 = A["node_se3_init"] if node_se3 is None else np.asarray(node_se3, dtype=np.float32).reshape(-1, 8)
        fi.node_se3 = ptr(se3)
        fi.canonical_nodes = ptr(A["nodes"])
        fi.node_graph = ptr(A["node_graph"])
        fi.cross_pairs = ptr(A["cross_pairs"])
        self.struct = fi


class DeviceFrame:
    """mvdr_frame_input over device memory (torch tensors on `device`)."""

    def __init__(self, scene, device="cuda", node_se3=None, corr_map=False):
        import torch

        self.scene = scene
        self.keep = []
        fi = _abi.FrameInput()
        _fill_common(fi, scene)

        def ptr(a):
            a = np.ascontiguousarray(a)
            if a.size == 0:
                return None
            t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).to(device)
            self.keep.append(t)
            return t.data_ptr()

        for v, view in enumerate(scene.views):
            for name, fld in MAP_FIELDS:
                getattr(fi, fld)[v] = ptr(view.maps[name])
            fi.feature_pairs[v] = ptr(view.feature_pairs)
            if corr_map:
                fi.corr_map[v] = ptr(view.maps["corr_map"])
        A = scene.arrays
        fi.surfel_knn = ptr(A["surfel_knn"])
        fi.surfel_knn_weight = ptr(A["surfel_knn_weight"])
        se3 = A["node_se3_init"] if node_se3 is None else np.asarray(node_se3, dtype=np.float32).reshape(-1, 8)
        fi.node_se3 = ptr(se3)
        fi.canonical_nodes = ptr(A["nodes"])
        fi.node_graph = ptr(A["node_graph"])
        fi.cross_pairs = ptr(A["cross_pairs"])
        self.struct = fi
