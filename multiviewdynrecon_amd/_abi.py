"""ctypes mirrors of include/mvdr/types.h, solver.h and scene.h.

Field order and types match the C headers exactly; `tests/test_abi.py`
checks the struct sizes against the C compiler's view of them.
"""
import ctypes as C

MAX_VIEWS = 3

u16 = C.c_uint16
u32 = C.c_uint32
f32 = C.c_float
vp = C.c_void_p


class Intrinsic(C.Structure):
    _fields_ = [("fx", f32), ("fy", f32), ("cx", f32), ("cy", f32)]


class Mat34(C.Structure):
    _fields_ = [("rot", f32 * 9), ("trans", f32 * 3)]


class DQ(C.Structure):
    _fields_ = [("q0", f32 * 4), ("q1", f32 * 4)]


class CrossPair(C.Structure):
    _fields_ = [("pixel", u16 * 4), ("view", u16 * 2)]


class SolverConfig(C.Structure):
    _fields_ = [
        ("max_views", u32), ("max_width", u32), ("max_height", u32), ("max_nodes", u32),
        ("max_graph_edges", u32), ("max_feature_pairs", u32), ("max_cross_pairs", u32),
        ("gn_iterations", u32), ("pcg_iterations", u32), ("global_iterations", u32),
        ("use_elastic_penalty", u32), ("use_density", u32), ("use_foreground", u32),
        ("use_cross_view", u32), ("density_in_jtj", u32), ("precision", u32), ("jtj_mode", u32),
        ("refine_steps", u32), ("use_graph", u32),
    ]


V = MAX_VIEWS


class FrameInput(C.Structure):
    _fields_ = [
        ("num_views", u32), ("width", u32), ("height", u32),
        ("intrinsic", Intrinsic * V), ("world2camera", Mat34 * V), ("initial_camera_se3", Mat34 * V),
        ("vertex_confid", vp * V), ("normal_radius", vp * V), ("filtered_foreground", vp * V),
        ("foreground_gradient", vp * V), ("gray", vp * V), ("gray_gradient", vp * V), ("edge_mask", vp * V),
        ("reference_vertex", vp * V), ("reference_normal", vp * V), ("warp_vertex", vp * V),
        ("index_map", vp * V), ("normalized_rgb", vp * V),
        ("surfel_knn", vp), ("surfel_knn_weight", vp), ("num_surfels", u32),
        ("node_se3", vp), ("canonical_nodes", vp), ("node_graph", vp), ("num_nodes", u32), ("num_graph_edges", u32),
        ("feature_pairs", vp * V), ("num_feature_pairs", u32 * V),
        ("cross_pairs", vp), ("num_cross_pairs", u32),
        ("corr_map", vp * V),
    ]


class LaunchParams(C.Structure):
    _fields_ = [("cg_workgroups", u32), ("eval_workgroups", u32), ("transform_workgroups", u32), ("gram_rows", u32)]


class SolverSizes(C.Structure):
    _fields_ = [
        ("num_dense_terms", u32), ("num_smooth_terms", u32), ("num_foreground_terms", u32),
        ("num_feature_terms", u32), ("num_cross_terms", u32), ("num_node2term", u32),
        ("num_nodepair_terms", u32), ("num_node_pairs", u32), ("num_nodes", u32), ("gn_iterations_done", u32),
    ]


class SceneParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("shape", u32), ("num_views", u32), ("width", u32), ("height", u32),
        ("fx", f32), ("fy", f32), ("target_surfels", u32), ("target_nodes", u32),
        ("cam_distance_min", f32), ("cam_distance_max", f32), ("features_per_view", u32), ("cross_pairs", u32),
        ("depth_noise_sigma", f32), ("outlier_fraction", f32), ("max_rot_deg", f32), ("max_trans", f32),
        ("feature_jitter_px", f32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("num_surfels", u32), ("num_nodes", u32), ("num_graph_edges", u32), ("num_views", u32),
        ("width", u32), ("height", u32), ("num_feature_pairs", u32 * V), ("num_cross_pairs", u32),
        ("voxel_size", f32),
    ]


# mvdr_tap
TAP = {
    "node2term_offset": 0, "node2term_index": 1, "nodepair_key": 2, "nodepair_range": 3,
    "nodepair_term_index": 4, "dense_terms": 5, "smooth_terms": 6, "feature_terms": 7, "cross_terms": 8,
    "foreground_terms": 9, "dense_pixels": 10, "jtj_diag": 11, "jtj_diag_inv": 12, "jtr": 13,
    "jtj_offdiag": 14, "twist": 15, "node_se3": 16, "penalty": 17, "density_terms": 18,
    "spmv_input": 19, "spmv_output": 20, "foreground_pixels": 21,
}

MAX_PCG_ITERATIONS = 64  # MVDR_MAX_PCG_ITERATIONS

STATUS = {0: "MVDR_OK", 1: "MVDR_E_INVALID", 2: "MVDR_E_SHAPE", 3: "MVDR_E_CAPACITY", 4: "MVDR_E_HIP",
          5: "MVDR_E_STATE"}


class CrossViewConfig(C.Structure):
    _fields_ = [("max_views", u32), ("max_width", u32), ("max_height", u32), ("max_pairs", u32),
                ("search_radius", u32), ("search_step", u32), ("track_radius", u32), ("max_distance2", f32)]


class CrossViewInput(C.Structure):
    _fields_ = [("num_views", u32), ("width", u32), ("height", u32), ("intrinsic", Intrinsic * V),
                ("initial_camera_se3", Mat34 * V), ("vertex", vp * V), ("edge_mask", vp * V), ("corr_map", vp * V)]


CROSSVIEW_LISTS = {"candidates": 0, "unique": 1, "backtraced": 2}


class RenderInput(C.Structure):
    _fields_ = [("num_views", u32), ("width", u32), ("height", u32), ("num_surfels", u32),
                ("intrinsic", Intrinsic * V), ("world2camera", Mat34 * V), ("reference_vertex", vp * V),
                ("reference_normal", vp * V), ("warp_vertex", vp * V), ("warp_normal", vp * V), ("color_time", vp),
                ("max_depth", f32), ("current_time", f32), ("confidence_threshold", f32), ("time_threshold", f32)]


class RenderOutput(C.Structure):
    _fields_ = [("reference_vertex", vp * V), ("reference_normal", vp * V), ("warp_vertex", vp * V),
                ("warp_normal", vp * V), ("index_map", vp * V), ("normalized_rgb", vp * V), ("zbuffer", vp * V)]


class RigidConfig(C.Structure):
    _fields_ = [("max_views", u32), ("max_width", u32), ("max_height", u32), ("reference_reduction", u32)]


class RigidInput(C.Structure):
    _fields_ = [("num_views", u32), ("width", u32), ("height", u32), ("intrinsic", Intrinsic * V),
                ("reference_vertex", vp * V), ("reference_normal", vp * V), ("conversion_vertex", vp * V),
                ("conversion_normal", vp * V), ("world2camera", Mat34 * V), ("relative_pose", Mat34 * V)]


class NodeCorrection(C.Structure):
    _fields_ = [("num_views", u32), ("width", u32), ("height", u32), ("intrinsic", Intrinsic * V),
                ("initial_camera_se3", Mat34 * V), ("world2camera", Mat34 * V), ("previous_vertex", vp * V),
                ("guided_se3", vp * V), ("guided_valid", vp * V), ("num_nodes", u32), ("node_error", vp),
                ("live_nodes", vp), ("node_se3", vp), ("node_knn", vp), ("node_knn_weight", vp),
                ("num_surfels", u32), ("live_vertex", vp), ("live_normal", vp), ("surfel_knn", vp),
                ("surfel_knn_weight", vp), ("correction", vp)]
