"""C-ABI boundary: the libraries load, export every declared entry point, and
the ctypes mirrors have the C compiler's layout.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from multiviewdynrecon_amd import _abi, config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", "mvdr", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mvdr_[a-z0-9_]+)\s*\(", text)))


def test_solver_exports_every_declared_symbol(built_libs):
    from multiviewdynrecon_amd.solver import lib

    L = lib()
    names = declared("solver.h")
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_crossview_exports_every_declared_symbol(built_libs):
    from multiviewdynrecon_amd.solver import lib

    L = lib()
    names = declared("crossview.h")
    assert len(names) == 5
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_warpfield_exports_every_declared_symbol(built_libs):
    from multiviewdynrecon_amd.solver import lib

    L = lib()
    names = declared("warpfield.h")
    assert len(names) == 5
    assert not [n for n in names if not hasattr(L, n)]


def test_warpfield_rejects_bad_arguments_without_gpu(built_libs):
    from multiviewdynrecon_amd.warpfield import lib

    p = C.c_void_p(16)
    assert lib().mvdr_warp_skin(p, 10, None, 5, 0.025, p, p, p, p, None) == 1
    assert lib().mvdr_warp_skin(p, 10, p, 70000, 0.025, p, p, p, p, None) == 2
    assert lib().mvdr_warp_build_node_graph(p, 70000, p, None) == 2
    assert lib().mvdr_warp_inverse(p, 0, p, p, p, p, 10, p, p, None) == 1
    nc = _abi.NodeCorrection()
    nc.num_views, nc.width, nc.height, nc.num_nodes = 4, 8, 8, 1
    assert lib().mvdr_warp_correct_nodes(C.byref(nc), None) == 2  # more views than MVDR_MAX_VIEWS
    nc.num_views = 1
    assert lib().mvdr_warp_correct_nodes(C.byref(nc), None) == 1  # node arrays missing
    nc.num_nodes = 0
    assert lib().mvdr_warp_correct_nodes(C.byref(nc), None) == 0  # nothing to correct


def test_render_exports_and_rejects_bad_shapes_without_gpu(built_libs):
    from multiviewdynrecon_amd.render import lib

    assert declared("render.h") == ["mvdr_render_solver_maps"]
    ri, ro = _abi.RenderInput(), _abi.RenderOutput()
    ri.num_views, ri.width, ri.height = 4, 8, 8
    assert lib().mvdr_render_solver_maps(C.byref(ri), C.byref(ro), None) == 2
    ri.num_views = 1
    assert lib().mvdr_render_solver_maps(C.byref(ri), C.byref(ro), None) == 1  # output maps missing


def test_rigid_exports_and_rejects_bad_config_without_gpu(built_libs):
    from multiviewdynrecon_amd.rigid import lib

    L = lib()
    names = declared("rigid.h")
    assert len(names) == 4 and not [n for n in names if not hasattr(L, n)]
    for bad in ((4, 64, 64, 0), (1, 0, 64, 0), (1, 64, 64, 2)):
        h = C.c_void_p()
        assert L.mvdr_rigid_create(C.byref(_abi.RigidConfig(*bad)), C.byref(h)) == 2 and not h.value


def test_crossview_create_rejects_bad_config_without_gpu(built_libs):
    from multiviewdynrecon_amd.crossview import default_config, lib

    for bad in (dict(max_views=4), dict(search_step=0), dict(max_pairs=0), dict(max_distance2=-1.0)):
        cfg = default_config(**bad)
        h = C.c_void_p()
        assert lib().mvdr_crossview_create(C.byref(cfg), C.byref(h)) == 2
        assert not h.value


def test_scene_exports_every_declared_symbol(built_libs):
    from multiviewdynrecon_amd.scene import lib

    L = lib()
    missing = [n for n in declared("scene.h") if not hasattr(L, n)]
    assert not missing, missing


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "mvdr/solver.h"
#include "mvdr/scene.h"
#include "mvdr/crossview.h"
#include "mvdr/render.h"
#include "mvdr/rigid.h"
#include "mvdr/warpfield.h"
#define P(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  P(mvdr_intrinsic) P(mvdr_mat34) P(mvdr_dq) P(mvdr_cross_pair) P(mvdr_solver_config) P(mvdr_launch_params)
  P(mvdr_frame_input) P(mvdr_solver_sizes) P(mvdr_scene_params) P(mvdr_scene_info)
  P(mvdr_rigid_config) P(mvdr_rigid_input) O(mvdr_rigid_input, relative_pose)
  P(mvdr_node_correction) O(mvdr_node_correction, num_nodes) O(mvdr_node_correction, correction)
  P(mvdr_render_input) P(mvdr_render_output) O(mvdr_render_input, time_threshold) O(mvdr_render_output, zbuffer)
  P(mvdr_crossview_config) P(mvdr_crossview_input) O(mvdr_crossview_input, corr_map) O(mvdr_crossview_config, max_distance2)
  O(mvdr_frame_input, intrinsic) O(mvdr_frame_input, world2camera) O(mvdr_frame_input, vertex_confid)
  O(mvdr_frame_input, surfel_knn) O(mvdr_frame_input, node_se3) O(mvdr_frame_input, feature_pairs)
  O(mvdr_frame_input, num_cross_pairs) O(mvdr_frame_input, corr_map) O(mvdr_solver_config, use_graph) O(mvdr_scene_params, feature_jitter_px)
  return 0;
}
"""


def test_ctypes_layout_matches_c_compiler():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write(LAYOUT_C)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    c_vals = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    py = {
        "mvdr_intrinsic": C.sizeof(_abi.Intrinsic), "mvdr_mat34": C.sizeof(_abi.Mat34), "mvdr_dq": C.sizeof(_abi.DQ),
        "mvdr_cross_pair": C.sizeof(_abi.CrossPair), "mvdr_solver_config": C.sizeof(_abi.SolverConfig),
        "mvdr_frame_input": C.sizeof(_abi.FrameInput), "mvdr_solver_sizes": C.sizeof(_abi.SolverSizes),
        "mvdr_launch_params": C.sizeof(_abi.LaunchParams),
        "mvdr_scene_params": C.sizeof(_abi.SceneParams), "mvdr_scene_info": C.sizeof(_abi.SceneInfo),
        "mvdr_crossview_config": C.sizeof(_abi.CrossViewConfig), "mvdr_crossview_input": C.sizeof(_abi.CrossViewInput),
        "mvdr_render_input": C.sizeof(_abi.RenderInput), "mvdr_render_output": C.sizeof(_abi.RenderOutput),
        "mvdr_rigid_config": C.sizeof(_abi.RigidConfig), "mvdr_rigid_input": C.sizeof(_abi.RigidInput),
        "mvdr_node_correction": C.sizeof(_abi.NodeCorrection),
    }
    for k, v in py.items():
        assert int(c_vals[k]) == v, k
    offs = {
        "mvdr_frame_input.intrinsic": _abi.FrameInput.intrinsic.offset,
        "mvdr_frame_input.world2camera": _abi.FrameInput.world2camera.offset,
        "mvdr_frame_input.vertex_confid": _abi.FrameInput.vertex_confid.offset,
        "mvdr_frame_input.surfel_knn": _abi.FrameInput.surfel_knn.offset,
        "mvdr_frame_input.node_se3": _abi.FrameInput.node_se3.offset,
        "mvdr_frame_input.feature_pairs": _abi.FrameInput.feature_pairs.offset,
        "mvdr_frame_input.num_cross_pairs": _abi.FrameInput.num_cross_pairs.offset,
        "mvdr_frame_input.corr_map": _abi.FrameInput.corr_map.offset,
        "mvdr_solver_config.use_graph": _abi.SolverConfig.use_graph.offset,
        "mvdr_scene_params.feature_jitter_px": _abi.SceneParams.feature_jitter_px.offset,
        "mvdr_crossview_input.corr_map": _abi.CrossViewInput.corr_map.offset,
        "mvdr_crossview_config.max_distance2": _abi.CrossViewConfig.max_distance2.offset,
        "mvdr_render_input.time_threshold": _abi.RenderInput.time_threshold.offset,
        "mvdr_render_output.zbuffer": _abi.RenderOutput.zbuffer.offset,
        "mvdr_rigid_input.relative_pose": _abi.RigidInput.relative_pose.offset,
        "mvdr_node_correction.num_nodes": _abi.NodeCorrection.num_nodes.offset,
        "mvdr_node_correction.correction": _abi.NodeCorrection.correction.offset,
    }
    for k, v in offs.items():
        assert int(c_vals[k]) == v, k


def test_default_config_matches_c(built_libs):
    from multiviewdynrecon_amd.solver import lib

    c = _abi.SolverConfig()
    lib().mvdr_solver_default_config(C.byref(c))
    d = config.default_config()
    for name, _ in _abi.SolverConfig._fields_:
        assert getattr(c, name) == getattr(d, name), name


def test_create_rejects_bad_config_without_gpu(built_libs):
    # shape validation happens before any device allocation
    from multiviewdynrecon_amd.solver import lib

    for bad in (dict(max_views=0), dict(max_views=4), dict(max_nodes=70000), dict(precision=5),
                dict(pcg_iterations=65)):
        cfg = config.default_config(**bad)
        h = C.c_void_p()
        assert lib().mvdr_solver_create(C.byref(cfg), C.byref(h)) == 2  # MVDR_E_SHAPE
        assert not h.value
    assert lib().mvdr_solver_create(None, None) == 1  # MVDR_E_INVALID


def test_noalloc_rejects_oversized_config_without_gpu(built_libs):
    # the scratch-size check precedes any device work
    from multiviewdynrecon_amd.frame import HostFrame
    from multiviewdynrecon_amd.solver import lib
    from tiny_scene import make_tiny_scene

    frame = HostFrame(make_tiny_scene(1, views=1))
    cfg = config.workload_config(1, precision=1, refine_steps=1)  # 3 x 640 x 480 x 8192 nodes: too large
    assert lib().mvdr_solve_noalloc(C.byref(cfg), C.byref(frame.struct), C.c_void_p(8), None, None) == 3
    assert lib().mvdr_solve_noalloc(None, C.byref(frame.struct), C.c_void_p(8), None, None) == 1
    bad = config.default_config(pcg_iterations=65)
    assert lib().mvdr_solve_noalloc(C.byref(bad), C.byref(frame.struct), C.c_void_p(8), None, None) == 2


def test_status_strings(built_libs):
    from multiviewdynrecon_amd.solver import lib

    for code, name in _abi.STATUS.items():
        assert lib().mvdr_status_string(code).decode() == name


def test_product_path_does_not_import_oracle():
    # only the build recipe names the oracle library (it compiles the checker)
    pkg = os.path.join(ROOT, "multiviewdynrecon_amd")
    pattern = re.compile(r"^\s*(from\s+oracle|import\s+oracle)|libmvdr_oracle|mvdr_oracle_", re.M)
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")) and f != "_build.py":
                text = open(os.path.join(dirpath, f)).read()
                assert not pattern.search(text), f
