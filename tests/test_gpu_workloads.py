"""The two larger workloads at their real sizes.

* Uneven batch: the 16 seeded scenes of the batched workload (N ~ U[1000, 8000]
  nodes, 1-3 views, mixed precision) solved as ONE batch, in a hipGraph: every
  member is bit-identical to the same scene solved alone, and the two largest
  members -- the largest overall (3 views) and the largest single-view scene --
  have per-GN twists within 1e-4 of the FP64 oracle from the same warp.
* Scaled body, matrix-free (~8k nodes, ~300k surfels, 3 views, 5 GN x 15 PCG),
  with the Jacobians recomputed inside every apply and with stored ones:
  per-GN twists of the matrix-free mixed solve within 1e-4 of the FP64 oracle
  linearised at its warp, and within 1e-4 of the materialized-J^T J solve run
  side by side from the same input (the two warps stay within rounding).
"""
import numpy as np
import pytest

from multiviewdynrecon_amd.config import workload_config
from multiviewdynrecon_amd.frame import DeviceFrame, HostFrame
from multiviewdynrecon_amd.scene import make_scene, make_uneven_scenes, uneven_scene_params
from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver
from oracle import Oracle

pytestmark = pytest.mark.gpu


def twist_errors_vs_oracle(cfg, scene, frame, gn):
    """Stepwise GPU solve; per GN iteration the relative twist error against the
    FP64 oracle linearised at the GPU's warp field."""
    solver = NonRigidSolver(cfg)
    solver.SetSolverInput(frame)
    solver.BuildSolverIndex()
    o = Oracle(cfg, HostFrame(scene), "fp64")
    o.build_index()
    errs = []
    for k in range(gn):
        se3 = solver.solved_node_se3_host().astype(np.float64)
        solver.SolverIteration()
        o.set("node_se3", se3)
        o.set_iteration(k)
        o.eval_terms()
        o.assemble()
        o.pcg(cfg.pcg_iterations)
        ref = o.get("twist").ravel()
        tw = solver.debug_get("twist").ravel().astype(np.float64)
        errs.append(float(np.linalg.norm(tw - ref) / np.linalg.norm(ref)))
    return errs


def test_uneven_batch_members_and_oracle():
    count = 16
    params = uneven_scene_params(count)
    scenes = make_uneven_scenes(count)
    frames = [DeviceFrame(sc) for sc in scenes]
    cfg = workload_config(4, use_graph=1)
    members = []
    for fr in frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        members.append(s)
    batch = NonRigidBatch(members)
    batch.SolveNonRigidAlignment()
    got = [(m.solved_node_se3_host(), m.debug_get("twist")) for m in members]
    single_cfg = workload_config(4, use_graph=0)
    for k, fr in enumerate(frames):
        s = NonRigidSolver(single_cfg)
        s.SetSolverInput(fr)
        s.SolveNonRigidAlignment()
        np.testing.assert_array_equal(got[k][0], s.solved_node_se3_host(), err_msg=f"scene {k}")
        np.testing.assert_array_equal(got[k][1], s.debug_get("twist"), err_msg=f"scene {k}")
    nodes = np.array([p["target_nodes"] for p in params])
    views = np.array([p["num_views"] for p in params])
    largest = int(np.argmax(nodes))
    single_view = int(np.argmax(np.where(views == 1, nodes, -1)))
    for k in (largest, single_view):
        errs = twist_errors_vs_oracle(single_cfg, scenes[k], frames[k], single_cfg.gn_iterations)
        print(f"scene {k} ({scenes[k].num_nodes} nodes, {views[k]} views): twist rel err {errs}")
        assert max(errs) < 1e-4, (k, errs)


@pytest.fixture(scope="module")
def scaled_body():
    scene = make_scene(3)
    return scene, DeviceFrame(scene)


@pytest.mark.parametrize("jtj_mode", [2, 1])
def test_scaled_body_matrix_free_against_oracle(scaled_body, jtj_mode):
    # 2: Jacobians recomputed inside every apply (the workload's form); 1: stored J
    scene, frame = scaled_body
    cfg = workload_config(3, precision=1, refine_steps=1, jtj_mode=jtj_mode)
    assert workload_config(3).jtj_mode == 2 and cfg.pcg_iterations == 15 and scene.num_nodes > 7000
    errs = twist_errors_vs_oracle(cfg, scene, frame, cfg.gn_iterations)
    print(f"scaled body, matrix-free ({jtj_mode}) mixed: twist rel err {errs}")
    assert max(errs) < 1e-4, errs


@pytest.mark.parametrize("jtj_mode", [2, 1])
def test_scaled_body_matrix_free_equals_materialized(scaled_body, jtj_mode):
    _, frame = scaled_body
    mf = NonRigidSolver(workload_config(3, precision=1, refine_steps=1, jtj_mode=jtj_mode))
    mat = NonRigidSolver(workload_config(3, precision=1, refine_steps=1, jtj_mode=0))
    for s in (mf, mat):
        s.SetSolverInput(frame)
        s.BuildSolverIndex()
    worst = 0.0
    for k in range(5):
        mf.SolverIteration()
        mat.SolverIteration()
        a = mf.debug_get("twist").ravel().astype(np.float64)
        b = mat.debug_get("twist").ravel().astype(np.float64)
        worst = max(worst, np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"scaled body: matrix-free vs materialized twist rel diff {worst:.3e}")
    assert worst < 1e-4
