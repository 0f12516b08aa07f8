"""The autotuner (mvdr_solver_autotune / mvdr_batch_autotune) through the C
ABI: it picks counts from its candidate sets, and neither the tuned table nor
any forced workgroup count changes a result -- every solve stays bit-identical
to the solve with the built-in rules, in a captured graph (re-captured when the
table changes) and without."""
import numpy as np
import pytest

from multiviewdynrecon_amd import solver as sv
from multiviewdynrecon_amd.config import workload_config
from multiviewdynrecon_amd.frame import DeviceFrame
from multiviewdynrecon_amd.scene import make_scene
from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def clean_table():
    sv.clear_launch_table()
    yield
    sv.clear_launch_table()


def solved(members):
    return [(m.solved_node_se3_host().copy(), m.debug_get("twist").copy()) for m in members]


def assert_same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        np.testing.assert_array_equal(x[0], y[0], err_msg=f"{what}: member {k} warp")
        np.testing.assert_array_equal(x[1], y[1], err_msg=f"{what}: member {k} twist")


@pytest.mark.parametrize("precision,use_graph", [(1, 1), (1, 0), (0, 1)])
def test_tuning_changes_no_result(scene1, precision, use_graph):
    scenes = [scene1, make_scene(1, seed=4242)]
    cfg = workload_config(2 if precision else 1, use_graph=use_graph)
    members = []
    for sc in scenes:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(DeviceFrame(sc))
        members.append(s)
    batch = NonRigidBatch(members)
    batch.SolveNonRigidAlignment()
    ref = solved(members)
    chosen = batch.autotune(reps=3)
    assert chosen["eval_workgroups"] in (256, 512, 1024, 2048)
    if precision:
        assert chosen["transform_workgroups"] in (256, 512, 1024, 2048)
        assert chosen["cg_workgroups"] % 8 == 0 and 8 <= chosen["cg_workgroups"] <= 128
    else:
        assert chosen["cg_workgroups"] == 0 and chosen["transform_workgroups"] == 0
    assert batch.lanes in (1, 2)  # the lane count is tuned too
    batch.SolveNonRigidAlignment()
    assert_same(solved(members), ref, "after autotune")
    batch.set_lanes(1)  # the forced shapes below are for launches of both frames
    nodes = max(sc.num_nodes for sc in scenes)
    for cg, ev, tr, gr in ((8, 256, 2048, 1), (32, 2048, 256, 4), (10000, 512, 512, 2)):
        sv.set_launch_params(nodes, len(scenes), cg_workgroups=cg, eval_workgroups=ev, transform_workgroups=tr,
                             gram_rows=gr)
        batch.SolveNonRigidAlignment()
        assert_same(solved(members), ref, f"forced cg {cg} eval {ev} transform {tr} gram rows {gr}")
    # a lone frame: its own shape, the same results as in the batch
    single = NonRigidSolver(cfg)
    single.SetSolverInput(DeviceFrame(scene1))
    single.SolveNonRigidAlignment()
    assert_same(solved([single]), ref[:1], "single vs batch member")
    single.autotune(reps=3)
    single.SolveNonRigidAlignment()
    assert_same(solved([single]), ref[:1], "single after autotune")


@pytest.mark.parametrize("jtj_mode", [1, 2])
def test_gram_rows_per_wave_change_no_result_matrix_free(scene1, jtj_mode):
    """The matrix-free solve assembles only the diagonal blocks; 1, 2 or 4 Gram
    rows per wave give the same solve bit for bit."""
    cfg = workload_config(3 if jtj_mode == 2 else 1, precision=1, jtj_mode=jtj_mode, use_graph=1)
    s = NonRigidSolver(cfg)
    s.SetSolverInput(DeviceFrame(scene1))
    s.SolveNonRigidAlignment()
    ref = solved([s])
    for gr in (2, 4, 1):
        sv.set_launch_params(scene1.num_nodes, 1, gram_rows=gr)
        s.SolveNonRigidAlignment()
        assert_same(solved([s]), ref, f"matrix-free mode {jtj_mode}, gram rows {gr}")
    chosen = s.autotune(reps=2)
    assert chosen["gram_rows"] in (1, 2, 4) and chosen["cg_workgroups"] == 0
    s.SolveNonRigidAlignment()
    assert_same(solved([s]), ref, f"matrix-free mode {jtj_mode} after autotune")
