"""Batched solves (several frames per launch) against solves of one frame.

Kernels index the frame by blockIdx.y and bound their work by that frame's own
sizes, and every per-frame reduction keeps its order, so a member of a batch
must produce bit-identical results to the same frame solved alone, whatever
the other members (different sizes, view counts, node counts).
"""
import numpy as np
import pytest

from multiviewdynrecon_amd.config import workload_config
from multiviewdynrecon_amd.frame import DeviceFrame
from multiviewdynrecon_amd.scene import make_scene
from multiviewdynrecon_amd.solver import NonRigidBatch, NonRigidSolver, SolverError
from tiny_scene import make_tiny_scene

pytestmark = pytest.mark.gpu


def shared_config(precision, **kw):
    # one schedule for every member; capacities cover the largest frame
    extra = dict(precision=1, refine_steps=1) if precision == "mixed" else dict(precision=0, refine_steps=0)
    extra.update(kw)
    return workload_config(1, **extra)


@pytest.fixture(scope="module")
def frames():
    scenes = [make_scene(1, seed=11), make_scene(0, seed=3), make_tiny_scene(5), make_scene(1, seed=12, num_views=2)]
    return [DeviceFrame(s) for s in scenes]


@pytest.mark.parametrize("precision,count", [("mixed", 4), ("fp32", 4), ("mixed", 8)])
def test_batch_members_equal_single_solves(frames, precision, count):
    """count = 8: a frame count divisible by the XCD count, where the CG row
    kernels keep each frame's workgroups on one XCD."""
    cfg = shared_config(precision)
    members_frames = [frames[k % len(frames)] for k in range(count)]
    single = []
    for fr in frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        s.SolveNonRigidAlignment()
        single.append((s.solved_node_se3_host(), s.debug_get("twist"), s.sizes()))
    single = [single[k % len(frames)] for k in range(count)]
    members = []
    for fr in members_frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        members.append(s)
    batch = NonRigidBatch(members)
    for rep in range(2):  # a second solve of the same inputs gives the same result
        batch.SolveNonRigidAlignment()
        for k, s in enumerate(members):
            se3, twist, sizes = single[k]
            np.testing.assert_array_equal(s.solved_node_se3_host(), se3, err_msg=f"member {k} rep {rep}")
            np.testing.assert_array_equal(s.debug_get("twist"), twist, err_msg=f"member {k} rep {rep}")
            assert s.sizes() == sizes


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_lanes_equal_single_solves(frames, graph):
    """Lanes solve runs of members concurrently on their own streams; every
    lane count gives each member its single-solve result bit for bit."""
    cfg = shared_config("mixed", use_graph=graph)
    count = 8
    single = []
    for fr in frames:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        s.SolveNonRigidAlignment()
        single.append((s.solved_node_se3_host(), s.debug_get("twist")))
    members = []
    for k in range(count):
        s = NonRigidSolver(cfg)
        s.SetSolverInput(frames[k % len(frames)])
        members.append(s)
    batch = NonRigidBatch(members)
    assert batch.lanes == 2  # the default from 8 members on
    for lanes in (1, 3, 4, 2):
        batch.set_lanes(lanes)
        assert batch.lanes == lanes
        for rep in range(2):
            batch.SolveNonRigidAlignment()
            for k, s in enumerate(members):
                se3, twist = single[k % len(frames)]
                np.testing.assert_array_equal(s.solved_node_se3_host(), se3, err_msg=f"lanes {lanes} member {k}")
                np.testing.assert_array_equal(s.debug_get("twist"), twist, err_msg=f"lanes {lanes} member {k}")
    for bad in (0, 5, 9):
        with pytest.raises(SolverError):
            batch.set_lanes(bad)
    small = NonRigidBatch(members[:3])
    assert small.lanes == 1
    with pytest.raises(SolverError):
        small.set_lanes(4)


def test_batch_graph_replays_match(frames):
    cfg = shared_config("mixed", use_graph=1)
    members = []
    for fr in frames[:3]:
        s = NonRigidSolver(cfg)
        s.SetSolverInput(fr)
        members.append(s)
    batch = NonRigidBatch(members)
    batch.SolveNonRigidAlignment()
    first = [m.solved_node_se3_host() for m in members]
    for _ in range(3):
        members[0].ComputeTotalResidualSynced()  # eager work between replays
        batch.SolveNonRigidAlignment()
        for m, ref in zip(members, first):
            np.testing.assert_array_equal(m.solved_node_se3_host(), ref)


def test_batch_rejects_mixed_schedules(frames):
    a = NonRigidSolver(shared_config("mixed"))
    b = NonRigidSolver(shared_config("mixed", gn_iterations=3))
    with pytest.raises(SolverError) as e:
        NonRigidBatch([a, b])
    assert e.value.status == 2
    with pytest.raises(SolverError):
        NonRigidBatch([a, a])
