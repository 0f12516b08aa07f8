"""Matrix-free J^T J apply (applyJtJDotWithIndexKernel, ApplyJtJMatrixFreeIndexed.cu:29-349).

* the apply itself: q = J^T J d read back through the SPMV taps and compared
  with the FP64 oracle's materialized J^T J applied to the same d, built from
  the GPU's own term records -- on the body frame and on tiny frames with
  duplicated knn slots and self edges (the multiplicity correction);
* the same check for the materialized plain-FP32 SpMV;
* GN iterations of the matrix-free mixed solve against the FP64 oracle (twist
  <= 1e-4, diagonal blocks and J^T r <= 1e-5);
* the ~8k-node workload with 15 PCG iterations: size-independent properties.
"""
import ctypes as C

import numpy as np
import pytest

from multiviewdynrecon_amd.config import workload_config
from multiviewdynrecon_amd.frame import DeviceFrame, HostFrame
from multiviewdynrecon_amd.solver import NonRigidSolver, _stream_handle, lib
from oracle import Oracle
from test_gpu_parity import block_rel, gpu_terms_into_oracle
from tiny_scene import make_tiny_scene

pytestmark = pytest.mark.gpu

STAGE_SPMV = 3


def config(workload, precision, jtj_mode, **kw):
    extra = dict(precision=1, refine_steps=1) if precision == "mixed" else dict(precision=0, refine_steps=0)
    extra.update(kw)
    return workload_config(workload, jtj_mode=jtj_mode, **extra)


def spmv_against_oracle(cfg, scene):
    """One GN iteration, then one J^T J apply (stage bench) on its data; returns
    the relative error of q against the oracle's FP64 A d."""
    solver = NonRigidSolver(cfg)
    solver.SetSolverInput(DeviceFrame(scene))
    solver.BuildSolverIndex()
    solver.SolverIteration()
    out = C.c_float()
    solver._check(lib().mvdr_solver_bench_stage(solver.h, STAGE_SPMV, 1, _stream_handle(None), C.byref(out)))
    d = solver.debug_get("spmv_input").reshape(-1).astype(np.float64)
    q = solver.debug_get("spmv_output").reshape(-1).astype(np.float64)
    o = Oracle(cfg, HostFrame(scene), "fp64")
    o.build_index()
    o.set_iteration(0)
    o.eval_terms()
    gpu_terms_into_oracle(solver, o)
    o.assemble()
    ref = o.apply_jtj(d)
    assert np.isfinite(q).all() and np.linalg.norm(d) > 0
    return np.linalg.norm(q - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_matrix_free_apply_body(scene1, precision):
    assert spmv_against_oracle(config(1, precision, 1), scene1) < 1e-5


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_matrix_free_apply_tiny_frames(seed):
    scene = make_tiny_scene(seed)
    cfg = config(1, "mixed", 1, max_width=scene.width, max_height=scene.height, max_nodes=64)
    assert spmv_against_oracle(cfg, scene) < 1e-5


def test_materialized_spmv_body(scene1):
    assert spmv_against_oracle(config(1, "fp32", 0), scene1) < 1e-5


def test_matrix_free_gn_iterations_match_fp64_oracle(scene1):
    cfg = config(1, "mixed", 1)
    solver = NonRigidSolver(cfg)
    solver.SetSolverInput(DeviceFrame(scene1))
    solver.BuildSolverIndex()
    assert solver.sizes()["num_node_pairs"] == 0  # no pair index in the matrix-free mode
    o = Oracle(cfg, HostFrame(scene1), "fp64")
    o.build_index()
    worst = 0.0
    for k in range(3):
        se3 = solver.solved_node_se3_host().astype(np.float64)
        solver.SolverIteration()
        o.set("node_se3", se3)
        o.set_iteration(k)
        o.eval_terms()
        gpu_terms_into_oracle(solver, o)
        o.assemble()
        assert block_rel(solver.debug_get("jtj_diag"), o.get("jtj_diag")).max() < 1e-5
        jtr = o.get("jtr").ravel()
        assert np.linalg.norm(solver.debug_get("jtr").ravel() - jtr) / np.linalg.norm(jtr) < 1e-5
        o.pcg(cfg.pcg_iterations)
        ref = o.get("twist").ravel()
        tw = solver.debug_get("twist").ravel().astype(np.float64)
        worst = max(worst, np.linalg.norm(tw - ref) / np.linalg.norm(ref))
        o.set("twist", tw)
        o.apply_update()
        np.testing.assert_allclose(solver.solved_node_se3_host(), o.get("node_se3"), rtol=0, atol=1e-6)
    assert worst < 1e-4, worst


def test_matrix_free_scaled_body():
    """~8k nodes / ~300k surfels, 5 GN x 15 PCG, matrix-free with recomputed
    Jacobians (plain FP32)."""
    from multiviewdynrecon_amd.scene import make_scene

    scene = make_scene(3)
    cfg = workload_config(3)
    assert cfg.jtj_mode == 2 and cfg.pcg_iterations == 15
    solver = NonRigidSolver(cfg)
    solver.SetSolverInput(DeviceFrame(scene))
    solver.BuildSolverIndex()
    c0 = solver.ComputeTotalResidualSynced()
    for _ in range(cfg.gn_iterations):
        solver.SolverIteration()
    assert np.isfinite(solver.debug_get("twist")).all()
    assert solver.ComputeTotalResidualSynced() < c0
