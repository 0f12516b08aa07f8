"""Rigid ICP on the GPU (mvdr_rigid_*, through the C ABI) against the oracle:
per-view poses after 1 and 8 iterations within 2e-6 (FP64 sums in a
different order, FP32 poses), with and without the reference's sampled
reduction; the known motion is recovered; two views of the same surface
give the same pose and their average equals it."""
import numpy as np
import pytest

from multiviewdynrecon_amd.rigid import RigidSolver
from planted import rigid34
from test_oracle_rigid import H, W, oracle_rigid, view_with_motion

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("reference_reduction", [0, 1])
@pytest.mark.parametrize("iterations", [1, 8])
def test_rigid_against_oracle(reference_reduction, iterations):
    T = rigid34([1, -2, 0.5], 0.03, [0.006, 0.004, -0.005])
    views = [view_with_motion(T), view_with_motion(rigid34([0, 1, 0], 0.01, [0.002, 0, 0.001]))]
    want, _ = oracle_rigid(views, iterations, reference_reduction)
    solver = RigidSolver(max_views=3, max_width=W, max_height=H, reference_reduction=reference_reduction)
    solver.SolveRigidAlignment(views, W, H, iterations)
    got, avg = solver.GetWorld2Camera()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)
    if iterations == 8 and not reference_reduction:
        np.testing.assert_allclose(got[0], T, atol=3e-4)


def test_average_of_identical_views():
    T = rigid34([0.3, 1, -1], 0.02, [0.003, -0.002, 0.004])
    v = view_with_motion(T)
    solver = RigidSolver(max_views=3, max_width=W, max_height=H)
    solver.SolveRigidAlignment([v, v, v], W, H, 6)
    got, avg = solver.GetWorld2Camera()
    np.testing.assert_array_equal(got[0], got[1])
    np.testing.assert_allclose(avg, got[0], atol=1e-6)
