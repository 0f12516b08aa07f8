"""Solver configuration (mirror of mvdr_solver_default_config in C).

The reference keeps these as compile-time macros and the values of the
absent `Constants` library (base/Constants.h:30-62); the defaults below are
the reference build's switches (ConfigParser.h:43-53: density on, foreground
off, cross view on; USE_MATERIALIZED_JTJ) and the benchmark schedule (5 GN x
10 PCG, one global iteration in three).
"""
from . import _abi

PRECISION = {"fp32": 0, "mixed": 1}
JTJ_MODE = {"materialized": 0, "matrix_free": 1, "recompute": 2}


def default_config(**overrides):
    c = _abi.SolverConfig()
    c.max_views = 3
    c.max_width = 640
    c.max_height = 480
    c.max_nodes = 8192
    c.max_graph_edges = 8 * 8192
    c.max_feature_pairs = 3 * 1024
    c.max_cross_pairs = 4096
    c.gn_iterations = 5
    c.pcg_iterations = 10
    c.global_iterations = 1
    c.use_elastic_penalty = 1
    c.use_density = 1
    c.use_foreground = 0
    c.use_cross_view = 1
    c.density_in_jtj = 0
    c.precision = 0
    c.jtj_mode = 0
    c.refine_steps = 0
    c.use_graph = 0
    for k, v in overrides.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISION[v]
        if k == "jtj_mode" and isinstance(v, str):
            v = JTJ_MODE[v]
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, int(v))
    return c


def workload_config(workload, **overrides):
    """Schedule of the named benchmark workloads (iteration counts per workload)."""
    base = {0: dict(gn_iterations=3, pcg_iterations=10, max_width=320, max_height=240, max_views=1),
            1: dict(gn_iterations=5, pcg_iterations=10),
            2: dict(gn_iterations=5, pcg_iterations=10, precision=1, refine_steps=1),
            3: dict(gn_iterations=5, pcg_iterations=15, jtj_mode=2),
            4: dict(gn_iterations=5, pcg_iterations=10, precision=1, refine_steps=1, max_nodes=12288,
                    max_graph_edges=8 * 12288)}[workload]
    base.update(overrides)
    return default_config(**base)
