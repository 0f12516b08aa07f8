"""Pin the oracle's cross-view pair production (rules in include/mvdr/crossview.h;
the reference ships only CrossViewEdgeCorrespondence.h, so these rules are the
library's reconstruction and the reference itself leaves them unpinned):

* planted frames with known lists: the nearest edge vertex in the sampled
  window (step 2 skips odd offsets), the squared-distance threshold, the
  per-target nearest source, a pair found from both ends kept once from the
  lower source pixel, and back-tracing through correspondence-map entries
  within +-1 pixel;
* random tiny frames against a brute-force Python enumeration written here.
"""
import ctypes as C

import numpy as np
import pytest

from multiviewdynrecon_amd import _abi
from multiviewdynrecon_amd.crossview import crossview_input, default_config
from oracle import lib as oracle_lib
from planted import PlantedCrossView, crossview_case_basic, crossview_case_window, rigid34


def oracle_lists(cv, cfg=None, corr_map=True):
    L = oracle_lib()
    L.mvdr_oracle_crossview.argtypes = [C.POINTER(_abi.CrossViewConfig), C.POINTER(_abi.CrossViewInput), C.c_int,
                                        C.c_void_p, C.c_uint32]
    L.mvdr_oracle_crossview.restype = C.c_longlong
    cfg = cfg or default_config()
    ci, keep = crossview_input(cv, corr_map=corr_map)
    out = []
    for lst in range(3):
        n = L.mvdr_oracle_crossview(C.byref(cfg), C.byref(ci), lst, None, 0)
        buf = np.zeros((max(n, 1), 6), np.uint16)
        L.mvdr_oracle_crossview(C.byref(cfg), C.byref(ci), lst, buf.ctypes.data, n)
        out.append([tuple(int(v) for v in r) for r in buf[:n]])
    return out


def test_planted_nearest_and_unique(built_libs):
    cv, cands, unique = crossview_case_basic()
    got = oracle_lists(cv)
    assert got[0] == cands and got[1] == unique and got[2] == []


def test_planted_window_and_threshold(built_libs):
    cv, cands, unique = crossview_case_window()
    got = oracle_lists(cv)
    assert got[0] == cands and got[1] == unique


@pytest.mark.parametrize("flows,keep", [([(0, 10, 10), (1, 10, 12)], True),      # both pixels traced
                                        ([(0, 11, 9), (1, 9, 13)], True),        # within +-1
                                        ([(0, 10, 10)], False),                  # only one end traced
                                        ([(0, 12, 10), (1, 10, 12)], False)])    # 2 pixels away
def test_planted_backtrace(built_libs, flows, keep):
    cv, _, unique = crossview_case_basic()
    for v, x, y in flows:
        cv.flow(v, x, y)
    got = oracle_lists(cv)
    assert got[2] == (unique if keep else [])
    assert oracle_lists(cv, corr_map=False)[2] == []


def test_single_view_has_no_pairs(built_libs):
    cv = PlantedCrossView(views=1)
    cv.edge(0, 5, 5, cv.ray(5, 5))
    assert oracle_lists(cv) == [[], [], []]


# ------------------------------------------------------------- brute force
def f32(x):
    return np.float32(x)


def rot34_f32(m, p):
    """The rotation of an mvdr_mat34 applied in FP32, each row summed left to right."""
    c = [m[0:3], m[3:6], m[6:9]]
    return [f32(f32(f32(c[0][r] * p[0]) + f32(c[1][r] * p[1])) + f32(c[2][r] * p[2])) for r in range(3)]


def apply34_f32(m, p):
    r = rot34_f32(m, p)
    return [f32(r[k] + m[9 + k]) for k in range(3)]


def inverse34_f32(m):
    """R^T, -R^T t (column-major storage: column c of R^T is row c of R)."""
    R = [[m[3 * col + r] for col in range(3)] for r in range(3)]  # R[r][col]
    inv = np.array([R[c][r] for c in range(3) for r in range(3)] + [0, 0, 0], np.float32)
    rt = rot34_f32(inv, m[9:12])
    inv[9:12] = [-rt[0], -rt[1], -rt[2]]
    return inv


def project(q, f, c):
    u = np.float32(float(q[0]) / (float(q[2]) + 1e-10) * float(f) + float(c))
    return int(np.rint(u))


def brute_force(cv, cfg):
    V, W, H = cv.num_views, cv.width, cv.height
    R, S, T, thr = cfg.search_radius, cfg.search_step, cfg.track_radius, np.float32(cfg.max_distance2)

    def edge(v, x, y):
        if not (0 <= x < W and 0 <= y < H):
            return None
        m = cv.views[v].maps
        if m["edge_mask"][y, x] != 1:
            return None
        p = m["vertex_confid"][y, x]
        return None if (abs(p[0]) < 1e-3 and abs(p[1]) < 1e-3 and abs(p[2]) < 1e-3) else p

    cand = {}
    for v in range(V if V > 1 else 0):
        init = cv.views[v].initial_camera_se3.astype(np.float32)
        for y in range(H):
            for x in range(W):
                src = edge(v, x, y)
                if src is None:
                    continue
                world = apply34_f32(init, src)
                best, found = np.float32(np.inf), None
                for n in dict.fromkeys([(v + 1) % V, (v + V - 1) % V]):
                    q = apply34_f32(inverse34_f32(cv.views[n].initial_camera_se3.astype(np.float32)), world)
                    if not q[2] > 0:
                        continue
                    K = cv.views[n].intrinsic
                    u, w = project(q, K[0], K[2]), project([q[1], q[0], q[2]], K[1], K[3])
                    for yy in range(w - R, w + R + 1, S):
                        for xx in range(u - R, u + R + 1, S):
                            t = edge(n, xx, yy)
                            if t is None:
                                continue
                            d = [f32(t[k] - q[k]) for k in range(3)]
                            d2 = f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))
                            if d2 < best:
                                best, found = d2, (xx, yy, n)
                if found is not None and best < thr:
                    cand[(v, y, x)] = (found, best)
    order = sorted(cand)
    lin = lambda v, y, x: (v * H + y) * W + x
    winner = {}
    for s in order:
        (x2, y2, n), d2 = cand[s]
        key = (n, y2, x2)
        if key not in winner or d2 < cand[winner[key]][1]:
            winner[key] = s
    cands, unique = [], []
    for s in order:
        (x2, y2, n), d2 = cand[s]
        v, y, x = s
        rec = (x, y, x2, y2, v, n)
        cands.append(rec)
        if winner[(n, y2, x2)] != s:
            continue
        back = cand.get((n, y2, x2))
        if back is not None and back[0] == (x, y, v) and winner.get((v, y, x)) == (n, y2, x2) and \
                lin(n, y2, x2) < lin(v, y, x):
            continue
        unique.append(rec)
    return cands, unique


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_frames_against_brute_force(built_libs, seed):
    rng = np.random.default_rng(seed)
    views = 3 if seed else 2
    cv = PlantedCrossView(views=views, W=24, H=20, f=200.0)
    for v in range(views):
        cv.views[v].initial_camera_se3 = rigid34(rng.normal(size=3), rng.uniform(0, 0.02), rng.normal(scale=0.005, size=3))
        for _ in range(150):
            x, y = int(rng.integers(0, 24)), int(rng.integers(0, 20))
            cv.edge(v, x, y, cv.ray(x + rng.normal(scale=0.8), y + rng.normal(scale=0.8), 1.0 + rng.normal(scale=0.003)))
    cfg = default_config(max_distance2=1e-4, search_radius=8)
    cands, unique = brute_force(cv, cfg)
    got = oracle_lists(cv, cfg)
    assert len(cands) > 20 and len(unique) > 10
    assert got[0] == cands and got[1] == unique
