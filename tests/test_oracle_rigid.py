"""Pin the oracle's rigid point-to-plane ICP (rules in include/mvdr/rigid.h):

* one iteration's normal equations against a vectorised numpy evaluation of
  the same rows (A = [p x n, n], b = -n . (p - q) over the pixels passing the
  reference's tests), with and without the reference's every-32nd-block
  sampling;
* convergence: conversion maps made from the reference maps by a known rigid
  motion T^-1 are aligned back, the solved pose equals T.
"""
import ctypes as C

import numpy as np
from scipy.spatial.transform import Rotation

from multiviewdynrecon_amd import _abi
from multiviewdynrecon_amd.rigid import mat34_array, rigid_input
from oracle import lib as oracle_lib
from planted import inverse34, rigid34

W, H, F = 80, 60, 70.0
IDENT = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)


def surface_maps():
    """Vertex / normal maps of a smooth curved surface seen by the camera."""
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    x, y = (u - W / 2) / F, (v - H / 2) / F
    z = 1.0 + 0.15 * np.sin(3 * x) * np.cos(2 * y) + 0.1 * x * x
    P = np.stack([x * z, y * z, z], -1)
    du = np.gradient(P, axis=1)
    dv = np.gradient(P, axis=0)
    n = np.cross(du, dv)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    n[n[..., 2] > 0] *= -1  # facing the camera
    vert = np.concatenate([P, np.ones((H, W, 1))], -1).astype(np.float32)
    norm = np.concatenate([n, np.zeros((H, W, 1))], -1).astype(np.float32)
    return vert, norm


def transform(m, vert, norm):
    R = np.asarray(m[:9], np.float64).reshape(3, 3).T
    t = np.asarray(m[9:12], np.float64)
    v2 = vert.copy()
    n2 = norm.copy()
    v2[..., :3] = vert[..., :3] @ R.T + t
    n2[..., :3] = norm[..., :3] @ R.T
    return v2.astype(np.float32), n2.astype(np.float32)


def oracle_rigid(views, iterations, reference_reduction=0):
    L = oracle_lib()
    L.mvdr_oracle_rigid.argtypes = [C.POINTER(_abi.RigidInput), C.c_int, C.c_uint32, C.POINTER(_abi.Mat34), C.c_void_p]
    keep = []

    def ptr(a):
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data

    ri = rigid_input(views, W, H, ptr)
    poses = (_abi.Mat34 * 3)()
    for v in range(len(views)):
        poses[v] = ri.world2camera[v]
    system = np.zeros(27)
    L.mvdr_oracle_rigid(C.byref(ri), reference_reduction, iterations, poses, system.ctypes.data)
    return np.stack([mat34_array(poses[v]) for v in range(len(views))]), system


def numpy_system(view, reference_reduction):
    cv, cn = transform(view["world2camera"], view["conversion_vertex"], view["conversion_normal"])
    p, pn = cv[..., :3].reshape(-1, 3).astype(np.float32), cn[..., :3].reshape(-1, 3)
    fx, fy, cx, cy = view["intrinsic"]
    u = (p[:, 0].astype(np.float64) / (p[:, 2] + 1e-10) * fx + cx).astype(np.float32)
    w = (p[:, 1].astype(np.float64) / (p[:, 2] + 1e-10) * fy + cy).astype(np.float32)
    x = np.where(u > 0, np.rint(u), 0).astype(np.int64) & 0xFFFF
    y = np.where(w > 0, np.rint(w), 0).astype(np.int64) & 0xFFFF
    ok = (x < W) & (y < H)
    if reference_reduction:
        ok &= (np.arange(W * H) // 256) % 32 == 0
    q = view["reference_vertex"].reshape(-1, 4)[np.where(ok, y * W + x, 0)]
    qn = view["reference_normal"].reshape(-1, 4)[np.where(ok, y * W + x, 0)]
    zero = (np.abs(q[:, :3]) < 1e-3).all(1)
    d = p - q[:, :3]
    ok &= ~zero & ((pn * qn[:, :3]).sum(1) >= 0.8) & ((d * d).sum(1) <= np.float32(0.03) * np.float32(0.03))
    A = np.concatenate([np.cross(p, qn[:, :3]), qn[:, :3]], 1).astype(np.float64)[ok]
    b = -(qn[:, :3] * d).sum(1).astype(np.float64)[ok]
    ata = A.T @ A
    s = [ata[i, j] for i in range(6) for j in range(i, 6)] + list(A.T @ b)
    return np.array(s) * (32.0 if reference_reduction else 1.0), ok.sum()


def view_with_motion(T):
    """Reference maps of the surface; conversion maps = T^-1 applied to them, so
    the pose that aligns them back is T."""
    vert, norm = surface_maps()
    cv, cn = transform(inverse34(T), vert, norm)
    return dict(intrinsic=(F, F, W / 2, H / 2), reference_vertex=vert, reference_normal=norm, conversion_vertex=cv,
                conversion_normal=cn, world2camera=IDENT)


def test_normal_equations_match_numpy(built_libs):
    T = rigid34([1, 2, 3], 0.02, [0.004, -0.003, 0.002])
    view = view_with_motion(T)
    for rr in (0, 1):
        _, system = oracle_rigid([view], 1, rr)
        want, rows = numpy_system(view, rr)
        assert rows > (20 if rr else 1000)
        np.testing.assert_allclose(system, want, rtol=1e-9, atol=1e-12)


def test_icp_recovers_a_known_motion(built_libs):
    T = rigid34([1, -2, 0.5], 0.03, [0.006, 0.004, -0.005])
    view = view_with_motion(T)
    poses, _ = oracle_rigid([view], 12)
    got = poses[0]
    Rg, Rt = got[:9].reshape(3, 3).T, np.asarray(T[:9]).reshape(3, 3).T
    ang = np.linalg.norm(Rotation.from_matrix(Rg @ Rt.T).as_rotvec())
    assert ang < 2e-4, ang
    np.testing.assert_allclose(got[9:12], T[9:12], atol=2e-5)
