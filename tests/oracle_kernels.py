"""Thin numpy wrappers over the CPU oracle's per-kernel functions, for the
CPU pin tests (test infrastructure only)."""
import ctypes as C

import numpy as np

from cellsim3d_amd import abi
from oracle import oracle as O

S = abi.SLOTS


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def zeros(n):
    return np.zeros(n * S, dtype=np.float32)


def bboxes(n, X, Y, Z):
    bb = np.zeros(6 * n, dtype=np.float32)
    O.lib().or_bboxes(n, f32(X), f32(Y), f32(Z), bb)
    return bb


def extent(n, bb):
    out = np.zeros(6, dtype=np.float32)
    O.lib().or_extent(n, f32(bb), out)
    return out


def grid(n, p, bb, budget=1 << 14):
    g = np.zeros(8, dtype=np.int32)
    counts = np.zeros(budget, dtype=np.int32)
    lists = np.zeros(budget * p.voxel_capacity, dtype=np.int32)
    st = O.lib().or_grid_build(n, C.byref(p), f32(bb), budget, g, counts, lists)
    return st, g, counts, lists


def com(n, X, Y, Z):
    cm = [np.zeros(n, dtype=np.float32) for _ in range(3)]
    O.lib().or_com(n, f32(X), f32(Y), f32(Z), *cm)
    return cm


def con_force(n, p, mesh, X, Y, Z, CM=None, P=None, stiff=None, bonded_only=False, budget=1 << 14):
    """CalculateConForce over fresh bboxes and grid; returns (status, FX, FY, FZ)."""
    X, Y, Z = f32(X), f32(Y), f32(Z)
    if CM is None:
        CM = com(n, X, Y, Z)
    P = f32(np.full(n, p.min_pressure) if P is None else P)
    stiff = f32(np.full(n, p.stiffness1) if stiff is None else stiff)
    bb = bboxes(n, X, Y, Z)
    st, g, counts, lists = grid(n, p, bb, budget)
    F = [zeros(n) for _ in range(3)]
    st |= O.lib().or_con_force(n, C.byref(p), C.byref(mesh), X, Y, Z, *(f32(c) for c in CM), P, stiff, bb, g,
                               counts, lists, 1 if bonded_only else 0, *F)
    return st, F


def dis_force(n, p, mesh, X, Y, Z, VX, VY, VZ, budget=1 << 14):
    X, Y, Z = f32(X), f32(Y), f32(Z)
    bb = bboxes(n, X, Y, Z)
    st, g, counts, lists = grid(n, p, bb, budget)
    F = [zeros(n) for _ in range(3)]
    st |= O.lib().or_dis_force(n, C.byref(p), C.byref(mesh), X, Y, Z, f32(VX), f32(VY), f32(VZ), bb, g, counts,
                               lists, *F)
    return st, F


def volumes(n, p, faces, X, Y, Z, CM=None, stiff=None):
    X, Y, Z = f32(X), f32(Y), f32(Z)
    if CM is None:
        CM = com(n, X, Y, Z)
    stiff = f32(np.full(n, p.stiffness1) if stiff is None else stiff)
    vol = np.zeros(n, dtype=np.float32)
    area = np.zeros(n, dtype=np.float32)
    flag = np.zeros(n, dtype=np.int8)
    st = O.lib().or_volumes(n, C.byref(p), np.ascontiguousarray(faces, dtype=np.int32), X, Y, Z,
                            *(f32(c) for c in CM), stiff, 0, vol, area, flag)
    return st, vol, area, flag


def integrate(n, p, X, Y, Z, V, FC, FD):
    X, Y, Z = (f32(a).copy() for a in (X, Y, Z))
    V = [f32(a).copy() for a in V]
    O.lib().or_integrate(n, C.byref(p), X, Y, Z, *V, *(f32(a) for a in FC), *(f32(a) for a in FD))
    return (X, Y, Z), V


def velocity_ab(n, p, V, FC, FD):
    V = [f32(a).copy() for a in V]
    O.lib().or_velocity_ab(n, C.byref(p), *V, *(f32(a) for a in FC), *(f32(a) for a in FD))
    return V


def pressure_update(n, p, P, stiff, step, inc, scale=None):
    P = f32(P).copy()
    sc = None if scale is None else f32(scale)
    O.lib().or_pressure_update(n, C.byref(p), P, f32(stiff), None if sc is None else sc.ctypes.data_as(C.c_void_p),
                               step, inc)
    return P
