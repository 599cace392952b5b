"""Per-kernel parity on injected state: each cs3d_k_* entry point (through the
C ABI) against the oracle's restatement of the same reference kernel on the
same inputs.  Bar: bit for bit (the kernels keep the oracle's operation order;
the per-element error and the fraction of identical values are printed)."""
import numpy as np
import pytest

from conftest import elemerr

pytestmark = pytest.mark.gpu

S = 192


def _state(n=27, seed=4, angles=1):
    from cellsim3d_amd import configs
    m, X, Y, Z, info = configs.lattice_tissue(n, 0.40, seed=seed)
    p = configs.base_params(seed)
    p.constrain_angles = angles
    p.division_vol = 1.5 * info["v0"]
    p.voxel_capacity = 64
    rng = np.random.default_rng(seed)
    V = [np.zeros(n * S, np.float32) for _ in range(3)]
    FC = [np.zeros(n * S, np.float32) for _ in range(3)]
    FD = [np.zeros(n * S, np.float32) for _ in range(3)]
    for A in V:
        A.reshape(n, S)[:, :180] = rng.normal(0, 0.3, (n, 180)).astype(np.float32)
    for A in FC + FD:
        A.reshape(n, S)[:, :180] = rng.normal(0, 40, (n, 180)).astype(np.float32)
    P = rng.uniform(50, 60, n).astype(np.float32)
    return p, X, Y, Z, V, FC, FD, P, n


def _system(oracle_mesh, p, X, Y, Z, n, V=None):
    from cellsim3d_amd import engine
    ctx = engine.Context(oracle_mesh, capacity_cells=n + 8, voxel_budget=1 << 22)
    s = ctx.add_system(p, X, Y, Z, n, max_cells=n, voxels=1 << 14, V=V)
    return ctx, s


def _check(name, a, b):
    err = elemerr(a, b)
    print("%s: per-element rel err %.3g, bit-exact fraction %.6f" % (name, err, float(np.mean(a == b))))
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b), "%s is not bit-identical (per-element err %g)" % (name, err)


@pytest.mark.parametrize("angles", [0, 1])
def test_bonded_force(oracle_mesh, angles):
    import oracle_kernels as K
    from cellsim3d_amd import abi
    p, X, Y, Z, V, FC, FD, P, n = _state(angles=angles)
    print("bonded kernel, angles=%d, %d cells" % (angles, n), flush=True)
    ctx, s = _system(oracle_mesh, p, X, Y, Z, n)
    s.set(abi.F_PRESSURE, P)
    CM = [s.get(f)[:n] for f in (abi.F_CMX, abi.F_CMY, abi.F_CMZ)]
    s.k_con_force_bonded()
    _, F = K.con_force(n, p, oracle_mesh, X, Y, Z, CM=CM, P=P, bonded_only=True)
    for a, f in enumerate((abi.F_FCX, abi.F_FCY, abi.F_FCZ)):
        got = s.get(f)[:n * S].reshape(n, S)[:, :180]
        _check("FC[%d]" % a, got, F[a].reshape(n, S)[:, :180])
    ctx.close()


def test_integrate_velocity_com_volumes_pressure(oracle_mesh):
    import oracle_kernels as K
    from cellsim3d_amd import abi
    p, X, Y, Z, V, FC, FD, P, n = _state()
    print("integrate/velocity/com/volumes/pressure, %d cells" % n, flush=True)
    ctx, s = _system(oracle_mesh, p, X, Y, Z, n, V=V)
    for f, A in zip((abi.F_FCX, abi.F_FCY, abi.F_FCZ, abi.F_FDX, abi.F_FDY, abi.F_FDZ), FC + FD):
        s.set(f, A)
    s.set(abi.F_PRESSURE, P)
    s.k_integrate()
    (X1, Y1, Z1), V1 = K.integrate(n, p, X, Y, Z, V, FC, FD)
    for f, ref in zip((abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX, abi.F_VY, abi.F_VZ), (X1, Y1, Z1) + tuple(V1)):
        _check("integrate field %d" % f, s.get(f)[:n * S], ref)
    s.k_velocity_ab()
    V2 = K.velocity_ab(n, p, V1, FC, FD)
    for f, ref in zip((abi.F_VX, abi.F_VY, abi.F_VZ), V2):
        _check("velocity_ab field %d" % f, s.get(f)[:n * S], ref)
    s.k_com()
    cm = K.com(n, X1, Y1, Z1)
    for f, ref in zip((abi.F_CMX, abi.F_CMY, abi.F_CMZ), cm):
        _check("com field %d" % f, s.get(f)[:n], ref)
    s.k_volumes(1)
    faces = np.ctypeslib.as_array(oracle_mesh.faces)
    st, vol, area, flag = K.volumes(n, p, faces, X1, Y1, Z1, CM=cm)
    _check("volume", s.get(abi.F_VOLUME)[:n], vol)
    _check("area", s.get(abi.F_AREA)[:n], area)
    assert np.array_equal(s.get(abi.F_DIVFLAG)[:n], flag)
    s.k_pressure_update(5)
    E = np.full(n, p.stiffness1, np.float32)
    from oracle import oracle as O
    import ctypes as C
    Pn = K.pressure_update(n, p, P, E, 5, np.float32(O.lib().or_rgrowth(C.byref(p), n)))
    _check("pressure", s.get(abi.F_PRESSURE)[:n], Pn)
    ctx.close()


def test_split_velocity_boxes_extent_pressure_reset_state(oracle_mesh):
    """VelocityUpdateA and VelocityUpdateB as entry points of their own
    (propagate.cu:714-736), bounding_boxes (propagatebound.cu:10-141),
    minmaxpre/post (:145-204, 266-323), PressureReset (PressureKernels.cu:45-65)
    and the bulk state getter, on injected state."""
    import oracle_kernels as K
    from cellsim3d_amd import abi, engine
    p, X, Y, Z, V, FC, FD, P, n = _state(seed=6)
    print("split entry points, %d cells" % n, flush=True)
    ctx, s = _system(oracle_mesh, p, X, Y, Z, n, V=V)
    for f, A in zip((abi.F_FCX, abi.F_FCY, abi.F_FCZ, abi.F_FDX, abi.F_FDY, abi.F_FDZ), FC + FD):
        s.set(f, A)
    # A alone, against the formula in the reference's precision: float products, double kick
    s.k_velocity_a()
    dt, m = np.float32(p.dt), np.float32(p.mass)
    VA = [(np.float64(V[a]) + 0.5 * np.float64(dt * FC[a]) / np.float64(m)).astype(np.float32) for a in range(3)]
    for a, f in enumerate((abi.F_VX, abi.F_VY, abi.F_VZ)):
        got = s.get(f)[:n * S]
        assert np.array_equal(got, VA[a]), "VelocityUpdateA component %d" % a
    # then B: A followed by B is the oracle's VelocityUpdateA/B pair
    s.k_velocity_b()
    VAB = K.velocity_ab(n, p, V, FC, FD)
    for a, f in enumerate((abi.F_VX, abi.F_VY, abi.F_VZ)):
        _check("VelocityUpdateA then B, component %d" % a, s.get(f)[:n * S], VAB[a])
    # boxes of the current positions, then their extent
    X1 = X.copy()
    X1.reshape(n, S)[:, :180] += np.float32(0.125)
    s.set(abi.F_X, X1)
    s.k_bboxes()
    bb_ref = K.bboxes(n, X1, Y, Z)
    assert np.array_equal(s.get(abi.F_BBOX)[:6 * n], bb_ref)
    assert np.array_equal(s.k_extent(), K.extent(n, bb_ref))
    # the grid built after the boxes matches the oracle's grid of the same boxes
    s.k_grid()
    _, g_ref, _, _ = K.grid(n, p, bb_ref, budget=1 << 14)
    assert np.array_equal(s.get(abi.F_GRID)[:6], g_ref[:6])
    # PressureReset of three cells; the others keep their pressure
    s.set(abi.F_PRESSURE, P)
    cells = [0, 5, n - 1]
    s.k_pressure_reset(cells, 12.5)
    want = P.copy()
    want[cells] = np.float32(12.5)
    assert np.array_equal(s.get(abi.F_PRESSURE)[:n], want)
    with pytest.raises(engine.EngineError):
        s.k_pressure_reset([n], 1.0)
    # bulk state = the fields one by one
    st = s.get_state()
    assert st["n"] == n and st["status"] == s.info()[2]
    for a, f in enumerate((abi.F_X, abi.F_Y, abi.F_Z)):
        assert np.array_equal(st["xyz"][a], s.get(f)[:n * S])
    for a, f in enumerate((abi.F_VX, abi.F_VY, abi.F_VZ)):
        assert np.array_equal(st["v"][a], s.get(f)[:n * S])
    for a, f in enumerate((abi.F_CMX, abi.F_CMY, abi.F_CMZ)):
        assert np.array_equal(st["CM"][a], s.get(f)[:n])
    assert np.array_equal(st["P"], want) and np.array_equal(st["V"], s.get(abi.F_VOLUME)[:n])
    ctx.close()


def test_refused_calls_leave_the_system_usable(oracle_mesh):
    """Invalid calls through the C ABI return a status and a message and change
    nothing: unknown system, more bytes than a field holds, the read-only system
    centre, inconsistent cell counts, an exhausted cell pool, a NULL output and a
    cell index past the live cells.  The system then steps as before, and its
    read-only candidate counts bound its partner counts."""
    import ctypes as C
    from cellsim3d_amd import abi, engine
    L = engine.lib()
    vp = C.c_void_p
    p, X, Y, Z, V, FC, FD, P, n = _state(seed=8)
    print("refused calls, %d cells" % n, flush=True)
    ctx, s = _system(oracle_mesh, p, X, Y, Z, n)
    x0 = s.get(abi.F_X).copy()
    assert L.cs3d_step(ctx.h, 57, 1) == abi.INVALID_ARG
    big = np.zeros(n * S * 2, np.float32)
    assert L.cs3d_get(ctx.h, s.id, abi.F_X, big.ctypes.data_as(vp), big.nbytes) == abi.INVALID_ARG
    assert L.cs3d_set(ctx.h, s.id, abi.F_X, big.ctypes.data_as(vp), big.nbytes) == abi.INVALID_ARG
    c3 = np.zeros(3, np.float32)
    assert L.cs3d_set(ctx.h, s.id, abi.F_SYSCM, c3.ctypes.data_as(vp), 12) == abi.INVALID_ARG
    assert b"read-only" in L.cs3d_last_error()
    out = C.c_int32()
    assert L.cs3d_system_add_raw(ctx.h, C.byref(p), 10, 5, 0, C.byref(out)) == abi.INVALID_ARG
    assert L.cs3d_system_add_raw(ctx.h, C.byref(p), 1, 100, 0, C.byref(out)) == abi.CAPACITY
    assert L.cs3d_k_extent(ctx.h, s.id, None) == abi.INVALID_ARG
    idx = np.array([n], np.int32)
    assert L.cs3d_k_pressure_reset(ctx.h, s.id, idx.ctypes.data_as(vp), 1, C.c_float(1.0)) == abi.INVALID_ARG
    assert np.array_equal(s.get(abi.F_X), x0)
    s.step(3)
    ctx.sync()
    n1, st1, status = s.info()
    assert (n1, st1) == (n, 3) and status & abi.SYS_FATAL_MASK == 0
    # the candidate counts are read-only and bound the partner counts node by node
    cc = s.get(abi.F_CANDCOUNT)[:n * S].reshape(n, S)[:, :180]
    pc = s.get(abi.F_PAIRCOUNT)[:n * S].reshape(n, S)[:, :180]
    assert (pc <= cc).all() and cc.max() <= 40 and cc.sum() > 0
    assert L.cs3d_set(ctx.h, s.id, abi.F_CANDCOUNT, cc.ctypes.data_as(vp), 4) == abi.INVALID_ARG
    ctx.close()
