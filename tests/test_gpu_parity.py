"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on the
same seeded, injected state.

Bar: neighbour grids, candidate lists, division events and cell counts are
bit-exact.  The kernels follow the oracle's operation order (no contraction,
correctly rounded float division and square root, the reference's double
promotions), so positions, velocities, forces, pressures, volumes and
centres of mass are compared bit for bit as well; where a test states a
tolerance instead, it is per element: |gpu - oracle| <= tol * max(|oracle|,
1e-6 * max|oracle|) with tol = 1e-5, the force tolerance of this engine.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import elemerr, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _engine():
    from cellsim3d_amd import engine
    return engine


def _oracle():
    from oracle import oracle as O
    return O


def _abi():
    from cellsim3d_amd import abi
    return abi


def _case(kind="tissue", seed=3, n=27, phi=0.40, growth=False, vcap=64):
    from cellsim3d_amd import configs
    if kind == "small":
        cfg = configs.plain_small(seed=seed)
        return cfg["params"], cfg["X"], cfg["Y"], cfg["Z"], cfg["n"], cfg["max_cells"]
    m, X, Y, Z, info = configs.lattice_tissue(n, phi, seed=seed)
    p = configs.base_params(seed)
    p.division_vol = 2.0 * info["v0"]
    p.div_time_steps = 10 ** 6
    p.voxel_capacity = vcap
    if not growth:
        p.growth_rate = 0.0
    return p, X, Y, Z, n, 4 * n


def _pair(oracle_mesh, p, X, Y, Z, n, max_cells, budget=1 << 16, V=None, extra_cells=16):
    E, O, A = _engine(), _oracle(), _abi()
    sim = O.Sim(p, oracle_mesh, X, Y, Z, n, max_cells=max_cells, budget=budget)
    if V is not None:
        for f, a in zip((A.F_VX, A.F_VY, A.F_VZ), V):
            full = np.zeros(max_cells * 192, np.float32)
            full[:len(a)] = a
            sim.set(f, full)
    sim.init()
    ctx = E.Context(oracle_mesh, capacity_cells=max_cells + extra_cells, voxel_budget=budget * 70)
    sys_ = ctx.add_system(p, X, Y, Z, n, max_cells=max_cells, voxels=budget, V=V)
    return sim, ctx, sys_


def _cmp(sim, sys_, fields, n, tol=None, label=""):
    """Compares fields of the first n cells: bit for bit (tol None), or per
    element within tol."""
    A = _abi()
    out = {}
    for f in fields:
        a = sys_.get(f)
        b = sim.get(f)
        if f in A.NODE_FIELDS:
            a = a[:n * 192]
            b = b[:n * 192]
        elif f in A.CELL_FIELDS:
            a = a[:n]
            b = b[:n]
        elif f == A.F_BBOX:
            a = a[:6 * n]
            b = b[:6 * n]
        err = relerr(a, b)
        eerr = elemerr(a, b)
        exact = float(np.mean(a == b))
        out[f] = (err, exact)
        print("%s field %d: scale-relative err %.3g, per-element err %.3g, bit-exact fraction %.6f" %
              (label, f, err, eerr, exact))
        assert np.all(np.isfinite(a)), "non-finite values in field %d" % f
        if tol is None:
            assert np.array_equal(a, b), "field %d is not bit-identical (per-element err %g)" % (f, eerr)
        else:
            assert eerr <= tol, "field %d per-element err %g > %g" % (f, eerr, tol)
    return out


def _cmp_grid(sim, sys_):
    A = _abi()
    g1, g0 = sys_.get(A.F_GRID), sim.get(A.F_GRID)
    assert np.array_equal(g1[:6], g0[:6]), (g1, g0)
    nv = int(g0[3] * g0[4] * g0[5])
    cap = int(g0[7])
    c1, c0 = sys_.get(A.F_VOXCOUNT)[:nv], sim.get(A.F_VOXCOUNT)[:nv]
    assert np.array_equal(c1, c0)
    l1 = sys_.get(A.F_VOXLIST)[:nv * cap].reshape(nv, cap)
    l0 = sim.get(A.F_VOXLIST)[:nv * cap].reshape(nv, cap)
    for v in range(nv):
        k = min(int(c0[v]), cap)
        assert np.array_equal(l1[v, :k], l0[v, :k]), "voxel %d list differs" % v
    return nv


def test_startup_matches_oracle(oracle_mesh):
    A = _abi()
    p, X, Y, Z, n, mc = _case("tissue", n=27)
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc)
    nv = _cmp_grid(sim, s)
    assert nv > 1
    _cmp(sim, s, [A.F_CMX, A.F_CMY, A.F_CMZ, A.F_BBOX], n, label="startup")
    _cmp(sim, s, [A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY, A.F_FDZ], n, label="startup")
    ctx.close()


def test_con_and_dis_force_injected_velocities(oracle_mesh):
    """Random velocities exercise bond damping, friction and drag."""
    A = _abi()
    p, X, Y, Z, n, mc = _case("tissue", n=64, seed=5)
    rng = np.random.default_rng(11)
    V = [np.zeros(n * 192, np.float32) for _ in range(3)]
    for a in V:
        a.reshape(n, 192)[:, :180] = rng.normal(0, 0.5, size=(n, 180)).astype(np.float32)
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc, V=V)
    _cmp_grid(sim, s)
    _cmp(sim, s, [A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY, A.F_FDZ], n, label="injected")
    # per-kernel entry points on the same state
    s.k_con_force()
    s.k_dis_force(0)
    _cmp(sim, s, [A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY, A.F_FDZ], n, label="k_*")
    ctx.close()


@pytest.mark.parametrize("steps", [1, 25])
def test_steps_match_oracle(oracle_mesh, steps):
    A = _abi()
    p, X, Y, Z, n, mc = _case("tissue", n=64, seed=7)
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc)
    sim.step(steps)
    ctx.set_graph(False)
    s.step(steps)
    ctx.sync()
    n1, st1, status = s.info()
    assert n1 == sim.n and st1 == sim.stepno
    assert status & A.SYS_FATAL_MASK == 0
    _cmp_grid(sim, s)
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY,
                  A.F_FDZ, A.F_CMX, A.F_CMY, A.F_CMZ, A.F_PRESSURE], n, label="steps=%d" % steps)
    ctx.close()


def test_division_matches_oracle(oracle_mesh):
    """Growth with a division volume just above V0: cells divide within a few
    hundred steps; event order, daughter indices and RANMAR normals must agree."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=2)
    p = cfg["params"]
    p.division_vol = np.float32(1.02 * cfg["v0"])
    n, mc = cfg["n"], 64
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, mc)
    ctx.set_graph(False)
    total = 0
    for k in range(6):
        sim.step(50)
        s.step(50)
        ctx.sync()
        total += 50
        n1, st1, status = s.info()
        assert status & A.SYS_FATAL_MASK == 0
        assert n1 == sim.n, "cell count differs after %d steps: gpu %d oracle %d" % (total, n1, sim.n)
    assert sim.n > n, "no division happened; test is not exercising division"
    nn = sim.n
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_PRESSURE, A.F_VOLUME], nn, label="division")
    r1, r0 = s.get(A.F_RANMAR), sim.get(A.F_RANMAR)
    assert r1.i97 == r0.i97 and r1.j97 == r0.j97
    assert np.array_equal(np.ctypeslib.as_array(r1.u), np.ctypeslib.as_array(r0.u))
    ctx.close()


def test_batched_equals_alone(oracle_mesh):
    """Two systems stepped in one batch give bit-identical state to each system
    stepped alone (per-node summation order does not depend on the batch)."""
    A, E = _abi(), _engine()
    cases = [_case("tissue", n=27, seed=3), _case("tissue", n=40, seed=4, phi=0.3)]
    cases[1][0].attraction_strength = np.float32(0.4)
    ctx = E.Context(oracle_mesh, capacity_cells=512, voxel_budget=1 << 24)
    ctx_a = E.Context(oracle_mesh, capacity_cells=512, voxel_budget=1 << 24)
    sb = [ctx.add_system(p, X, Y, Z, n, max_cells=mc) for (p, X, Y, Z, n, mc) in cases]
    sa = [ctx_a.add_system(p, X, Y, Z, n, max_cells=mc) for (p, X, Y, Z, n, mc) in cases]
    ctx.step(30)  # batched (graph path)
    for s in sa:
        s.step(30)  # one system per launch
    ctx.sync()
    ctx_a.sync()
    for a, b, cse in zip(sa, sb, cases):
        n = cse[4]
        for f in (A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FDX):
            assert np.array_equal(a.get(f)[:n * 192], b.get(f)[:n * 192]), "field %d differs" % f
    ctx.close()
    ctx_a.close()


def test_graph_equals_eager(oracle_mesh):
    A, E = _abi(), _engine()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=9)
    c1 = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    c2 = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    s1 = c1.add_system(p, X, Y, Z, n, max_cells=mc)
    s2 = c2.add_system(p, X, Y, Z, n, max_cells=mc)
    c2.set_graph(False)
    for _ in range(3):  # replays with eager work and syncs in between
        s1.step(20)
        s2.step(20)
        c1.sync()
        c2.sync()
        for f in (A.F_X, A.F_VX, A.F_FCZ):
            assert np.array_equal(s1.get(f), s2.get(f))
    c1.close()
    c2.close()


def test_timed_search_leaves_results_unchanged(oracle_mesh):
    """The kernel timer's full searches (cs3d_time_kernel, which = 2) fill the search
    bins, search every cell, clear the bins and make the next step search again: a
    run with them between steps ends bit-identical to one without."""
    A, E = _abi(), _engine()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=6)
    p.grid_interval = 10
    c1 = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    c2 = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    s1 = c1.add_system(p, X, Y, Z, n, max_cells=mc)
    s2 = c2.add_system(p, X, Y, Z, n, max_cells=mc)
    for k in (7, 13, 20):
        s1.step(k)
        assert s1.time_kernel(2, reps=3) > 0
        s2.step(k)
    c1.sync()
    c2.sync()
    assert s1.info()[1] == s2.info()[1] == 40
    for f in (A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FDX):
        assert np.array_equal(s1.get(f)[:n * 192], s2.get(f)[:n * 192]), "field %d differs" % f
    c1.close()
    c2.close()


@pytest.mark.parametrize("graph", [False, True])
def test_random_force_matches_oracle(oracle_mesh, graph):
    """add_rands=1: the counter-based normals (Philox4x32-10 + polar method) are
    drawn per (cell, node, step) on the device and in the oracle; the random
    force enters Integrate with the previous step's draw and VelocityUpdateA
    with this step's, so positions and velocities must agree as without noise,
    and the FR fields themselves must be bit-identical."""
    A = _abi()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=5)
    p.add_rands = 1
    p.rand_scale_factor = np.float32(0.05)
    p.noise_seed = 11
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc)
    ctx.set_graph(graph)
    for k in (1, 19):
        sim.step(k)
        s.step(k)
        ctx.sync()
        n1, st1, status = s.info()
        assert n1 == sim.n and st1 == sim.stepno
        assert status & A.SYS_FATAL_MASK == 0
        for f in (A.F_FRX, A.F_FRY, A.F_FRZ):
            a, b = s.get(f)[:n * 192], sim.get(f)[:n * 192]
            assert np.array_equal(a, b), "random force field %d differs at step %d" % (f, st1)
        _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_PRESSURE], n,
             label="add_rands steps=%d" % st1)
    # the noise moved the system: a run without it differs
    p0 = _case("tissue", n=27, seed=5)[0]
    sim0, ctx0, s0 = _pair(oracle_mesh, p0, X, Y, Z, n, mc)
    s0.step(20)
    ctx0.sync()
    assert not np.array_equal(s0.get(A.F_X)[:n * 192], s.get(A.F_X)[:n * 192])
    ctx0.close()
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_correct_com_matches_oracle(oracle_mesh, graph):
    """correct_com = 1 (GPUbounce.cu:1431-1441): the division phase sums the cell
    centres in cell order and a sixth launch shifts every node; positions and the
    system centre must agree with the oracle bit for bit, across a division."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=2)
    p = cfg["params"]
    p.division_vol = np.float32(1.02 * cfg["v0"])
    p.correct_com = 1
    n, mc = cfg["n"], 64
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, mc)
    ctx.set_graph(graph)
    for k in (1, 30, 270):
        sim.step(k)
        s.step(k)
        ctx.sync()
        n1, st1, status = s.info()
        assert n1 == sim.n and st1 == sim.stepno and status & A.SYS_FATAL_MASK == 0
        assert np.array_equal(s.get(A.F_SYSCM), sim.get(A.F_SYSCM)), "system centre at step %d" % st1
        _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_CMX, A.F_PRESSURE], n1,
             label="correct_com steps=%d" % st1)
    assert sim.n > n, "no division: the test does not cover daughters"
    ctx.close()


def test_resume_after_horizon_equals_uninterrupted(oracle_mesh):
    """A system that stops at its horizon on a grid-rebuild step while others in
    the batch go on, then resumes, continues bit-identically to a run without the
    stop (the stopped system keeps its pending grid adoption)."""
    A, E = _abi(), _engine()
    pa, Xa, Ya, Za, na, mca = _case("tissue", n=27, seed=3)
    pb, Xb, Yb, Zb, nb, mcb = _case("tissue", n=40, seed=4, phi=0.3)
    for p_ in (pa, pb):
        p_.grid_interval = 10
    ctx = E.Context(oracle_mesh, capacity_cells=512, voxel_budget=1 << 24)
    sa = ctx.add_system(pa, Xa, Ya, Za, na, max_cells=mca)
    sb = ctx.add_system(pb, Xb, Yb, Zb, nb, max_cells=mcb)
    sa.set_horizon(20)
    ctx.step(35)  # a stops after step 20 (a rebuild step), b goes on to 35
    ctx.sync()
    assert sa.info()[1] == 20 and sb.info()[1] == 35
    sa.set_horizon(-1)
    sa.step(30)
    ctx.sync()
    ref = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    sr = ref.add_system(pa, Xa, Ya, Za, na, max_cells=mca)
    sr.step(50)
    ref.sync()
    assert sa.info()[1] == 50 and sr.info()[1] == 50
    for f in (A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_FCX, A.F_FDX, A.F_BBOX):
        assert np.array_equal(sa.get(f)[:na * 192], sr.get(f)[:na * 192]), "field %d differs after the resume" % f
    ctx.close()
    ref.close()


def test_system_without_cells(oracle_mesh):
    """A system with no cells steps alone and in a batch without any status bit,
    and leaves the other systems' results unchanged."""
    A, E = _abi(), _engine()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=3)
    ctx = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    empty = np.zeros(0, np.float32)
    s0 = ctx.add_system(p, empty, empty, empty, 0, max_cells=8)
    s1 = ctx.add_system(p, X, Y, Z, n, max_cells=mc)
    s0.step(12)
    ctx.step(25)
    ctx.sync()
    n0, st0, status0 = s0.info()
    assert (n0, st0, status0) == (0, 37, 0)
    assert s1.info()[1] == 25 and s1.info()[2] & A.SYS_FATAL_MASK == 0
    ref = E.Context(oracle_mesh, capacity_cells=256, voxel_budget=1 << 22)
    sr = ref.add_system(p, X, Y, Z, n, max_cells=mc)
    sr.step(25)
    ref.sync()
    for f in (A.F_X, A.F_VX, A.F_FCX):
        assert np.array_equal(s1.get(f)[:n * 192], sr.get(f)[:n * 192])
    ctx.close()
    ref.close()


def test_full_size_tissue_matches_oracle(oracle_mesh):
    """The throughput workload at its full size (2000 cells, 360,000 nodes, volume
    fraction 0.40), 30 steps with the grid rebuilt every 2 (the candidate lists are
    re-searched along the way): positions, velocities, forces, centres and the voxel
    grid against the oracle."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.single_tissue(2000, 0.40, seed=1, grid_interval=2, steps=10)
    p, n = cfg["params"], cfg["n"]
    print("full-size tissue: %d cells" % n, flush=True)
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, n, budget=1 << 14)
    sim.step(30)
    s.step(30)
    ctx.sync()
    n1, st1, status = s.info()
    assert n1 == n and st1 == 30 and status & A.SYS_FATAL_MASK == 0
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY,
                  A.F_FDZ, A.F_CMX], n, label="2000 cells")
    _cmp_grid(sim, s)
    ctx.close()


def test_throughput_workload_at_length_matches_oracle(oracle_mesh):
    """The throughput workload as the bench steps it (2000 cells, grid rebuilt every
    10 steps, the captured graph) for 300 steps, through about a dozen candidate
    searches, 30 grid rebuilds and the skin decisions between them: the state and the
    voxel grid bit for bit at the end."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.single_tissue(2000, 0.40, seed=1, grid_interval=10, steps=300)
    p, n = cfg["params"], cfg["n"]
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, n, budget=1 << 14)
    ctx.set_graph_steps(100)
    for k in range(3):
        sim.step(100)
        s.step(100)
        ctx.sync()
        n1, st1, status = s.info()
        assert n1 == n and st1 == 100 * (k + 1) and status & A.SYS_FATAL_MASK == 0
        print("2000 cells: %d steps" % st1, flush=True)
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY,
                  A.F_FDZ, A.F_CMX, A.F_CMY, A.F_CMZ], n, label="2000 cells, 300 steps")
    _cmp_grid(sim, s)
    ctx.close()


def test_random_force_with_divisions_matches_oracle(oracle_mesh):
    """add_rands=1 through divisions: the daughters' first Integrate and the random
    force of the new cells, in the graph path, against the oracle."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=3)
    p = cfg["params"]
    p.division_vol = np.float32(1.02 * cfg["v0"])
    p.add_rands = 1
    p.rand_scale_factor = np.float32(0.05)
    p.noise_seed = 7
    n, mc = cfg["n"], 64
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, mc)
    for k in range(6):
        sim.step(50)
        s.step(50)
        ctx.sync()
        n1, st1, status = s.info()
        assert status & A.SYS_FATAL_MASK == 0
        assert n1 == sim.n and st1 == sim.stepno, "after %d steps: gpu %d cells, oracle %d" % (st1, n1, sim.n)
    assert sim.n > n, "no division happened"
    nn = sim.n
    for f in (A.F_FRX, A.F_FRY, A.F_FRZ):
        assert np.array_equal(s.get(f)[:nn * 192], sim.get(f)[:nn * 192]), "random force field %d" % f
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_PRESSURE, A.F_VOLUME], nn,
         label="add_rands with divisions")
    ctx.close()


def _flags(st):
    A = _abi()
    return st & A.SYS_FATAL_MASK


def test_capacity_flag_matches_oracle(oracle_mesh):
    """Division past max_cells (GPUbounce.cu:1264-1270): CAPACITY is set, the
    system stops, and the cell count is the capacity, as in the oracle."""
    A = _abi()
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=2)
    p = cfg["params"]
    p.division_vol = np.float32(0.5 * cfg["v0"])  # every cell divides at the first growth step
    n = cfg["n"]
    sim, ctx, s = _pair(oracle_mesh, p, cfg["X"], cfg["Y"], cfg["Z"], n, n + 3)
    st_o = sim.step(5)
    s.step(5)
    ctx.sync()
    n1, step1, status = s.info()
    print("capacity: gpu status 0x%x oracle 0x%x, cells %d / %d" % (status, st_o, n1, sim.n), flush=True)
    assert status & A.SYS_CAPACITY and st_o & A.SYS_CAPACITY
    assert n1 == sim.n == n + 3
    ctx.close()


def test_grid_flags_match_oracle(oracle_mesh):
    """Voxel lists past voxel_capacity (propagatebound.cu:247-259, unchecked in the
    reference) and a grid larger than the voxel budget are flagged at start-up,
    as in the oracle."""
    A, E, O = _abi(), _engine(), _oracle()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=3, vcap=2)
    sim = O.Sim(p, oracle_mesh, X, Y, Z, n, max_cells=mc, budget=1 << 14)
    st_o = sim.init()
    ctx = E.Context(oracle_mesh, capacity_cells=2 * mc + 8, voxel_budget=1 << 20)
    s = ctx.add_system(p, X, Y, Z, n, max_cells=mc, voxels=1 << 14)
    status = s.info()[2]
    print("voxel capacity 2: gpu 0x%x oracle 0x%x" % (status, st_o), flush=True)
    assert status & A.SYS_VOXEL_OVERFLOW and st_o & A.SYS_VOXEL_OVERFLOW
    p2, X2, Y2, Z2, n2, mc2 = _case("tissue", n=27, seed=3)
    sim2 = O.Sim(p2, oracle_mesh, X2, Y2, Z2, n2, max_cells=mc2, budget=8)
    st_o2 = sim2.init()
    s2 = ctx.add_system(p2, X2, Y2, Z2, n2, max_cells=mc2, voxels=8)
    status2 = s2.info()[2]
    print("voxel budget 8: gpu 0x%x oracle 0x%x" % (status2, st_o2), flush=True)
    assert status2 & A.SYS_GRID_TOO_LARGE and st_o2 & A.SYS_GRID_TOO_LARGE
    s2.step(3)  # a flagged system does not step
    ctx.sync()
    assert s2.info()[1] == 0
    ctx.close()


def test_nan_flag_matches_oracle(oracle_mesh):
    """A NaN position (propagate.cu:241-248, a device trap in the reference) sets
    the NaN bit and stops the system; the other systems of a batch go on."""
    A = _abi()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=3)
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc, extra_cells=64)
    pb, Xb, Yb, Zb, nb, mcb = _case("tissue", n=8, seed=5)
    sb = ctx.add_system(pb, Xb, Yb, Zb, nb, max_cells=mcb)
    x = s.get(A.F_X)
    x[5 * 192 + 17] = np.nan
    s.set(A.F_X, x)
    xo = sim.get(A.F_X)
    xo[5 * 192 + 17] = np.nan
    sim.set(A.F_X, xo)
    st_o = sim.step(2)
    ctx.step(4)
    ctx.sync()
    status = s.info()[2]
    print("nan: gpu 0x%x oracle 0x%x" % (status, st_o), flush=True)
    assert status & A.SYS_NAN and st_o & A.SYS_NAN
    assert s.info()[1] == 0  # the step that finds it is not counted, and no step follows
    assert sb.info()[1] == 4 and sb.info()[2] & A.SYS_FATAL_MASK == 0
    ctx.close()


def test_search_list_overflow_is_flagged(oracle_mesh):
    """More contact candidates than a node's list holds, or more nodes than a
    search bin holds (a skin of 2 puts hundreds of nodes in reach): flagged, and
    the system stops instead of dropping contacts."""
    A, E = _abi(), _engine()
    p, X, Y, Z, n, mc = _case("tissue", n=27, seed=3)
    p.pair_skin = np.float32(2.0)
    ctx = E.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 22)
    s = ctx.add_system(p, X, Y, Z, n, max_cells=mc, voxels=1 << 14)
    s.step(2)
    ctx.sync()
    status = s.info()[2]
    print("skin 2: gpu 0x%x" % status, flush=True)
    assert status & (A.SYS_CAND_OVERFLOW | A.SYS_BIN_OVERFLOW)
    ctx.close()


@pytest.mark.parametrize("opt", ["soft_walls", "same_stiffness_no_sphericity", "basis_z", "basis_oblique",
                                 "pop_model", "no_angles", "daughter_com_fix"])
def test_parameter_branches_match_oracle(oracle_mesh, opt):
    """Branches off the default parameter set, with growth and division, through the
    fused step (graph path):
    * soft_walls: soft cells chosen before the start-up forces (d_Youngs_mod is uploaded
      at GPUbounce.cu:691, before the initial CalculateConForce at :978-1004), the
      variable-stiffness PressureUpdate (PressureKernels.cu:15-34) switching at
      phase_count, recalc_r0 keeping soft cells from dividing (volume.cu:155-160),
      daughters taking stiffness1 (GPUbounce.cu:1257-1262), and the rigid-box walls (propagate.cu:473-510)
      with nodes inside the threshold distance of all six walls;
    * same_stiffness_no_sphericity: daughters keep the parent's stiffness and
      checkSphericity is off (volume.cu:145-153);
    * basis_z, basis_oblique: division planes from GetRandomVectorBasis
      (RandomVector.h:26-78, called at GPUbounce.cu:1230-1231) about the basis
      (0, 0, 1) and about an oblique one, which take its two branches;
    * pop_model: the growth rate of the population model, rMax (1 - n / maxPop)
      (GPUbounce.cu:1077-1088), which stops the growth at maxPop cells;
    * no_angles: constrainAngles off, no bending forces (propagate.cu:332-337);
    * daughter_com_fix: the engine's opt-in fix of the stale daughter centre (the
      centres of parent and daughter from their new nodes right after division)."""
    A, E, O = _abi(), _engine(), _oracle()
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=4)
    p = cfg["params"]
    p.division_vol = np.float32(1.02 * cfg["v0"])
    n, mc = cfg["n"], 64
    X, Y, Z = cfg["X"].copy(), cfg["Y"].copy(), cfg["Z"].copy()
    p.use_different_stiffnesses = 1
    p.stiffness2 = np.float32(0.5 * p.stiffness1)
    soft = [1, 4, 6]
    if opt == "soft_walls":
        p.recalc_r0 = 1
        p.daught_same_stiffness = 0
        p.phase_count = 120
        p.use_rigid_box = 1
        p.thresh_dist = np.float32(0.1)
        live = np.zeros((n, 192), bool)
        live[:, :180] = True
        for a, arr in enumerate((X, Y, Z)):  # lowest node 0.03 above the wall at 0
            v = arr.reshape(n, 192)
            v[live] += np.float32(0.03) - v[live].min()
            p.box_max[a] = float(v[live].max() + np.float32(0.03))
    elif opt == "same_stiffness_no_sphericity":
        p.recalc_r0 = 0
        p.daught_same_stiffness = 1
        p.check_sphericity = 0
        p.phase_count = 10 ** 6
    elif opt in ("no_angles", "daughter_com_fix"):
        p.recalc_r0 = 0
        p.daught_same_stiffness = 1
        p.phase_count = 10 ** 6
        if opt == "no_angles":
            p.constrain_angles = 0
        else:
            p.daughter_com_fix = 1
    elif opt == "pop_model":
        p.recalc_r0 = 0
        p.daught_same_stiffness = 1
        p.phase_count = 10 ** 6
        p.do_pop_model = 1
        p.max_pop = np.float32(12.0)
    else:
        p.recalc_r0 = 0
        p.daught_same_stiffness = 1
        p.phase_count = 10 ** 6
        p.use_div_plane_basis = 1
        p.div_plane_basis[:] = [0.0, 0.0, 1.0] if opt == "basis_z" else [0.2, 0.6, 0.77]
    Ecell = np.full(mc, p.stiffness1, np.float32)
    Ecell[soft] = p.stiffness2
    sim = O.Sim(p, oracle_mesh, X, Y, Z, n, max_cells=mc, budget=1 << 16)
    sim.set(A.F_STIFFNESS, Ecell)
    sim.init()
    ctx = E.Context(oracle_mesh, capacity_cells=mc + 16, voxel_budget=(1 << 16) * 70)
    s = ctx.add_system(p, X, Y, Z, n, max_cells=mc, voxels=1 << 16)
    s.set(A.F_STIFFNESS, Ecell[:n])
    s.startup()
    _cmp(sim, s, [A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY, A.F_FDZ], n, label=opt + " start-up")
    total = 0
    for k in range(6):
        sim.step(50)
        s.step(50)
        ctx.sync()
        total += 50
        n1, st1, status = s.info()
        print("%s: %d steps, gpu %d cells status 0x%x, oracle %d cells" % (opt, total, n1, status, sim.n), flush=True)
        assert status & A.SYS_FATAL_MASK == 0
        assert n1 == sim.n and st1 == sim.stepno
    nn = sim.n
    assert nn > n, "no division happened; the test does not exercise the branches"
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_PRESSURE,
                  A.F_VOLUME], nn, label=opt)
    e1, e0 = s.get(A.F_STIFFNESS)[:nn], sim.get(A.F_STIFFNESS)[:nn]
    assert np.array_equal(e1, e0)
    if opt == "soft_walls":
        assert np.all(s.get(A.F_PRESSURE)[soft] == 0)  # soft cells lose their pressure after phase_count
        assert np.all(e1[n:] == p.stiffness1) and np.all(e1[soft] == p.stiffness2)
        x = s.get(A.F_X).reshape(-1, 192)[:nn, :180]
        assert x.min() < p.thresh_dist, "no node near the lower wall"
    elif opt == "same_stiffness_no_sphericity":
        assert np.sum(e1 == p.stiffness2) > len(soft), "no soft daughter"
    ctx.close()


def test_staggered_divisions_in_a_tissue_match_oracle(oracle_mesh):
    """In a 64-cell tissue (growth scales U[0.5, 1.5]) the cells divide a few at
    a time over about 20 steps, as their volumes pass the division volume one after
    another.  After each such step only the cells near a divided cell search their
    contact candidates again; cells farther away keep their lists.  Cell counts must agree with the
    oracle after every 5 steps, and the state bit for bit at the end."""
    A = _abi()
    from cellsim3d_amd import configs
    m, X, Y, Z, info = configs.lattice_tissue(64, 0.40, seed=11)
    p = configs.base_params(11)
    p.division_vol = np.float32(1.02 * info["v0"])
    p.div_time_steps = 10 ** 6
    p.voxel_capacity = 64
    n, mc = 64, 256
    gs = np.ones(mc, np.float32)
    gs[:n] = np.random.default_rng(5).uniform(0.5, 1.5, n).astype(np.float32)
    sim, ctx, s = _pair(oracle_mesh, p, X, Y, Z, n, mc)
    sim.set(A.F_GROWTH_SCALE, gs)
    s.set(A.F_GROWTH_SCALE, gs[:n])
    ctx.set_graph(False)
    counts = []
    for k in range(20):
        sim.step(5)
        s.step(5)
        ctx.sync()
        n1, st1, status = s.info()
        assert status & A.SYS_FATAL_MASK == 0
        assert n1 == sim.n and st1 == sim.stepno, "after %d steps: gpu %d cells, oracle %d" % (st1, n1, sim.n)
        counts.append(n1)
    print("cell counts every 5 steps:", counts, flush=True)
    jumps = [b - a for a, b in zip([n] + counts, counts)]
    assert counts[-1] > n and sum(1 for j in jumps if j) >= 3 and max(jumps) < n, "divisions not staggered: %s" % counts
    nn = sim.n
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FDX, A.F_PRESSURE], nn,
         label="staggered divisions")
    ctx.close()


def test_results_do_not_depend_on_the_skin(oracle_mesh):
    """The contact-candidate skin is a performance knob only (the autotuner picks it:
    it sets the edge of the search bins and how often candidates are searched
    again).  The 64-cell tissue of the staggered-division test, stepped 100 steps
    through its divisions with skins 0.01, 0.05 and 0.15 as three systems of one
    batch (graph path), ends bit for bit where the oracle ends; the oracle scans
    every node of every cell in the voxel lists and has no skin."""
    A, E = _abi(), _engine()
    from cellsim3d_amd import configs
    m, X, Y, Z, info = configs.lattice_tissue(64, 0.40, seed=11)
    p = configs.base_params(11)
    p.division_vol = np.float32(1.02 * info["v0"])
    p.div_time_steps = 10 ** 6
    p.voxel_capacity = 64
    n, mc = 64, 256
    gs = np.ones(mc, np.float32)
    gs[:n] = np.random.default_rng(5).uniform(0.5, 1.5, n).astype(np.float32)
    O = _oracle()
    sim = O.Sim(p, oracle_mesh, X, Y, Z, n, max_cells=mc, budget=1 << 16)
    sim.set(A.F_GROWTH_SCALE, gs)
    sim.init()
    ctx = E.Context(oracle_mesh, capacity_cells=3 * mc + 16, voxel_budget=1 << 24)
    systems = []
    for skin in (0.01, 0.05, 0.15):
        q = p.copy()
        q.pair_skin = np.float32(skin)
        sy = ctx.add_system(q, X, Y, Z, n, max_cells=mc)
        sy.set(A.F_GROWTH_SCALE, gs[:n])
        systems.append((skin, sy))
    sim.step(100)
    ctx.step(100)
    ctx.sync()
    nn = sim.n
    assert nn > n, "no division happened"
    for skin, sy in systems:
        n1, st1, status = sy.info()
        print("skin %.2f: %d cells after %d steps, status 0x%x (oracle %d cells)" % (skin, n1, st1, status, nn),
              flush=True)
        assert status & A.SYS_FATAL_MASK == 0
        assert n1 == nn and st1 == sim.stepno
        _cmp(sim, sy, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FDX, A.F_PRESSURE], nn,
             label="skin %.2f" % skin)
    ctx.close()


def test_division_burst_with_aliased_bins_matches_oracle(oracle_mesh):
    """The large-tissue paths of the colony at a size the oracle steps in seconds: a
    400-cell tissue at volume fraction 0.40 whose cells reach the division volume
    within a few dozen steps of each other (up to about 20 divisions in one step, 400
    -> about 780 cells in 200 steps), voxel lists of 128, the grid rebuilt every 10
    steps and after divisions, and a bin budget of 98,304 search bins where the
    tissue's extent needs about 128,000 bins of edge attraction_range + skin: the
    periodic bins alias distant regions into the same bins (at most 16 nodes share a
    bin through the run, by a host model of the binning; a budget of 32,768 overflows
    the 32-node bins once cells divide, which the engine flags), and lists of divided
    and moved cells longer than the partial search takes force full searches.  Cell
    counts must agree with the oracle every 10 steps, and the state and the voxel grid
    bit for bit at the end."""
    A, E, O = _abi(), _engine(), _oracle()
    from cellsim3d_amd import configs
    n, mc, bins = 400, 1024, 98304
    m, X, Y, Z, info = configs.lattice_tissue(n, 0.40, seed=9)
    p = configs.base_params(9)
    p.division_vol = np.float32(1.01 * info["v0"])
    p.div_time_steps = 10 ** 6
    p.voxel_capacity = 128
    p.grid_interval = 10
    h = (float(p.attraction_range) + float(p.pair_skin)) * 1.001
    need = np.prod([(float(q.max()) - float(q.min())) / h + 5
                    for q in (X.reshape(n, 192)[:, :180], Y.reshape(n, 192)[:, :180], Z.reshape(n, 192)[:, :180])])
    assert need > 1.2 * bins, "the bin budget does not alias: %d bins needed" % need
    budget = 1 << 15
    sim = O.Sim(p, oracle_mesh, X, Y, Z, n, max_cells=mc, budget=budget)
    assert sim.init() == 0
    ctx = E.Context(oracle_mesh, capacity_cells=mc, voxel_budget=budget * (p.voxel_capacity + 1),
                    bin_budget=bins * A.BIN_WORDS + 64)
    s = ctx.add_system(p, X, Y, Z, n, max_cells=mc, voxels=budget)
    counts = []
    for k in range(20):
        assert sim.step(10) & A.SYS_FATAL_MASK == 0
        s.step(10)
        ctx.sync()
        n1, st1, status = s.info()
        assert status & A.SYS_FATAL_MASK == 0, hex(status)
        assert n1 == sim.n and st1 == sim.stepno, "after %d steps: gpu %d cells, oracle %d" % (st1, n1, sim.n)
        counts.append(n1)
    print("cell counts every 10 steps:", counts, flush=True)
    jumps = [b - a for a, b in zip([n] + counts, counts)]
    assert max(jumps) >= 30 and counts[-1] > 1.8 * n, counts
    nn = sim.n
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY,
                  A.F_FDZ, A.F_PRESSURE, A.F_VOLUME, A.F_CMX, A.F_CMY, A.F_CMZ], nn, label="division burst")
    _cmp_grid(sim, s)
    ctx.close()


def test_growing_colony_matches_oracle(oracle_mesh):
    """The growing-colony workload at reduced size: one cell with per-cell growth
    scales U[0.8, 1.2], division at twice the cell volume, voxel lists of 128, the
    grid rebuilt every 10 steps and after divisions.  It grows past 100 cells in
    about 11,000 steps, dividing in parent order first and then a few cells at a
    time, with partial searches after most divisions.  Cell counts must agree with
    the oracle after every 250 steps; at the end positions, velocities, forces,
    pressures, volumes, centres and the voxel grid are compared bit for bit."""
    A = _abi()
    E, O = _engine(), _oracle()
    from cellsim3d_amd import configs
    mc = 256
    cfg = configs.colony(seed=1, max_cells=mc)
    p = cfg["params"]
    p.grid_interval = 10
    p.voxel_capacity = 128
    budget = 1 << 14
    sim = O.Sim(p, oracle_mesh, cfg["X"], cfg["Y"], cfg["Z"], 1, max_cells=mc, budget=budget)
    sim.set(A.F_GROWTH_SCALE, cfg["growth_scale"])
    assert sim.init() == 0
    ctx = E.Context(oracle_mesh, capacity_cells=mc, voxel_budget=budget * (p.voxel_capacity + 1))
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], 1, max_cells=mc, voxels=budget)
    s.set(A.F_GROWTH_SCALE, cfg["growth_scale"])
    counts = []
    for k in range(56):
        assert sim.step(250) & A.SYS_FATAL_MASK == 0
        s.step(250)
        ctx.sync()
        n1, st1, status = s.info()
        assert status & A.SYS_FATAL_MASK == 0, hex(status)
        assert n1 == sim.n and st1 == sim.stepno, "after %d steps: gpu %d cells, oracle %d" % (st1, n1, sim.n)
        counts.append(n1)
        print("colony: %d steps, %d cells" % (st1, n1), flush=True)
        if n1 >= 100:
            break
    print("colony cell counts every 250 steps:", counts, flush=True)
    assert counts[-1] >= 100
    n = sim.n
    _cmp(sim, s, [A.F_X, A.F_Y, A.F_Z, A.F_VX, A.F_VY, A.F_VZ, A.F_FCX, A.F_FCY, A.F_FCZ, A.F_FDX, A.F_FDY,
                  A.F_FDZ, A.F_PRESSURE, A.F_VOLUME, A.F_AREA, A.F_CMX, A.F_CMY, A.F_CMZ], n, label="colony")
    _cmp_grid(sim, s)
    ctx.close()
