"""The batched ensemble and the serving loop against the CPU oracle.

* The 256-tissue ensemble workload (N ~ log-U[1, 512], parameters x U[0.7,
  1.3], growth on) stepped batched through the captured all-systems graph:
  every tissue bit for bit against its own oracle run.
* Small tissues in one batch whose cells divide, with cell rooms below the
  partial-search list size: each tissue against its oracle run and against
  itself stepped alone.
* Served requests (Poisson-like arrivals, a pool small enough that requests
  queue, admission while others run): each result at its horizon against the
  oracle run of that request alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NODE = ("F_X", "F_Y", "F_Z", "F_VX", "F_VY", "F_VZ")


def _mesh():
    from cellsim3d_amd import configs, engine
    from cellsim3d_amd import mesh as M
    m = configs.unit_mesh()
    return engine.build_mesh(M.tables(m), configs.rest_positions(m))


def _oracle_run(gm, params, X, Y, Z, n, max_cells, steps, growth_scale=None):
    from cellsim3d_amd import abi
    from oracle import oracle as O
    sim = O.Sim(params, gm, X, Y, Z, n, max_cells=max_cells, budget=max(4096, 32 * max_cells))
    if growth_scale is not None:
        sim.set(abi.F_GROWTH_SCALE, growth_scale)
    sim.init()
    st = sim.step(steps)
    return sim, st


def _same(label, sys_or_arrays, sim, n, fields):
    from cellsim3d_amd import abi
    for name in fields:
        f = getattr(abi, name)
        k = n * 192 if f in abi.NODE_FIELDS else n
        a = sys_or_arrays[f][:k] if isinstance(sys_or_arrays, dict) else sys_or_arrays.get(f)[:k]
        b = sim.get(f)[:k]
        assert np.array_equal(a, b), "%s: field %s differs from the oracle (max |d| %g)" % (
            label, name, float(np.max(np.abs(a.astype(np.float64) - b))))


def test_ensemble_256_tissues_match_oracle():
    from cellsim3d_amd import abi, ensemble
    steps = 60
    gm, tissues = ensemble.build(256, 1, 1, 512, steps=5000, grid_interval=10)
    cells = sum(t["n"] for t in tissues)
    print("ensemble: 256 tissues, %d cells, %d batched steps" % (cells, steps), flush=True)
    dt, ctx, systems, _ = ensemble.run_batched(gm, tissues, steps, wg_per_cu=16, graph_steps=10)
    got = []
    for s in systems:
        n, step, status = s.info()
        got.append((n, step, status, {getattr(abi, f): s.get(getattr(abi, f))[:n * 192].copy()
                                      for f in NODE + ("F_PRESSURE",)}))
    ctx.close()
    for i, (t, (n, step, status, arrays)) in enumerate(zip(tissues, got)):
        sim, st = _oracle_run(gm, t["params"], t["X"], t["Y"], t["Z"], t["n"], t["max_cells"], steps)
        assert (n, step) == (sim.n, sim.stepno) and (status & abi.SYS_FATAL_MASK) == (st & abi.SYS_FATAL_MASK), \
            "tissue %d: gpu (%d cells, step %d, 0x%x) oracle (%d, %d, 0x%x)" % (i, n, step, status, sim.n,
                                                                              sim.stepno, st)
        _same("tissue %d (%d cells)" % (i, n), arrays, sim, n, NODE + ("F_PRESSURE",))
        if i % 32 == 31:
            print("ensemble: %d tissues match the oracle" % (i + 1), flush=True)


def test_batched_divisions_match_oracle_and_alone():
    """Tissues of 1..6 cells with room for one division per cell (cap 2n + 2,
    below the 64-entry partial-search list), division just above the cell
    volume: several divisions per tissue within the run, in one batch."""
    from cellsim3d_amd import abi, configs, engine
    gm = _mesh()
    tissues = []
    for i, n in enumerate((1, 2, 3, 4, 5, 6, 3, 2)):
        _, X, Y, Z, info = configs.lattice_tissue(n, 0.30, seed=40 + i)
        p = configs.base_params(5 + i)
        p.division_vol = np.float32(1.02 * info["v0"])
        p.div_time_steps = 10 ** 6
        p.voxel_capacity = 64
        p.grid_interval = 10
        tissues.append({"params": p, "X": X, "Y": Y, "Z": Z, "n": n, "max_cells": 2 * n + 2})
    steps = 400
    cap = sum(t["max_cells"] for t in tissues)
    ctx = engine.Context(gm, capacity_cells=cap, voxel_budget=1 << 22, bin_budget=len(tissues) * 16384 * abi.BIN_WORDS + 64)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    ctx.set_graph_steps(10)
    for k in range(steps // 50):
        ctx.step(50)
        ctx.sync()
    batched = []
    for s in systems:
        n, step, status = s.info()
        batched.append((n, step, status, {getattr(abi, f): s.get(getattr(abi, f))[:n * 192].copy()
                                          for f in NODE + ("F_PRESSURE",)}))
    ctx.close()
    grew = 0
    for i, (t, (n, step, status, arrays)) in enumerate(zip(tissues, batched)):
        sim, st = _oracle_run(gm, t["params"], t["X"], t["Y"], t["Z"], t["n"], t["max_cells"], steps)
        print("tissue %d: %d -> %d cells (oracle %d), status 0x%x" % (i, t["n"], n, sim.n, status), flush=True)
        assert (n, step, status) == (sim.n, sim.stepno, st)
        _same("batched tissue %d" % i, arrays, sim, n, NODE + ("F_PRESSURE",))
        grew += n > t["n"]
        c1 = engine.Context(gm, capacity_cells=t["max_cells"] + 4, voxel_budget=1 << 20,
                            bin_budget=16384 * abi.BIN_WORDS + 64)
        s1 = c1.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"])
        s1.step(steps)
        c1.sync()
        assert s1.info() == (n, step, status)
        for f, a in arrays.items():
            assert np.array_equal(s1.get(f)[:len(a)], a), "tissue %d field %d: batched differs from alone" % (i, f)
        c1.close()
    assert grew >= 6, "too few tissues divided (%d)" % grew


def test_served_requests_match_oracle():
    """Ten requests arriving at once; a 40-cell pool admits the first few and
    makes the others wait for finished ones, so later requests are admitted
    while earlier ones are still being stepped.  The scheduler steps only the
    resident jobs with the least work left, up to 12 cells per chunk, and pauses
    the others on the device, so most jobs are paused and resumed on the way."""
    from cellsim3d_amd import abi, serving
    gm = _mesh()
    specs = [(0.0, 1, 60), (0.0, 8, 140), (0.0, 20, 90), (0.0, 3, 230), (0.0, 12, 110), (0.0, 2, 75),
             (0.0, 27, 50), (0.0, 5, 300), (0.0, 9, 45), (0.0, 4, 160)]
    reqs = [{"arrival": a, "n": n, "horizon": h, "seed": 21 + i} for i, (a, n, h) in enumerate(specs)]
    jobs = serving.make_jobs(reqs, grid_interval=10)
    print("serving: %d requests into a 40-cell pool" % len(jobs), flush=True)
    fields = tuple(getattr(abi, f) for f in NODE)
    srv = serving.Server(gm, hbm_budget_bytes=1e12, max_cells_total=40, chunk_steps=20, graph_steps=10,
                         fields=fields, policy="least_work", active_cells=12)
    srv.run(jobs)
    srv.close()
    assert srv.pauses >= 3, "the scheduler paused too few jobs (%d)" % srv.pauses
    waited = sum(1 for j in jobs if j.admitted > j.arrival + 1e-3)
    print("serving: %d of %d requests waited for room" % (waited, len(jobs)), flush=True)
    assert waited >= 1
    for j in jobs:
        assert j.result is not None and (j.result["status"] & abi.SYS_FATAL_MASK) == 0
        sim, st = _oracle_run(gm, j.params, j.X, j.Y, j.Z, j.n, j.max_cells, j.horizon)
        assert j.result["n"] == sim.n
        _same("request %d (%d cells, %d steps)" % (j.id, j.n, j.horizon), j.result, sim, sim.n, NODE)


def test_room_growth_keeps_results():
    """Tissues that outgrow their room are moved to larger rooms between chunks
    (cs3d_system_grow) while the others keep stepping: every tissue ends bit for
    bit where its oracle run (with room to spare) ends."""
    from cellsim3d_amd import abi, configs, engine, ensemble
    gm = _mesh()
    tissues = []
    for i, n in enumerate((1, 2, 4, 3)):
        _, X, Y, Z, info = configs.lattice_tissue(n, 0.30, seed=60 + i)
        p = configs.base_params(9 + i)
        p.division_vol = np.float32(1.02 * info["v0"])
        p.div_time_steps = 10 ** 6
        p.voxel_capacity = 64
        tissues.append({"params": p, "X": X, "Y": Y, "Z": Z, "n": n, "max_cells": 2 * n + 2})
    steps = 600
    ctx = engine.Context(gm, capacity_cells=512, voxel_budget=1 << 24)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    ctx.set_graph_steps(10)
    moves = ensemble.run_chunks(ctx, abi.ALL, steps, systems, chunk=10)
    ctx.sync()
    print("room moves: %d, cells %s, rooms %s" % (moves, [s.n for s in systems], [s.max_cells for s in systems]),
          flush=True)
    assert moves >= 2
    for i, (t, s) in enumerate(zip(tissues, systems)):
        n, step, status = s.info()
        sim, st = _oracle_run(gm, t["params"], t["X"], t["Y"], t["Z"], t["n"], 64, steps)
        assert (n, step, status) == (sim.n, sim.stepno, st), "tissue %d" % i
        _same("grown tissue %d" % i, s, sim, n, NODE + ("F_PRESSURE", "F_VOLUME"))
    work = ctx.work(0, len(systems))
    assert all(w > 0 for w in work)
    ctx.close()
