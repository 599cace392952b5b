"""Batched ensembles and the serving loop on the GPU: every tissue's result is
bit-identical whether it runs alone on the drop-in path, in a batch, or in the
continuously batched request loop (admitted mid-run, sharing launches with
tissues of other sizes and horizons)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mesh():
    from cellsim3d_amd import configs, engine
    from cellsim3d_amd import mesh as M
    m = configs.unit_mesh()
    return engine.build_mesh(M.tables(m), configs.rest_positions(m))


def _alone(gm, job):
    from cellsim3d_amd import abi, engine
    ctx = engine.Context(gm, capacity_cells=job.max_cells + 4)
    s = ctx.add_system(job.params, job.X, job.Y, job.Z, job.n, max_cells=job.max_cells)
    s.step(job.horizon)
    ctx.sync()
    n = s.n
    out = {f: s.get(f)[:n * 192].copy() for f in (abi.F_X, abi.F_Y, abi.F_Z)}
    out["n"] = n
    ctx.close()
    return out


def _jobs(specs):
    from cellsim3d_amd import serving
    reqs = [{"arrival": a, "n": n, "horizon": h, "seed": 11 + i} for i, (a, n, h) in enumerate(specs)]
    return serving.make_jobs(reqs, grid_interval=1)


@pytest.mark.parametrize("pool_cells,policy,active", [(4096, "all", 0), (48, "all", 0), (4096, "least_work", 16),
                                                       (48, "least_work", 8)])
def test_serving_results_equal_alone(pool_cells, policy, active):
    """pool_cells = 48 forces queueing: later jobs wait for finished ones.  The
    least-work scheduler with a small active_cells pauses and resumes jobs; its
    requests arrive together, so that several are resident at once."""
    from cellsim3d_amd import abi, serving
    gm = _mesh()
    late = 0.0 if policy == "least_work" else 1.0
    jobs = _jobs([(0.0, 1, 40), (0.0, 8, 75), (0.01 * late, 27, 60), (0.02 * late, 3, 123), (0.03 * late, 12, 90)])
    print("serving jobs:", [(j.n, j.horizon, j.max_cells) for j in jobs], "pool", pool_cells, policy, flush=True)
    srv = serving.Server(gm, hbm_budget_bytes=1e12, max_cells_total=pool_cells, chunk_steps=20, graph_steps=10,
                         policy=policy, active_cells=active or 16384)
    srv.run(jobs)
    srv.close()
    if policy == "least_work":
        assert srv.pauses >= 1
    for j in jobs:
        assert j.result is not None and j.result["status"] & abi.SYS_FATAL_MASK == 0
        ref = _alone(gm, j)
        assert j.result["n"] == ref["n"]
        for f in (abi.F_X, abi.F_Y, abi.F_Z):
            assert np.array_equal(j.result[f], ref[f]), "job %d field %d differs from its solo run" % (j.id, f)


def test_paused_systems_resume_where_an_uninterrupted_run_ends():
    """cs3d_systems_set_horizon pauses systems at or below their step and resumes
    them with a later horizon; a paused and resumed tissue ends bit for bit where
    the same tissue stepped without a pause ends, and the steps of each system
    follow its horizons."""
    from cellsim3d_amd import abi, engine
    gm = _mesh()
    jobs = _jobs([(0.0, 5, 100), (0.0, 9, 100), (0.0, 2, 100)])
    ctx = engine.Context(gm, capacity_cells=64)
    sys_ = [ctx.add_system(j.params, j.X, j.Y, j.Z, j.n, max_cells=j.max_cells) for j in jobs]
    ids = [s.id for s in sys_]
    hi = max(ids) + 1
    h = np.full(hi, -1, dtype=np.int64)
    h[ids[0]], h[ids[1]], h[ids[2]] = 30, 0, -1  # stop at 30, paused from the start, no limit
    ctx.set_horizons(0, h)
    ctx.step(40)
    steps = ctx.poll(0, hi)[1]
    assert [int(steps[i]) for i in ids] == [30, 0, 40]
    h[ids[0]], h[ids[1]], h[ids[2]] = 60, 45, 40  # resume the first two, pause the third
    ctx.set_horizons(0, h)
    ctx.step(70)
    steps = ctx.poll(0, hi)[1]
    assert [int(steps[i]) for i in ids] == [60, 45, 40]
    ctx.set_horizons(0, np.full(hi, -1, dtype=np.int64))
    ctx.step(20)
    ctx.sync()
    for j, s, total in zip(jobs, sys_, (80, 65, 60)):
        ref = engine.Context(gm, capacity_cells=j.max_cells + 4)
        r = ref.add_system(j.params, j.X, j.Y, j.Z, j.n, max_cells=j.max_cells)
        r.step(total)
        ref.sync()
        assert s.info()[1] == total and r.n == s.n
        for f in (abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX):
            assert np.array_equal(s.get(f)[:s.n * 192], r.get(f)[:r.n * 192]), "system %d field %d" % (s.id, f)
        ref.close()
    ctx.close()


def test_ensemble_batched_equals_plain():
    from cellsim3d_amd import abi, ensemble
    gm, tissues = ensemble.build(n_tissues=6, seed=3, nmin=1, nmax=40, steps=60, grid_interval=10)
    print("ensemble sizes:", [t["n"] for t in tissues], flush=True)
    dt, ctx, systems, _ = ensemble.run_batched(gm, tissues, 60, wg_per_cu=4, graph_steps=10)
    batched = [s.get(abi.F_X)[:s.n * 192].copy() for s in systems]
    ctx.close()
    for t, xb in zip(tissues, batched):
        from cellsim3d_amd import engine
        c = engine.Context(gm, capacity_cells=t["max_cells"] + 4)
        s = c.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"])
        s.step(60)
        c.sync()
        assert np.array_equal(s.get(abi.F_X)[:s.n * 192], xb)
        c.close()
