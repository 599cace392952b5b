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


@pytest.mark.parametrize("pool_cells", [4096, 48])
def test_serving_results_equal_alone(pool_cells):
    """pool_cells = 48 forces queueing: later jobs wait for finished ones."""
    from cellsim3d_amd import abi, serving
    gm = _mesh()
    jobs = _jobs([(0.0, 1, 40), (0.0, 8, 75), (0.01, 27, 60), (0.02, 3, 123), (0.03, 12, 90)])
    print("serving jobs:", [(j.n, j.horizon, j.max_cells) for j in jobs], "pool", pool_cells, flush=True)
    srv = serving.Server(gm, hbm_budget_bytes=1e12, max_cells_total=pool_cells, chunk_steps=20, graph_steps=10)
    srv.run(jobs)
    srv.close()
    for j in jobs:
        assert j.result is not None and j.result["status"] & abi.SYS_FATAL_MASK == 0
        ref = _alone(gm, j)
        assert j.result["n"] == ref["n"]
        for f in (abi.F_X, abi.F_Y, abi.F_Z):
            assert np.array_equal(j.result[f], ref[f]), "job %d field %d differs from its solo run" % (j.id, f)


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
