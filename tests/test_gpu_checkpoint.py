"""Checkpoint / restart.  A system saved mid-run and loaded into a fresh
context continues bit-identically to the one that kept running (growth and
division included, and with growth off, where volumes are no longer
recomputed and must come from the file), and it ends where the CPU oracle
run straight through ends.  A mismatched mesh is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("F_X", "F_Y", "F_Z", "F_VX", "F_VY", "F_VZ", "F_FCX", "F_FDZ", "F_PRESSURE", "F_CMX", "F_VOLUME", "F_AREA")


def _case(growth, seed=3):
    from cellsim3d_amd import configs
    cfg = configs.plain_small(seed=seed)
    p = cfg["params"]
    p.grid_interval = 7
    if growth:
        p.division_vol = np.float32(1.05 * cfg["v0"])  # divisions within the run
    else:
        # volumes are computed while growing (GPUbounce.cu:1168); growth ends at step 100
        p.div_time_steps = 100
        p.non_div_time_steps = 1000
    return cfg, p


@pytest.mark.parametrize("growth", [True, False])
def test_restart_continues_bit_identically(oracle_mesh, tmp_path, growth):
    from cellsim3d_amd import abi, engine
    from oracle import oracle as O
    cfg, p = _case(growth)
    mc = 64
    fields = [getattr(abi, f) for f in FIELDS]
    print("checkpoint test: 8 cells, 150 + 150 steps, growth %s" % growth, flush=True)
    ctx = engine.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 22)
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], cfg["n"], max_cells=mc, voxels=1 << 14)
    s.step(150)
    ctx.sync()
    path = tmp_path / "state.ckpt"
    s.save(path)
    s.step(150)
    ctx.sync()
    ref = {f: s.get(f) for f in fields}
    n_ref, step_ref, st_ref = s.info()
    ctx.close()
    if growth:
        assert n_ref > cfg["n"], "the run was meant to include divisions"
    ctx2 = engine.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 22)
    t = ctx2.load_system(path, voxels=1 << 14)
    assert t.info()[1] == 150
    t.step(150)
    ctx2.sync()
    n2, step2, st2 = t.info()
    assert (n2, step2, st2) == (n_ref, step_ref, st_ref)
    for f, a in ref.items():
        assert np.array_equal(t.get(f), a), "field %d differs after restart" % f
    # the restarted run ends where the oracle, run straight through, ends
    sim = O.Sim(p, oracle_mesh, cfg["X"], cfg["Y"], cfg["Z"], cfg["n"], max_cells=mc, budget=1 << 14)
    sim.init()
    sim.step(300)
    assert sim.n == n2 and sim.stepno == 300
    for f in fields:
        a, b = t.get(f), sim.get(f)
        k = n2 * 192 if f in abi.NODE_FIELDS else n2
        assert np.array_equal(a[:k], b[:k]), "field %d differs from the oracle after restart" % f
    ctx2.close()


def test_checkpoint_refuses_another_mesh(oracle_mesh, tmp_path):
    from cellsim3d_amd import configs, engine
    from cellsim3d_amd import mesh as M
    cfg = configs.plain_small(seed=1)
    ctx = engine.Context(oracle_mesh, capacity_cells=40)
    s = ctx.add_system(cfg["params"], cfg["X"], cfg["Y"], cfg["Z"], cfg["n"], max_cells=16, voxels=1 << 12)
    s.save(tmp_path / "a.ckpt")
    ctx.close()
    big = configs.meshmod.gp30_mesh(2.0)
    other = engine.build_mesh(M.tables(big), configs.rest_positions(big))
    ctx2 = engine.Context(other, capacity_cells=40)
    with pytest.raises(engine.EngineError):
        ctx2.load_system(tmp_path / "a.ckpt")
    ctx2.close()
