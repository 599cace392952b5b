"""Checkpoint / restart: a system saved mid-run and loaded into a fresh context
continues bit-identically to the one that kept running (growth and division
included), and a mismatched mesh is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_restart_continues_bit_identically(oracle_mesh, tmp_path):
    from cellsim3d_amd import abi, configs, engine
    cfg = configs.plain_small(seed=3)
    p = cfg["params"]
    p.division_vol = np.float32(1.05 * cfg["v0"])  # divisions within the run
    p.grid_interval = 7
    mc = 64
    print("checkpoint test: 8 cells, 150 + 150 steps", flush=True)
    ctx = engine.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 22)
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], cfg["n"], max_cells=mc, voxels=1 << 14)
    s.step(150)
    ctx.sync()
    path = tmp_path / "state.ckpt"
    s.save(path)
    s.step(150)
    ctx.sync()
    ref = {f: s.get(f) for f in (abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX, abi.F_FCX, abi.F_FDZ, abi.F_PRESSURE,
                                   abi.F_CMX)}
    n_ref, step_ref, st_ref = s.info()
    ctx.close()
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
