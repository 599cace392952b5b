"""Long-run parity on the plain small configuration (8 unit-radius cells,
2x2x2 lattice, growth and division at 2 V0): the GPU engine and the CPU
oracle stepped side by side, compared every 100 steps (cell counts exactly;
positions, velocities and pressures within 1e-5 relative).  It runs for
minutes on the CPU side, so it is opt-in: set CS3D_LONG_STEPS (e.g. 3000)."""
import os

import numpy as np
import pytest

from conftest import relerr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.environ.get("CS3D_LONG_STEPS"),
                                                  reason="opt-in long run (set CS3D_LONG_STEPS)")]


def test_plain_small_long_run(oracle_mesh):
    from cellsim3d_amd import abi, configs, engine
    from oracle import oracle as O
    steps = int(os.environ["CS3D_LONG_STEPS"])
    cfg = configs.plain_small(seed=1, steps=steps)
    p, n, mc = cfg["params"], cfg["n"], cfg["max_cells"]
    sim = O.Sim(p, oracle_mesh, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, budget=1 << 16)
    sim.init()
    ctx = engine.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 24)
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, voxels=1 << 16)
    worst = 0.0
    exact = []
    done = 0
    while done < steps:
        k = min(100, steps - done)
        st_o = sim.step(k)
        s.step(k)
        ctx.sync()
        done += k
        n_g, step_g, status = s.info()
        assert n_g == sim.n, "cell counts differ at step %d: %d vs %d" % (done, n_g, sim.n)
        assert (status & abi.SYS_FATAL_MASK) == (st_o & abi.SYS_FATAL_MASK)
        nn = n_g * 192
        for f in (abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX, abi.F_VY, abi.F_VZ):
            a, b = s.get(f)[:nn], sim.get(f)[:nn]
            e = relerr(a, b)
            worst = max(worst, e)
            exact.append(float(np.mean(a == b)))
            assert e <= 1e-5, "field %d rel err %g at step %d" % (f, e, done)
        pa, pb = s.get(abi.F_PRESSURE)[:n_g], sim.get(abi.F_PRESSURE)[:n_g]
        assert relerr(pa, pb) <= 1e-5
        print("step %d: %d cells, worst rel err so far %.3g, bit-exact fraction %.4f" %
              (done, n_g, worst, exact[-1]), flush=True)
        if status & abi.SYS_FATAL_MASK:
            break
    ctx.close()
