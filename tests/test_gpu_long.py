"""Long-run parity on the plain small configuration at its stated length: 8
unit-radius cells on a 2x2x2 lattice, growth and division at 2 V0, 10^4 steps
(the 8 cells grow to 287).  The GPU engine and the CPU oracle step side by side
and are compared every 100 steps: cell counts, steps and status exactly;
positions, velocities and pressures bit for bit (the 1e-5 per-element force
tolerance is checked first so that a failure states the size of a difference).
The kernels and the oracle share the angle force's double acos (cs3d_acos.h):
with the device and host math libraries' acos instead, one angle near a float
midpoint (step 8277) made a few values differ by an ulp from there on.
CS3D_LONG_STEPS sets another length."""
import os

import numpy as np
import pytest

from conftest import elemerr

pytestmark = pytest.mark.gpu


def test_plain_small_long_run(oracle_mesh):
    from cellsim3d_amd import abi, configs, engine
    from oracle import oracle as O
    steps = int(os.environ.get("CS3D_LONG_STEPS", "10000"))
    cfg = configs.plain_small(seed=1, steps=steps)
    p, n, mc = cfg["params"], cfg["n"], cfg["max_cells"]
    print("C1 long run: 8 cells, %d steps, division at 2 V0" % steps, flush=True)
    sim = O.Sim(p, oracle_mesh, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, budget=1 << 16)
    sim.init()
    ctx = engine.Context(oracle_mesh, capacity_cells=mc + 8, voxel_budget=1 << 24)
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, voxels=1 << 16)
    worst = 0.0
    first_diff = None
    done = 0
    while done < steps:
        k = min(100, steps - done)
        st_o = sim.step(k)
        s.step(k)
        ctx.sync()
        done += k
        n_g, step_g, status = s.info()
        assert n_g == sim.n, "cell counts differ at step %d: %d vs %d" % (done, n_g, sim.n)
        assert step_g == sim.stepno == done
        assert (status & abi.SYS_FATAL_MASK) == (st_o & abi.SYS_FATAL_MASK)
        nn = n_g * 192
        exact = True
        for f in (abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX, abi.F_VY, abi.F_VZ, abi.F_PRESSURE):
            k_ = nn if f in abi.NODE_FIELDS else n_g
            a, b = s.get(f)[:k_], sim.get(f)[:k_]
            e = elemerr(a, b)
            worst = max(worst, e)
            exact = exact and bool(np.array_equal(a, b))
            assert e <= 1e-5, "field %d per-element rel err %g at step %d" % (f, e, done)
        if not exact and first_diff is None:
            first_diff = done
        assert exact, "fields differ (per element within %g) at step %d" % (worst, done)
        if done % 1000 == 0 or done == steps:
            print("step %d: %d cells, worst per-element rel err so far %.3g, bit-exact so far: %s" %
                  (done, n_g, worst, "yes" if first_diff is None else "no (first differing check at step %d)"
                   % first_diff), flush=True)
        if status & abi.SYS_FATAL_MASK:
            break
    ctx.close()
