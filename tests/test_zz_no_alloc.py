"""The allocation-free PressureReset (cs3d_k_pressure_reset_noalloc) through the
C ABI against cs3d_k_pressure_reset on identical systems: two calls in a row,
the second with a shorter cell list, leave bit-identical pressures; lists past
the staging array's 65536 cells are refused with INVALID_ARG."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_pressure_reset_matches_the_allocating_one(oracle_mesh):
    from cellsim3d_amd import abi, configs, engine
    L = engine.lib()
    L.cs3d_k_pressure_reset_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_float]
    L.cs3d_k_pressure_reset_noalloc.restype = C.c_int
    m, X, Y, Z, info = configs.lattice_tissue(64, 0.35, seed=3)
    p = configs.base_params(3)
    with engine.Context(oracle_mesh, capacity_cells=160) as ctx:
        a = ctx.add_system(p, X, Y, Z, 64, max_cells=64)
        b = ctx.add_system(p, X, Y, Z, 64, max_cells=64)
        P0 = np.random.default_rng(1).uniform(50, 65, 64).astype(np.float32)
        a.set(abi.F_PRESSURE, P0)
        b.set(abi.F_PRESSURE, P0)
        calls = [(np.arange(0, 64, 3, dtype=np.int32), 51.5), (np.array([5, 7, 60], dtype=np.int32), 49.25)]
        for cells, value in calls:
            print("pressure reset of %d cells to %g" % (len(cells), value), flush=True)
            engine.check(L.cs3d_k_pressure_reset(ctx.h, a.id, cells.ctypes.data_as(C.c_void_p), len(cells),
                                                 C.c_float(value)), "cs3d_k_pressure_reset")
            engine.check(L.cs3d_k_pressure_reset_noalloc(ctx.h, b.id, cells.ctypes.data_as(C.c_void_p),
                                                         len(cells), C.c_float(value)),
                         "cs3d_k_pressure_reset_noalloc")
            pa, pb = a.get(abi.F_PRESSURE)[:64], b.get(abi.F_PRESSURE)[:64]
            assert np.array_equal(pa, pb)
            assert np.all(pb[cells] == np.float32(value))
        big = np.zeros(65537, dtype=np.int32)
        st = L.cs3d_k_pressure_reset_noalloc(ctx.h, b.id, big.ctypes.data_as(C.c_void_p), len(big), C.c_float(1.0))
        assert st == abi.INVALID_ARG
        assert np.array_equal(a.get(abi.F_PRESSURE)[:64], b.get(abi.F_PRESSURE)[:64])
