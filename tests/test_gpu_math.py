"""The kernels' own arithmetic functions on the device, through the C ABI's
self-check entry points:
  * the correctly rounded float square root of the kernels (sqrt_cr: hardware
    estimate plus one correction step) against sqrtf over all 2^32 inputs;
  * the angle force's double acos (cellsim3d_amd/csrc/cs3d_acos.h, propagate.cu:36)
    against the oracle's, which is the same source compiled for the host, bit for
    bit on 4M arguments (uniform on [-1, 1], dense near -1/2 where the hexagon
    angles of the mesh lie, near -0.3 and +-1, and the special values)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(oracle_mesh):
    from cellsim3d_amd import engine
    c = engine.Context(oracle_mesh, capacity_cells=1, voxel_budget=1 << 10)
    yield c
    c.close()


def test_sqrt_matches_sqrtf_over_every_float(ctx):
    print("sqrt_cr against sqrtf over all 2^32 floats", flush=True)
    assert ctx.check_sqrt() == 0


def test_device_acos_equals_the_oracles(ctx):
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    n = 1 << 20
    x = np.concatenate([rng.uniform(-1, 1, n), -0.5 + 0.05 * rng.uniform(-1, 1, n),
                        -0.3 + 0.2 * rng.uniform(-1, 1, n),
                        np.where(rng.uniform(size=n) < 0.5, 1.0, -1.0) * (1 - 1e-3 * rng.uniform(0, 1, n)),
                        [0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 1.0000000000000002, -1.0000000000000002, np.nan]])
    print("device acos against the oracle's on %d arguments" % x.size, flush=True)
    dev = ctx.device_acos(x)
    host = O.acos(x)
    same = (dev.view(np.uint64) == host.view(np.uint64)) | (np.isnan(dev) & np.isnan(host))
    assert same.all(), (int((~same).sum()), x[~same][:5])
