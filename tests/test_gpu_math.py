"""The kernels' own arithmetic functions on the device, through the C ABI's
self-check entry points:
  * the correctly rounded float square root of the kernels (sqrt_cr: hardware
    estimate plus one correction step) against sqrtf over all 2^32 inputs;
  * the quotients from a reciprocal (cellsim3d_amd/csrc/cs3d_div.h: hardware
    reciprocal estimate plus one Newton step, then one Markstein correction)
    against IEEE division: the reciprocal over every float in its range, the
    quotient over every pair of significands (2^46 pairs) and over random pairs
    across the exponent range the kernels admit;
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


def test_quotients_from_a_reciprocal_equal_ieee_division(ctx):
    print("reciprocal, 2^46 significand pairs and 2^30 scaled pairs against IEEE division", flush=True)
    rcp_bad, rcp_n, quot_bad, quot_ctl, scaled_bad, scaled_ctl = ctx.check_div()
    print("differing: reciprocal %d of %d, quotient %d (uncorrected %d), scaled %d (uncorrected %d)" %
          (rcp_bad, rcp_n, quot_bad, quot_ctl, scaled_bad, scaled_ctl), flush=True)
    # every float of either sign with exponent in [-40, 39], and +-2^40
    assert rcp_n == 2 * 80 * (1 << 23) + 2
    # the controls show that every pair was compared and that a difference is seen:
    # the uncorrected product misses about a quarter of the quotients
    assert 0.2 * 2 ** 46 < quot_ctl < 0.35 * 2 ** 46
    assert 0.2 * 2 ** 30 < scaled_ctl < 0.35 * 2 ** 30
    assert (rcp_bad, quot_bad, scaled_bad) == (0, 0, 0)


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
