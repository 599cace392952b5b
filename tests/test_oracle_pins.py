"""Pins of the CPU oracle (CPU only).

The oracle restates the reference kernels; these tests check it against what
fixes each part independently of the restatement itself:
  * the RANMAR known-answer test of marsaglia.h:68-74 and the reference's own
    marsaglia.h / RandomVector.h compiled into oracle/_ref (bit-exact);
  * the reference's mesh files (topology, volume/area magic numbers);
  * energies whose finite-difference gradients must equal the bonded forces
    (springs: U = sum 1/2 (E/R0)(R-R0)^2; angles: U = sum 1/2 k (theta-theta0)^2);
  * closed forms and invariants (pressure magnitude, momentum conservation of
    contact forces, volume scaling, kinematics of one node);
  * brute force on small inputs (all-pairs contact and friction sums, voxel
    membership from box overlap, min/max boxes).
A plausible slip in the oracle (wrong neighbour index or angle mapping, a sign,
a dropped term, an inclusive/exclusive cut) fails at least one of them.
"""
import ctypes as C
import os

import numpy as np
import pytest

from cellsim3d_amd import abi, configs
from cellsim3d_amd import mesh as M
from conftest import REF_ROOT, ref_available
import oracle_kernels as K
from oracle import oracle as O

S = abi.SLOTS
KAT = [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0, 10633180.0]


# ------------------------------------------------------------------ RANMAR
def test_ranmar_known_answer():
    """marsaglia.h:68-74: seeds 1802/9373, 20000 draws, then six values x 2^24."""
    r = O.Ranmar(1802, 9373)
    r.draw(20000)
    assert list(r.draw(6).astype(np.float64) * 4096 * 4096) == KAT


def test_ranmar_seed_range():
    with pytest.raises(ValueError):
        O.Ranmar(1, 30082)  # kl > 30081 is rejected (marsaglia.h:79-83)
    with pytest.raises(ValueError):
        O.Ranmar(31329, 100)  # ij > 31328


@pytest.mark.parametrize("seeds", [(1, 100), (7, 700), (300, 30000), (1802, 9373)])
def test_ranmar_and_random_vectors_match_reference_code(seeds):
    """The oracle's RANMAR, GetRandomVector and GetRandomVectorBasis against the
    reference's own marsaglia.h / RandomVector.h compiled from its tree.  The
    library is loaded here, not at collection, so that collecting the test
    files never loads it."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    mine = O.Ranmar(*seeds)
    R.ref_rmarin(*seeds)
    a = mine.draw(5000)
    b = np.zeros(5000, dtype=np.float32)
    R.ref_ranmar(b, 5000)
    assert np.array_equal(a, b)
    for _ in range(200):
        v = np.zeros(3, dtype=np.float32)
        R.ref_random_vector(v)
        assert np.array_equal(mine.vector(), v)
    for basis in ([0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.267261, 0.534522, 0.801784]):
        bz = np.asarray(basis, dtype=np.float32)
        for _ in range(100):
            v = np.zeros(3, dtype=np.float32)
            R.ref_random_vector_basis(v, bz)
            assert np.array_equal(mine.vector_basis(bz), v)


def test_random_vector_is_unit():
    r = O.Ranmar(5, 500)
    vs = np.array([r.vector() for _ in range(2000)], dtype=np.float64)
    assert np.max(np.abs(np.linalg.norm(vs, axis=1) - 1.0)) < 1e-6
    assert np.all(np.abs(vs.mean(0)) < 0.06)  # uniform on the sphere


# ------------------------------------------------------------------ mesh
def _check_topology(m):
    nn, pent, hexa = m["nn"], m["pent"], m["hexa"]
    assert nn.shape == (180, 3) and nn.min() >= 0 and nn.max() < 180
    for i in range(180):
        assert len(set(nn[i])) == 3 and i not in nn[i]
        for j in nn[i]:
            assert i in nn[j], "bond %d-%d is one-sided" % (i, j)
    bonds = {tuple(sorted((i, int(j)))) for i in range(180) for j in nn[i]}
    assert len(bonds) == 270
    assert len(pent) == 12 and len(hexa) == 80
    seen = np.zeros(180, dtype=int)
    for ring in list(pent) + list(hexa):
        for a, b in zip(ring, np.roll(ring, -1)):
            assert (min(a, b), max(a, b)) in bonds, "face edge %d-%d is not a bond" % (a, b)
        seen[ring] += 1
    assert np.all(seen == 3)  # every node is on three faces
    assert 180 - 270 + 92 == 2
    x = np.asarray(m["x"], dtype=np.float64)
    x = x - x.mean(0)
    for i in range(180):
        n1, n2, n3 = nn[i]
        t = np.dot(np.cross(x[n2] - x[n1], x[n3] - x[n1]), x[i]) * m["sign"][i]
        assert t > 0, "sign[%d] does not point outward" % i


def test_generated_mesh_topology(gp30):
    _check_topology(gp30)
    r = np.linalg.norm(gp30["x"], axis=1)
    assert np.allclose(r, 1.0, atol=1e-12)


@pytest.mark.skipif(not ref_available(), reason="reference mesh files absent")
def test_reference_mesh_files_follow_the_same_conventions():
    m = M.read_reference_files(REF_ROOT)
    _check_topology(m)
    assert int(np.sum(m["sign"] == -1)) == 84 and int(np.sum(m["sign"] == 1)) == 96


def test_mesh_file_round_trip(gp30, tmp_path):
    M.write_reference_files(str(tmp_path), gp30)
    back = M.read_reference_files(str(tmp_path))
    for k in ("nn", "sign", "pent", "hexa", "bonds"):
        assert np.array_equal(np.asarray(back[k]), np.asarray(gp30[k])), k
    assert np.allclose(back["x"], gp30["x"], atol=1e-8)


def _cell(mesh_dict, centre=(0.0, 0.0, 0.0)):
    X, Y, Z = configs.place(mesh_dict, [centre])
    return X, Y, Z


def _signed_volume_area(x, pent, hexa):
    """Divergence-theorem volume and fan area of the closed surface, float64."""
    vol = 0.0
    area = 0.0
    c = x.mean(0)
    for ring in list(pent) + list(hexa):
        a = x[ring].mean(0)
        for i in range(len(ring)):
            p0, p1 = x[ring[i]], x[ring[(i + 1) % len(ring)]]
            vol += np.linalg.det(np.stack([p0 - c, p1 - c, a - c])) / 6.0
            area += 0.5 * np.linalg.norm(np.cross(p0 - a, p1 - a))
    return abs(vol), area


@pytest.mark.skipif(not ref_available(), reason="reference mesh files absent")
def test_volume_of_reference_mesh():
    """V0 = 0.78665, A0 = 4.16814 and psi0 = 0.98870 for the reference's C180
    (its own magic constant 0.786 at GPUbounce.cu:1656 agrees)."""
    m = M.read_reference_files(REF_ROOT)
    nn, sign, faces = M.tables(m)
    p = configs.base_params()
    X, Y, Z = _cell(m)
    st, vol, area, _ = K.volumes(1, p, faces, X, Y, Z)
    assert st == 0
    assert abs(vol[0] - 0.78665) < 2e-5
    assert abs(area[0] - 4.16814) < 2e-5
    psi = 4.835975862049408 * np.cbrt(vol[0]) ** 2 / area[0]
    assert abs(psi - 0.98870) < 2e-5


def test_volume_area_against_divergence_theorem(gp30):
    nn, sign, faces = M.tables(gp30)
    p = configs.base_params()
    rng = np.random.default_rng(4)
    for scale, jit in ((1.0, 0.0), (1.0, 0.02), (2.5, 0.03), (0.6, 0.01)):
        m = dict(gp30)
        m["x"] = gp30["x"] * scale
        X, Y, Z = configs.place(m, [(3.0, -1.0, 7.5)], jitter=jit, rng=rng)
        st, vol, area, _ = K.volumes(1, p, faces, X, Y, Z)
        assert st == 0
        x = np.stack([X[:180], Y[:180], Z[:180]], 1).astype(np.float64)
        v_ref, a_ref = _signed_volume_area(x, gp30["pent"], gp30["hexa"])
        assert abs(vol[0] - v_ref) / v_ref < 2e-6
        assert abs(area[0] - a_ref) / a_ref < 2e-6
    # the generated unit cell sits inside its circumscribed sphere
    X, Y, Z = _cell(gp30)
    _, vol, area, _ = K.volumes(1, p, faces, X, Y, Z)
    assert 0.93 * 4.18879 < vol[0] < 4.18879 and area[0] < 4 * np.pi
    assert abs(vol[0] - configs.mesh_volume(gp30)) / vol[0] < 1e-6


def test_volume_scaling_and_division_flag(gp30):
    nn, sign, faces = M.tables(gp30)
    p = configs.base_params()
    X, Y, Z = _cell(gp30, (1.0, 2.0, 3.0))
    _, v1, a1, _ = K.volumes(1, p, faces, X, Y, Z)
    for s in (0.5, 1.7):
        cx, cy, cz = (np.float32(1.0), np.float32(2.0), np.float32(3.0))
        Xs = ((X - cx) * s + cx).astype(np.float32)
        Ys = ((Y - cy) * s + cy).astype(np.float32)
        Zs = ((Z - cz) * s + cz).astype(np.float32)
        for A in (Xs, Ys, Zs):
            A[180:] = 0
        _, v2, a2, _ = K.volumes(1, p, faces, Xs, Ys, Zs)
        assert abs(v2[0] / v1[0] - s ** 3) < 1e-5 * s ** 3
        assert abs(a2[0] / a1[0] - s ** 2) < 1e-5 * s ** 2
    # flag: V > division_vol, cleared when |1 - psi| > 0.05 under check_sphericity
    p.division_vol = float(v1[0]) * 0.99
    assert K.volumes(1, p, faces, X, Y, Z)[3][0] == 1
    p.division_vol = float(v1[0]) * 1.01
    assert K.volumes(1, p, faces, X, Y, Z)[3][0] == 0
    Zf = Z.copy()
    Zf[:180] = (Z[:180] - 3.0) * 0.4 + 3.0  # flattened: psi drops well below 0.95
    p.division_vol = 0.1
    assert K.volumes(1, p, faces, X, Y, Zf)[3][0] == 0
    p.check_sphericity = 0
    assert K.volumes(1, p, faces, X, Y, Zf)[3][0] == 1


# ------------------------------------------------------------------ CoM, boxes, grid
def _tissue(n=8, seed=2, spacing=2.2, jitter=0.05):
    m = configs.unit_mesh()
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    cs = [(i * spacing, j * spacing, k * spacing) for i in range(side) for j in range(side) for k in range(side)][:n]
    X, Y, Z = configs.place(m, cs, jitter=jitter, rng=rng)
    return X, Y, Z


def test_com_is_the_node_mean():
    X, Y, Z = _tissue(8)
    cm = K.com(8, X, Y, Z)
    for a, A in enumerate((X, Y, Z)):
        ref = A.reshape(8, S)[:, :180].astype(np.float64).mean(1)
        assert np.max(np.abs(cm[a] - ref)) < 4e-6 * max(1.0, np.max(np.abs(ref)))
    Xc = np.full(S, 1.2345678, dtype=np.float32)
    Xc[180:] = 0
    assert abs(K.com(1, Xc, Xc, Xc)[0][0] - np.float32(1.2345678)) <= 2 * np.spacing(np.float32(1.2345678))


def test_bboxes_and_extent_are_exact_min_max():
    X, Y, Z = _tissue(8)
    bb = K.bboxes(8, X, Y, Z).reshape(8, 6)
    for a, A in enumerate((X, Y, Z)):
        v = A.reshape(8, S)[:, :180]
        assert np.array_equal(bb[:, 2 * a], v.min(1))
        assert np.array_equal(bb[:, 2 * a + 1], v.max(1))
    e = K.extent(8, bb.reshape(-1))
    assert np.array_equal(e[0::2], bb[:, 0::2].min(0)) and np.array_equal(e[1::2], bb[:, 1::2].max(0))
    # every node index is the extreme of some cell on every side
    n = 180
    rng = np.random.default_rng(3)
    A = [rng.uniform(-1, 1, (n, S)).astype(np.float32) for _ in range(3)]
    for c in range(n):
        A[0][c, c] = 5.0
        A[1][c, c] = -5.0
        A[2][c, (c + 90) % 180] = 5.0
    bb = K.bboxes(n, *(a.reshape(-1) for a in A)).reshape(n, 6)
    for a in range(3):
        assert np.array_equal(bb[:, 2 * a], A[a][:, :180].min(1))
        assert np.array_equal(bb[:, 2 * a + 1], A[a][:, :180].max(1))


def test_grid_membership_by_box_overlap():
    """b is listed in voxel v exactly when v's cube overlaps b's bbox widened by
    attraction_range (grid edges clamp); lists are ascending."""
    n = 27
    X, Y, Z = _tissue(n, seed=5, jitter=0.3)
    p = configs.base_params()
    p.voxel_capacity = 64
    bb = K.bboxes(n, X, Y, Z)
    st, g, counts, lists = K.grid(n, p, bb)
    assert st == 0
    gmin = g[:3].view(np.float32).astype(np.float64)
    dims = g[3:6]
    e = K.extent(n, bb)
    DL, ar = 2.9, 0.3
    for a in range(3):
        assert dims[a] == int(np.float32((e[2 * a + 1] - e[2 * a]) / np.float32(DL)) + 1)
    xd, yd, zd = dims
    B = bb.reshape(n, 6).astype(np.float64)
    for v in range(xd * yd * zd):
        i, j, k = v % xd, (v // xd) % yd, v // (xd * yd)
        idx = (i, j, k)
        want = []
        for b in range(n):
            ok = True
            for a in range(3):
                lo, hi = B[b, 2 * a] - ar, B[b, 2 * a + 1] + ar
                c_lo = gmin[a] + idx[a] * DL
                c_hi = c_lo + DL
                if idx[a] == 0:
                    c_lo = -np.inf
                if idx[a] == dims[a] - 1:
                    c_hi = np.inf
                ok &= (hi >= c_lo) and (lo < c_hi)
            if ok:
                want.append(b)
        got = list(lists[v * 64:v * 64 + counts[v]])
        assert got == want, "voxel %d: %s vs %s" % (v, got, want)


# ------------------------------------------------------------------ bonded forces
def _params_bonded(E=1000.0, P=0.0, angles=False, k=1000.0):
    p = configs.base_params()
    p.stiffness1 = E
    p.min_pressure = P
    p.constrain_angles = 1 if angles else 0
    p.angle_k = k
    return p


def _perturbed(m, amp, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = configs.place(m, [(0.0, 0.0, 0.0)])
    for A in (X, Y, Z):
        A[:180] += rng.uniform(-amp, amp, 180).astype(np.float32)
    return X, Y, Z


def _fd_grad(U, x, h):
    g = np.zeros_like(x)
    for i in range(x.shape[0]):
        for a in range(3):
            xp = x.copy()
            xp[i, a] += h
            xm = x.copy()
            xm[i, a] -= h
            g[i, a] = (U(xp) - U(xm)) / (2 * h)
    return g


def test_spring_force_zero_at_rest(gp30, oracle_mesh):
    p = _params_bonded()
    X, Y, Z = configs.place(gp30, [(0.0, 0.0, 0.0)])
    st, F = K.con_force(1, p, oracle_mesh, X, Y, Z, bonded_only=True)
    assert st == 0
    for A in F:
        assert np.all(A == 0.0)


def test_spring_force_is_the_energy_gradient(gp30, oracle_mesh):
    """F_i = -dU/dx_i with U = sum over bonds of (E/2 R0)(R - R0)^2 (per-cell E)."""
    E = 700.0
    p = _params_bonded(E=E)
    X, Y, Z = _perturbed(gp30, 0.02, 11)
    st, F = K.con_force(1, p, oracle_mesh, X, Y, Z, bonded_only=True)
    assert st == 0
    R0 = np.ctypeslib.as_array(oracle_mesh.R0).astype(np.float64)
    nn = np.ctypeslib.as_array(oracle_mesh.nn)
    bonds = [(i, int(nn[k * S + i]), R0[k * S + i]) for i in range(180) for k in range(3) if nn[k * S + i] > i]
    bi = np.array([b[0] for b in bonds])
    bj = np.array([b[1] for b in bonds])
    r0 = np.array([b[2] for b in bonds])

    def U(x):
        R = np.linalg.norm(x[bj] - x[bi], axis=1)
        return np.sum(0.5 * (E / r0) * (R - r0) ** 2)

    x = np.stack([X[:180], Y[:180], Z[:180]], 1).astype(np.float64)
    g = _fd_grad(U, x, 1e-6)
    f = np.stack([F[0][:180], F[1][:180], F[2][:180]], 1)
    scale = np.max(np.abs(g))
    assert scale > 1.0
    assert np.max(np.abs(f + g)) < 2e-4 * scale
    assert np.max(np.abs(f.sum(0))) < 1e-4 * scale  # internal forces sum to zero


def test_pressure_force_is_radial_with_magnitude_3P(gp30, oracle_mesh):
    p = _params_bonded(E=0.0, P=0.0)
    X, Y, Z = _perturbed(gp30, 0.03, 3)
    CM = [np.array([0.01], np.float32), np.array([-0.02], np.float32), np.array([0.005], np.float32)]
    Pr = np.array([57.5], np.float32)
    st, F = K.con_force(1, p, oracle_mesh, X, Y, Z, CM=CM, P=Pr, bonded_only=True)
    f = np.stack([F[0][:180], F[1][:180], F[2][:180]], 1).astype(np.float64)
    r = np.stack([X[:180] - CM[0][0], Y[:180] - CM[1][0], Z[:180] - CM[2][0]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(f, axis=1), 3 * 57.5, rtol=1e-6)
    cosang = np.sum(f * r, 1) / (np.linalg.norm(f, axis=1) * np.linalg.norm(r, axis=1))
    assert np.min(cosang) > 1 - 1e-6  # along x - CM of the supplied centre, outward


def _angle_triples(nn):
    """(vertex, arm a, arm b, theta0 slot) for every angle of the mesh: slot 0 is
    between neighbours 0 and 1, slot 1 between 1 and 2, slot 2 between 0 and 2."""
    out = []
    for i in range(180):
        n0, n1, n2 = (int(nn[k * S + i]) for k in range(3))
        out += [(i, n0, n1, 3 * i + 0), (i, n1, n2, 3 * i + 1), (i, n0, n2, 3 * i + 2)]
    return np.array(out)


def _big_cell_mesh(radius=50.0):
    big = configs.meshmod.gp30_mesh(radius)
    return big, O.build_mesh(M.tables(big), configs.rest_positions(big))


def test_angle_force_is_the_bending_energy_gradient():
    """F = -grad sum over angles of k/2 (theta - theta0)^2 (propagate.cu:26-181),
    checked on a radius-50 cell where the 1e-3 regularisers are negligible."""
    big, mesh = _big_cell_mesh()
    k = 1000.0
    p = _params_bonded(E=0.0, angles=True, k=k)
    X, Y, Z = _perturbed(big, 0.3, 21)
    st, F = K.con_force(1, p, mesh, X, Y, Z, bonded_only=True, budget=1 << 17)
    assert st == 0
    nn = np.ctypeslib.as_array(mesh.nn)
    th0 = np.ctypeslib.as_array(mesh.theta0).astype(np.float64)
    T = _angle_triples(nn)

    def U(x):
        a = x[T[:, 1]] - x[T[:, 0]]
        b = x[T[:, 2]] - x[T[:, 0]]
        c = np.sum(a * b, 1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
        th = np.arccos(np.clip(c, -1, 1))
        return np.sum(0.5 * k * (th - th0[T[:, 3]]) ** 2)

    x = np.stack([X[:180], Y[:180], Z[:180]], 1).astype(np.float64)
    g = _fd_grad(U, x, 1e-5)
    f = np.stack([F[0][:180], F[1][:180], F[2][:180]], 1)
    scale = np.max(np.abs(g))
    assert scale > 0.1
    assert np.max(np.abs(f + g)) < 3e-3 * scale
    # a permuted theta0 assignment would not match
    th_bad = th0.copy()
    th_bad[0::3], th_bad[1::3] = th0[1::3], th0[0::3]

    def U_bad(x):
        a = x[T[:, 1]] - x[T[:, 0]]
        b = x[T[:, 2]] - x[T[:, 0]]
        c = np.sum(a * b, 1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
        return np.sum(0.5 * k * (np.arccos(np.clip(c, -1, 1)) - th_bad[T[:, 3]]) ** 2)

    assert np.max(np.abs(f + _fd_grad(U_bad, x, 1e-5))) > 0.05 * scale


def test_angle_force_at_rest_and_net_zero(gp30, oracle_mesh):
    """Rest state: no bending force where the 1e-3 regularisers are negligible
    (radius 50).  At unit radius they are not (|r_i||r_k| ~ 0.05 next to the 1e-3
    in the acos argument, propagate.cu:36), so the reference's rest mesh carries a
    residual bending force; its sum over the cell is still zero."""
    big, mesh = _big_cell_mesh()
    p = _params_bonded(E=0.0, angles=True)
    X, Y, Z = configs.place(big, [(0.0, 0.0, 0.0)])
    _, F = K.con_force(1, p, mesh, X, Y, Z, bonded_only=True, budget=1 << 17)
    assert max(np.max(np.abs(A)) for A in F) < 5e-3
    X, Y, Z = _perturbed(gp30, 0.01, 8)
    _, F = K.con_force(1, p, oracle_mesh, X, Y, Z, bonded_only=True)
    f = np.stack([A[:180] for A in F], 1).astype(np.float64)
    assert np.max(np.abs(f)) > 1.0
    assert np.max(np.abs(f.sum(0))) < 1e-4 * np.max(np.abs(f))


# ------------------------------------------------------------------ contact and friction
def _contact_case(n=8, seed=9):
    X, Y, Z = _tissue(n, seed=seed, spacing=2.06, jitter=0.08)
    return X, Y, Z


def _nodes(n, X, Y, Z):
    x = np.stack([X.reshape(n, S)[:, :180], Y.reshape(n, S)[:, :180], Z.reshape(n, S)[:, :180]], -1)
    return x.astype(np.float64)


def test_contact_forces_brute_force_and_momentum(oracle_mesh):
    """All-pairs sum of the pair law of propagate.cu:406-442 (no grid) equals the
    oracle's grid-based contact force; total contact force is zero."""
    n = 8
    p = configs.base_params()
    p.voxel_capacity = 64
    X, Y, Z = _contact_case(n)
    st, Fall = K.con_force(n, p, oracle_mesh, X, Y, Z)
    assert st & abi.SYS_FATAL_MASK == 0
    _, Fb = K.con_force(n, p, oracle_mesh, X, Y, Z, bonded_only=True)
    fc = np.stack([(Fall[a] - Fb[a]).reshape(n, S)[:, :180] for a in range(3)], -1).astype(np.float64)
    x = _nodes(n, X, Y, Z)
    ar, rr, a_s, r_s, E = 0.3, 0.2, p.attraction_strength, p.repulsion_strength, p.youngs_mod
    ref = np.zeros_like(x)
    pairs = 0
    for c in range(n):
        for b in range(n):
            if b == c:
                continue
            d = x[c][:, None, :] - x[b][None, :, :]
            R = np.linalg.norm(d, axis=-1)
            att = (R < ar) * (-a_s * E * (ar - R) / np.where(R > 0, R, 1))
            rep = (R <= rr) * (r_s * E * (rr - R) / np.where(R > 0, R, 1))
            ref[c] += np.sum((att + rep)[..., None] * d, axis=1)
            pairs += int(np.sum(R < ar))
    assert pairs > 500  # contacts are active
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(fc - ref)) < 2e-4 * scale
    assert np.max(np.abs(fc.reshape(-1, 3).sum(0))) < 1e-3 * scale


def test_contact_pair_law_shape():
    """One node pair (radius-20 cells touching at their extreme nodes, so no other
    pair is within range): attraction a_s E (a_r - R) toward the partner for
    r_r < R < a_r, repulsion r_s E (r_r - R) added at R <= r_r, nothing from a_r on
    (propagate.cu:421-437; the global Youngs_mod, not the per-cell stiffness)."""
    big = configs.meshmod.gp30_mesh(20.0)
    mesh = O.build_mesh(M.tables(big), configs.rest_positions(big))
    p = configs.base_params()
    p.constrain_angles = 0
    p.stiffness1 = 0.0
    p.min_pressure = 0.0
    ar, rr, a_s, r_s, E = 0.3, 0.2, p.attraction_strength, p.repulsion_strength, p.youngs_mod
    base = configs.meshmod.centred(big)
    # node 0 is the unique extreme node along u = x0 / |x0| (all nodes lie on the
    # sphere); the partner is the extreme node of the second cell along -u
    u = big["x"][0] / np.linalg.norm(big["x"][0])
    i = 0
    j = int(np.argmin(big["x"] @ u))
    for dist in (0.25, 0.15, 0.2, 0.31):
        t = (base[i] - base[j] + dist * u).astype(np.float32)
        X, Y, Z = (np.zeros(2 * S, np.float32) for _ in range(3))
        for a, A in enumerate((X, Y, Z)):
            A[:180] = base[:, a]
            A[S:S + 180] = base[:, a] + t[a]
        st, F = K.con_force(2, p, mesh, X, Y, Z)
        assert st & abi.SYS_FATAL_MASK == 0
        d = np.array([X[i] - X[S + j], Y[i] - Y[S + j], Z[i] - Z[S + j]], np.float64)
        R = np.linalg.norm(d)
        mag = (-a_s * E * (ar - R) if R < ar else 0.0) + (r_s * E * (rr - R) if R <= rr else 0.0)
        want = mag / R * d  # force on node i of cell 0
        got = np.array([F[0][i], F[1][i], F[2][i]], np.float64)
        assert np.allclose(got, want, rtol=1e-4, atol=1e-4), (dist, got, want)
        if dist == 0.25:
            assert got @ u > 0  # pulled toward the partner
        if dist == 0.15:
            assert got @ u < 0  # pushed away
        # the partner feels the opposite force
        got_j = np.array([F[0][S + j], F[1][S + j], F[2][S + j]], np.float64)
        assert np.allclose(got_j, -want, rtol=1e-4, atol=1e-4)
        # every other node is force free
        for A in F:
            others = np.delete(A.reshape(2, S)[:, :180].ravel(), [i, 180 + j])
            assert np.max(np.abs(others)) < 1e-6


def test_friction_brute_force_and_rigid_translation(oracle_mesh):
    """CalculateDisForce against an all-pairs sum (propagate.cu:547-658), and a
    rigid translation: only the drag -gamma_visc * v remains."""
    n = 8
    p = configs.base_params()
    p.voxel_capacity = 64
    X, Y, Z = _contact_case(n, seed=13)
    rng = np.random.default_rng(1)
    V = [np.zeros(n * S, np.float32) for _ in range(3)]
    for A in V:
        A.reshape(n, S)[:, :180] = rng.normal(0, 0.1, (n, 180)).astype(np.float32)
    st, F = K.dis_force(n, p, oracle_mesh, X, Y, Z, *V)
    assert st & abi.SYS_FATAL_MASK == 0
    x = _nodes(n, X, Y, Z)
    v = _nodes(n, *V)
    nn = np.ctypeslib.as_array(oracle_mesh.nn).reshape(3, S)[:, :180].T
    sign = np.ctypeslib.as_array(oracle_mesh.sign)[:180]
    gi, ge, go, ar = p.internal_damping, p.viscotic_damping, p.gamma_visc, 0.3
    ref = np.zeros_like(x)
    for c in range(n):
        ref[c] -= gi * np.sum(v[c][:, None, :] - v[c][nn], axis=1)
        nrm = np.cross(x[c][nn[:, 1]] - x[c][nn[:, 0]], x[c][nn[:, 2]] - x[c][nn[:, 0]])
        nrm = sign[:, None] * nrm / np.linalg.norm(nrm, axis=1)[:, None]
        for b in range(n):
            if b == c:
                continue
            d = x[c][:, None, :] - x[b][None, :, :]
            inside = np.sum(d * d, -1) <= ar * ar
            vij = v[c][:, None, :] - v[b][None, :, :]
            vt = vij - np.sum(vij * nrm[:, None, :], -1)[..., None] * nrm[:, None, :]
            ref[c] -= ge * np.sum(inside[..., None] * vt, axis=1)
        ref[c] -= go * v[c]
    f = np.stack([A.reshape(n, S)[:, :180] for A in F], -1)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(f - ref)) < 2e-4 * scale
    # rigid translation
    u = np.array([0.3, -0.2, 0.1], np.float32)
    Vt = [np.zeros(n * S, np.float32) for _ in range(3)]
    for a in range(3):
        Vt[a].reshape(n, S)[:, :180] = u[a]
    _, Ft = K.dis_force(n, p, oracle_mesh, X, Y, Z, *Vt)
    for a in range(3):
        assert np.allclose(Ft[a].reshape(n, S)[:, :180], -go * u[a], rtol=1e-6, atol=1e-7)


# ------------------------------------------------------------------ integration and pressure
def test_integrate_kinematics_of_free_nodes():
    """Half kick then drift on the 180 real nodes: v' = v + dt (Fc + Fd) / 2m,
    x' = x + v' dt; padding slots are left alone."""
    p = configs.base_params()
    rng = np.random.default_rng(6)
    n = 2
    X, Y, Z = (rng.normal(0, 3, n * S).astype(np.float32) for _ in range(3))
    V = [rng.normal(0, 0.5, n * S).astype(np.float32) for _ in range(3)]
    FC = [rng.normal(0, 50, n * S).astype(np.float32) for _ in range(3)]
    FD = [rng.normal(0, 5, n * S).astype(np.float32) for _ in range(3)]
    (X1, Y1, Z1), V1 = K.integrate(n, p, X, Y, Z, V, FC, FD)
    dt, m = 1e-4, 0.04
    real = (np.arange(n * S) % S) < 180

    def r(a):
        return np.asarray(a)[real].astype(np.float64)

    for a, (x0, x1) in enumerate(((X, X1), (Y, Y1), (Z, Z1))):
        v_ref = r(V[a]) + dt * (r(FC[a]) + r(FD[a])) / (2 * m)
        assert np.max(np.abs(r(V1[a]) - v_ref)) <= 2e-7 * np.max(np.abs(v_ref))
        x_ref = r(x0) + r(V1[a]) * dt
        assert np.max(np.abs(r(x1) - x_ref)) <= 1e-6 * np.max(np.abs(x_ref))
        assert np.array_equal(x1[~real], x0[~real]) and np.array_equal(V1[a][~real], V[a][~real])
    Va = K.velocity_ab(n, p, V, FC, FD)
    for a in range(3):
        v_ref = r(V[a]) + dt * r(FC[a]) / (2 * m) + dt * r(FD[a]) / (2 * m)
        assert np.max(np.abs(r(Va[a]) - v_ref)) <= 2e-7 * np.max(np.abs(v_ref))


def test_pressure_update_rules():
    p = configs.base_params()
    P = np.array([50.0, 64.999, 65.0, 70.0], np.float32)
    out = K.pressure_update(4, p, P, np.full(4, p.stiffness1, np.float32), 1, np.float32(0.002))
    assert out[0] == np.float32(50.0) + np.float32(0.002)
    assert out[1] == np.float32(64.999) + np.float32(0.002)  # may overshoot by < inc
    assert out[2] == 65.0 and out[3] == 70.0
    # population model: growth scales with (1 - n / max_pop), never below 0
    p.do_pop_model = 1
    p.max_pop = 100.0
    assert abs(O.lib().or_rgrowth(C.byref(p), 50) - 0.001) < 1e-9
    assert O.lib().or_rgrowth(C.byref(p), 150) == 0.0


# ------------------------------------------------------------------ division
def test_division_half_spaces_and_bookkeeping(gp30, oracle_mesh):
    """postscriptinit.cu:148-256: parent nodes end on the +n side of the plane
    through CM + h n, daughter nodes on the -n side of CM - h n; daughters take
    the next indices in ascending parent order with consecutive RANMAR normals."""
    p = configs.base_params(seed=4)
    v0 = configs.mesh_volume(gp30)
    p.division_vol = 0.5 * v0  # every cell is above the threshold
    p.growth_rate = 0.002
    p.div_time_steps = 1000
    p.voxel_capacity = 64
    centres = [(0.0, 0.0, 0.0), (6.0, 0.0, 0.0), (0.0, 6.0, 0.0)]
    X, Y, Z = configs.place(gp30, centres)
    sim = O.Sim(p, oracle_mesh, X, Y, Z, 3, max_cells=8)
    sim.init()
    sim.step(1)
    assert sim.n == 6
    rng = O.Ranmar(4, 400)
    h = np.float32(0.5 * 0.2)
    X1, Y1, Z1 = sim.get(abi.F_X), sim.get(abi.F_Y), sim.get(abi.F_Z)
    CMx, CMy, CMz = sim.get(abi.F_CMX), sim.get(abi.F_CMY), sim.get(abi.F_CMZ)
    VX = sim.get(abi.F_VX)
    P = sim.get(abi.F_PRESSURE)
    for parent in range(3):
        nrm = rng.vector().astype(np.float64)
        d = 3 + parent
        cm = np.array([CMx[parent], CMy[parent], CMz[parent]], np.float64)
        xp = np.stack([X1[parent * S:parent * S + 180], Y1[parent * S:parent * S + 180],
                       Z1[parent * S:parent * S + 180]], 1).astype(np.float64)
        xd = np.stack([X1[d * S:d * S + 180], Y1[d * S:d * S + 180], Z1[d * S:d * S + 180]], 1).astype(np.float64)
        assert np.min((xp - cm - h * nrm) @ nrm) > -1e-5
        assert np.min((xd - cm + h * nrm) @ (-nrm)) > -1e-5
        assert np.array_equal(VX[d * S:d * S + 180], VX[parent * S:parent * S + 180])
        assert P[parent] == np.float32(p.min_pressure) and P[d] == np.float32(p.min_pressure)
        # the split is a real cut: some nodes were projected onto each plane
        assert np.sum(np.abs((xp - cm - h * nrm) @ nrm) < 1e-5) > 10
        # daughter CM stays stale (0) until the next CenterOfMass (Appendix hazard A5)
        assert CMx[d] == 0 and CMy[d] == 0 and CMz[d] == 0


# ------------------------------------------------------------------ random force
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123 (Salmon et al.,
    SC'11): counter, key -> output."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert [int(v) for v in O.philox(ctr, key)] == out


def test_log_of_the_polar_method():
    import ctypes as C
    s = np.concatenate([np.linspace(1e-6, 1 - 1e-6, 20001), 2.0 ** -np.arange(1, 48)]).astype(np.float32)
    got = np.array([O.lib().or_ln_unit(C.c_float(v)) for v in s], dtype=np.float64)
    ref = np.log(s.astype(np.float64))
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)) < 4e-7


def test_normals_are_standard_normal_and_reproducible():
    from scipy import stats
    z = np.array([O.normals3(7, node, step) for node in range(3000) for step in (1, 2)]).ravel()
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert np.array_equal(O.normals3(7, 5, 11), O.normals3(7, 5, 11))
    assert not np.array_equal(O.normals3(7, 5, 11), O.normals3(7, 5, 12))
    assert not np.array_equal(O.normals3(7, 5, 11), O.normals3(8, 5, 11))


# ------------------------------------------------------------------ correct_com
def test_correct_com_shifts_every_node_by_the_mean_centre(gp30):
    """GPUbounce.cu:1431-1441: at the end of each step the mean of the cell centres
    (computed before the shift) is subtracted from every node.  In double precision:
    the centres recomputed from the shifted nodes plus the system centre give back
    the stored pre-shift centres, the stored centre is their mean, and the shifted
    system is centred at 0."""
    nn, sign, faces = M.tables(gp30)
    from cellsim3d_amd import engine  # host mesh build only (no device)
    mesh = engine.build_mesh((nn, sign, faces), configs.rest_positions(gp30))
    cents = [(4.0, 5.0, 6.0), (9.0, 5.5, 6.0), (4.5, 10.0, 7.0)]
    X, Y, Z = configs.place(gp30, cents)
    n = len(cents)
    p = configs.base_params(3)
    p.correct_com = 1
    p.growth_rate = 0.0
    sim = O.Sim(p, mesh, X, Y, Z, n, max_cells=n, precision=1)
    rng = np.random.default_rng(1)
    for f in (abi.F_VX, abi.F_VY, abi.F_VZ):  # a drift so the centre moves
        v = np.zeros(n * S)
        v.reshape(n, S)[:, :180] = 0.5 + rng.normal(0, 0.05, (n, 180))
        sim.set(f, v.astype(np.float32))
    sim.init()
    for k in range(3):
        sim.step(1)
        x = [sim.get64(f).reshape(n, S)[:, :180] for f in (abi.F_X, abi.F_Y, abi.F_Z)]
        cm = [sim.get64(f)[:n] for f in (abi.F_CMX, abi.F_CMY, abi.F_CMZ)]
        sc = sim.get(abi.F_SYSCM).astype(np.float64)
        for a in range(3):
            fresh = x[a].mean(1)
            assert np.allclose(fresh + sc[a], cm[a], rtol=0, atol=1e-6), (k, a)
            assert abs(cm[a].mean() - sc[a]) < 1e-6 * max(1.0, abs(sc[a]))
            assert abs(fresh.mean()) < 1e-6
        assert np.all(sim.get(abi.F_X).reshape(n, S)[:, 180:] == 0)
        if k == 0:
            assert np.all(np.abs(sc) > 1.0)  # the first step moves the system to the origin


# ------------------------------------------------------------------ angle acos
def test_acos_against_extended_precision():
    """The angle force's double acos (propagate.cu:36; cellsim3d_amd/csrc/cs3d_acos.h,
    shared by the kernels and the oracle) against long double arccos: within 1.2
    double ulps everywhere, the same float after rounding wherever the true value is
    not within 2 double ulps of a float midpoint, exact special values, monotone."""
    rng = np.random.default_rng(5)
    n = 400000
    x = np.concatenate([rng.uniform(-1, 1, n), -0.5 + 0.05 * rng.uniform(-1, 1, n),
                        -0.3 + 0.2 * rng.uniform(-1, 1, n), 1 - 1e-3 * rng.uniform(0, 1, n),
                        -1 + 1e-3 * rng.uniform(0, 1, n), 1e-6 * rng.uniform(-1, 1, n),
                        np.nextafter(0.5, np.array([0.0, 1.0])), np.nextafter(-0.5, np.array([0.0, -1.0]))])
    y = O.acos(x)
    ref = np.arccos(x.astype(np.longdouble))
    r64 = ref.astype(np.float64)
    ulp = np.spacing(np.abs(r64))
    err = np.abs((y.astype(np.longdouble) - ref) / ulp.astype(np.longdouble)).astype(np.float64)
    assert err.max() <= 1.2, (err.max(), x[np.argmax(err)])
    # float rounding: equal unless the true value lies within 2 double ulps of a float midpoint
    f_ours = y.astype(np.float32)
    f_true = ref.astype(np.float32)
    lo = np.minimum(f_ours, f_true).astype(np.float64)
    mid = lo + np.spacing(lo.astype(np.float32)).astype(np.float64) / 2
    near = np.abs(r64 - mid) <= 2 * ulp
    assert np.all((f_ours == f_true) | near)
    sp = np.array([1.0, -1.0, 0.0, -0.0, 1.0000000000000002, -1.0000000000000002, 2.0, np.nan])
    ys = O.acos(sp)
    assert ys[0] == 0.0 and ys[1] == np.pi and ys[2] == np.pi / 2 and ys[3] == np.pi / 2
    assert np.all(np.isnan(ys[4:]))
    g = np.linspace(-1, 1, 200001)
    assert np.all(np.diff(O.acos(g)) <= 0)
