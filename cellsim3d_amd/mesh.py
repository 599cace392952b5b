"""C180 cell mesh: the engine's own icosahedral Goldberg GP(3,0) generator and
readers/writers for the reference's mesh file formats.

The reference reads four files from the working directory
(src/GPUbounce.cu:1616, 1822, 1840, 1858):
  C180            180 lines "x y z"
  C180NN          180 lines "n1, n2, n3, sign" (1-based neighbours)
  C180C           270 lines "a, b" (1-based bonds)
  C180_pentahexa  12 lines of 5 ints then 80 lines of 6 ints (0-based rings)
Conventions kept by the generator: every face row is a cyclic ring of bonded
nodes, and sign[i] * ((x[n2]-x[n1]) x (x[n3]-x[n1])) points away from the cell
centre (as in the reference's files).

GP(3,0) construction: an icosahedron is subdivided with frequency 3 (92
vertices, 180 triangles); the cell nodes are the triangle centroids projected
onto the unit sphere (the dual polyhedron), which gives 180 nodes of degree 3,
270 bonds, 12 pentagons and 80 hexagons.  Nodes are numbered by recursive
median bisection so that every run of 16 consecutive indices is a compact
patch of the surface: neighbouring lanes of a wavefront then hold nearby
nodes, whose contact candidates lie in the same few search bins.
"""
import os

import numpy as np

from . import abi

PATCH = 16  # nodes per compact patch of the numbering (a quarter of a wavefront)


def _icosahedron():
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    v = []
    for a in (-1.0, 1.0):
        for b in (-phi, phi):
            v += [(0.0, a, b), (a, b, 0.0), (b, 0.0, a)]
    v = np.array(v)
    v /= np.linalg.norm(v, axis=1)[:, None]
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
    edge = d[d > 1e-9].min()
    faces = []
    n = len(v)
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                if abs(d[i, j] - edge) < 1e-9 and abs(d[j, k] - edge) < 1e-9 and abs(d[i, k] - edge) < 1e-9:
                    faces.append((i, j, k))
    assert len(faces) == 20
    return v, faces


def _geodesic(freq=3):
    """Class-I geodesic sphere: returns (vertices (92,3), triangles (180,3))."""
    iv, ifaces = _icosahedron()
    verts, key2id, tris = [], {}, []

    def vid(p):
        p = p / np.linalg.norm(p)
        key = tuple(np.round(p, 9))
        if key not in key2id:
            key2id[key] = len(verts)
            verts.append(p)
        return key2id[key]

    for a, b, c in ifaces:
        A, B, Cc = iv[a], iv[b], iv[c]

        def P(i, j):
            return vid(A + (B - A) * (i / freq) + (Cc - A) * (j / freq))

        for i in range(freq):
            for j in range(freq - i):
                tris.append((P(i, j), P(i + 1, j), P(i, j + 1)))
                if j < freq - 1 - i:
                    tris.append((P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)))
    return np.array(verts), np.array(tris)


def _bisect_order(x, idx, out):
    n = len(idx)
    if n <= PATCH:
        out.extend(idx)
        return
    first = PATCH * max(1, int(round(n / (2.0 * PATCH))))
    first = min(first, n - 1)
    pts = x[idx]
    axis = int(np.argmax(pts.max(0) - pts.min(0)))
    order = np.argsort(pts[:, axis], kind="stable")
    left = [idx[i] for i in order[:first]]
    right = [idx[i] for i in order[first:]]
    _bisect_order(x, left, out)
    _bisect_order(x, right, out)


def gp30_mesh(radius=1.0):
    """Returns a dict with x (180,3) float64 centred at 0, nn (180,3) 0-based,
    sign (180,), pent (12,5), hexa (80,6), bonds (270,2) (0-based, a < b)."""
    gv, tris = _geodesic(3)
    assert len(gv) == 92 and len(tris) == 180
    cen = gv[tris].mean(axis=1)
    cen /= np.linalg.norm(cen, axis=1)[:, None]
    # dual adjacency: triangles sharing an edge
    edge2t = {}
    for t, (a, b, c) in enumerate(tris):
        for e in ((a, b), (b, c), (a, c)):
            edge2t.setdefault(tuple(sorted(e)), []).append(t)
    adj = [[] for _ in range(180)]
    for ts in edge2t.values():
        assert len(ts) == 2
        adj[ts[0]].append(ts[1])
        adj[ts[1]].append(ts[0])
    assert all(len(a) == 3 for a in adj)
    # faces: triangles around each geodesic vertex, in cyclic order
    rings = []
    for v in range(92):
        ts = [t for t in range(180) if v in tris[t]]
        nrm = gv[v]
        ref = cen[ts[0]] - nrm * np.dot(cen[ts[0]], nrm)
        ref /= np.linalg.norm(ref)
        ortho = np.cross(nrm, ref)
        ang = [np.arctan2(np.dot(cen[t], ortho), np.dot(cen[t], ref)) for t in ts]
        rings.append([ts[i] for i in np.argsort(ang)])
    # numbering: compact 16-node runs
    order = []
    _bisect_order(cen, list(range(180)), order)
    new = np.empty(180, dtype=np.int64)
    new[np.array(order)] = np.arange(180)
    x = cen[np.array(order)] * radius
    nn = np.zeros((180, 3), dtype=np.int64)
    for old in range(180):
        nn[new[old]] = sorted(new[a] for a in adj[old])
    sign = np.zeros(180, dtype=np.int64)
    for i in range(180):
        n1, n2, n3 = nn[i]
        t = np.dot(np.cross(x[n2] - x[n1], x[n3] - x[n1]), x[i])
        sign[i] = 1 if t > 0 else -1
    pent = [[int(new[t]) for t in r] for r in rings if len(r) == 5]
    hexa = [[int(new[t]) for t in r] for r in rings if len(r) == 6]
    assert len(pent) == 12 and len(hexa) == 80
    bonds = sorted({tuple(sorted((i, int(j)))) for i in range(180) for j in nn[i]})
    assert len(bonds) == 270
    return {"x": x, "nn": nn, "sign": sign, "pent": np.array(pent), "hexa": np.array(hexa),
            "bonds": np.array(bonds)}


def write_reference_files(directory, mesh):
    """Writes C180, C180NN, C180C and C180_pentahexa in the reference's formats."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "C180"), "w") as f:
        for p in mesh["x"]:
            f.write("%.9f  %.9f  %.9f\n" % tuple(p))
    with open(os.path.join(directory, "C180NN"), "w") as f:
        for (a, b, c), s in zip(mesh["nn"], mesh["sign"]):
            f.write("%4d, %4d, %4d, %2d\n" % (a + 1, b + 1, c + 1, s))
    with open(os.path.join(directory, "C180C"), "w") as f:
        for a, b in mesh["bonds"]:
            f.write("%4d, %4d\n" % (a + 1, b + 1))
    with open(os.path.join(directory, "C180_pentahexa"), "w") as f:
        for r in mesh["pent"]:
            f.write("".join("%4d" % v for v in r) + "\n")
        for r in mesh["hexa"]:
            f.write("".join("%4d" % v for v in r) + "\n")


def read_reference_files(directory):
    """Reads the four mesh files of the reference's format (GPUbounce.cu:1616-1886)."""
    x = np.loadtxt(os.path.join(directory, "C180"))[:180]
    t = np.loadtxt(os.path.join(directory, "C180NN"), delimiter=",", dtype=np.int64)[:180]
    nn, sign = t[:, :3] - 1, t[:, 3]
    bonds = np.loadtxt(os.path.join(directory, "C180C"), delimiter=",", dtype=np.int64)[:270] - 1
    rows = [list(map(int, line.split())) for line in open(os.path.join(directory, "C180_pentahexa")) if line.strip()]
    pent, hexa = np.array(rows[:12]), np.array(rows[12:92])
    return {"x": x, "nn": nn, "sign": sign, "pent": pent, "hexa": hexa, "bonds": bonds}


def tables(mesh):
    """Device-layout tables: nn (3*192, k*192+i), sign (192), faces (92*7) with
    pentagon rows N1..N5,N1,N1 and hexagon rows N1..N6,N1 (GPUbounce.cu:1861-1886)."""
    nn = np.zeros(3 * abi.SLOTS, dtype=np.int32)
    for k in range(3):
        nn[k * abi.SLOTS:k * abi.SLOTS + 180] = mesh["nn"][:, k]
    sign = np.zeros(abi.SLOTS, dtype=np.int32)
    sign[:180] = mesh["sign"]
    faces = np.zeros(abi.FACES * 7, dtype=np.int32)
    for i, r in enumerate(mesh["pent"]):
        faces[7 * i:7 * i + 7] = [r[0], r[1], r[2], r[3], r[4], r[0], r[0]]
    for i, r in enumerate(mesh["hexa"]):
        faces[84 + 7 * i:84 + 7 * i + 7] = [r[0], r[1], r[2], r[3], r[4], r[5], r[0]]
    return nn, sign, faces


def centred(mesh):
    """Node positions with the mean removed, in float32 arithmetic like
    initialize_C180s (GPUbounce.cu:1631-1653)."""
    x = np.asarray(mesh["x"], dtype=np.float32)
    out = np.empty_like(x)
    for a in range(3):
        s = np.float32(0)
        for v in x[:, a]:
            s = np.float32(s + v)
        s = np.float32(np.float64(s) / 180.0)
        out[:, a] = x[:, a] - s
    return out
