"""Seeded synthetic workloads (initial states and parameters).

  plain_small     8 unit-radius C180 cells on a 2x2x2 lattice, spacing 2.2,
                  +-0.01 per-coordinate jitter, dt 1e-4, division at 2*V0
  single_tissue   2000 cells, volume fraction 0.40 (jittered simple-cubic
                  lattice, random subset of sites), division off
  colony          1 cell growing by division, per-cell growth factor U[0.8,1.2]
  ensemble        many independent tissues, N ~ log-U[1, 512], parameters +-30 %
  serving         Poisson arrivals, N ~ log-U[8, 4096], horizons log-U[1e3, 1e5]

Random draws use numpy's PCG64 (np.random.default_rng(seed)).  Positions are
float32 SoA arrays with a 192-slot stride (padding slots 0), the reference's
layout (GPUbounce.cu:365-373).
"""
import numpy as np

from . import abi
from . import mesh as meshmod

_MESH_CACHE = {}


def unit_mesh():
    if "gp30" not in _MESH_CACHE:
        _MESH_CACHE["gp30"] = meshmod.gp30_mesh(1.0)
    return _MESH_CACHE["gp30"]


# volume of the reference's own C180 mesh (radius ~0.58); inp.json's division_Vol
# of 2.9 is 3.69 times it (inp.json:14, GPUbounce.cu:1656 uses 0.786)
REF_C180_VOLUME = 0.78665
REF_DIVISION_VOL = 2.9


def division_vol_unit(v0):
    """inp.json's division volume carried over to a mesh of volume v0."""
    return REF_DIVISION_VOL / REF_C180_VOLUME * v0


def mesh_volume(m):
    """Polyhedral volume of a centred mesh in float64 (signed fan decomposition
    about each face centroid; independent of the kernel formula's abs value)."""
    x = np.asarray(m["x"], dtype=np.float64)
    x = x - x.mean(0)
    vol = 0.0
    for ring in list(m["pent"]) + list(m["hexa"]):
        a = x[ring].mean(0)
        for i in range(len(ring)):
            p0, p1 = x[ring[i]], x[ring[(i + 1) % len(ring)]]
            vol += abs(np.dot(p0, np.cross(p1, a))) / 6.0
    return vol


def place(m, centres, jitter=0.0, rng=None):
    """Copies the centred mesh to each centre. Returns X, Y, Z (n*192 float32)."""
    base = meshmod.centred(m)
    n = len(centres)
    X = np.zeros((n, abi.SLOTS), dtype=np.float32)
    Y = np.zeros_like(X)
    Z = np.zeros_like(X)
    c = np.asarray(centres, dtype=np.float32)
    X[:, :180] = base[None, :, 0] + c[:, None, 0]
    Y[:, :180] = base[None, :, 1] + c[:, None, 1]
    Z[:, :180] = base[None, :, 2] + c[:, None, 2]
    if jitter > 0:
        rng = rng or np.random.default_rng(0)
        for A in (X, Y, Z):
            A[:, :180] += rng.uniform(-jitter, jitter, size=(n, 180)).astype(np.float32)
    return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)


def rest_positions(m):
    """Positions of one placed cell used for R0/theta0 (cell at the origin)."""
    b = meshmod.centred(m)
    return b[:, 0].copy(), b[:, 1].copy(), b[:, 2].copy()


def base_params(seed=1):
    p = abi.default_params()
    p.use_rigid_box = 0
    p.ranmar_ij = seed
    p.ranmar_kl = 100 * seed
    return p


def plain_small(seed=1, steps=10000):
    m = unit_mesh()
    v0 = mesh_volume(m)
    rng = np.random.default_rng(seed)
    centres = [(2.2 * i, 2.2 * j, 2.2 * k) for i in range(2) for j in range(2) for k in range(2)]
    X, Y, Z = place(m, centres, 0.01, rng)
    p = base_params(seed)
    p.division_vol = 2.0 * v0
    p.div_time_steps = steps
    p.non_div_time_steps = 0
    # the population grows to a few hundred cells; dense voxels exceed the
    # reference's 32-entry lists (silently overflowing there), so lists hold 64
    p.voxel_capacity = 64
    return {"mesh": m, "params": p, "X": X, "Y": Y, "Z": Z, "n": 8, "steps": steps, "v0": v0,
            "max_cells": 1024}


def lattice_tissue(n_cells, phi, seed=1, jitter=0.02):
    """Random subset of a simple-cubic lattice with the side chosen so that the
    mesh volume fraction is phi; centres jittered uniformly by +-jitter."""
    m = unit_mesh()
    v0 = mesh_volume(m)
    side = (n_cells * v0 / phi) ** (1.0 / 3.0)
    k = int(np.ceil(n_cells ** (1.0 / 3.0)))
    spacing = side / k
    rng = np.random.default_rng(seed)
    sites = np.array([(i, j, l) for i in range(k) for j in range(k) for l in range(k)], dtype=np.float64)
    pick = np.sort(rng.choice(len(sites), size=n_cells, replace=False))
    centres = (sites[pick] + 0.5) * spacing
    centres += rng.uniform(-jitter, jitter, size=centres.shape)
    X, Y, Z = place(m, centres)
    return m, X, Y, Z, {"side": side, "spacing": spacing, "v0": v0}


def single_tissue(n_cells=2000, phi=0.40, seed=1, grid_interval=10, steps=1000):
    m, X, Y, Z, info = lattice_tissue(n_cells, phi, seed)
    p = base_params(seed)
    p.growth_rate = 0.0  # division off: volumes/division are skipped when rGrowth = 0 (GPUbounce.cu:1168)
    p.division_vol = 2.0 * info["v0"]
    p.div_time_steps = steps
    p.non_div_time_steps = 0
    p.grid_interval = grid_interval
    return {"mesh": m, "params": p, "X": X, "Y": Y, "Z": Z, "n": n_cells, "steps": steps, **info}


def colony(seed=1, max_cells=20000):
    m = unit_mesh()
    v0 = mesh_volume(m)
    X, Y, Z = place(m, [(0.0, 0.0, 0.0)])
    p = base_params(seed)
    p.division_vol = 2.0 * v0
    p.div_time_steps = 10 ** 8
    p.non_div_time_steps = 0
    rng = np.random.default_rng(seed)
    scale = rng.uniform(0.8, 1.2, size=max_cells).astype(np.float32)
    return {"mesh": m, "params": p, "X": X, "Y": Y, "Z": Z, "n": 1, "max_cells": max_cells,
            "growth_scale": scale, "v0": v0}


def ensemble(n_tissues=256, seed=1, nmin=1, nmax=512, phi=0.30, steps=5000):
    """Independent tissues: sizes log-uniform in [nmin, nmax]; stiffness, growth
    rate and adhesion strength each scaled by U[0.7, 1.3]; division at inp.json's
    division volume carried over to the unit-radius mesh."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_tissues))).astype(int)
    sizes = np.clip(sizes, nmin, nmax)
    out = []
    m = unit_mesh()
    v0 = mesh_volume(m)
    for t, n in enumerate(sizes):
        _, X, Y, Z, info = lattice_tissue(int(n), phi, seed=seed * 1000 + t)
        p = base_params(1 + (seed * 7 + t) % 300)
        f = rng.uniform(0.7, 1.3, size=3)
        # bond stiffness = stiffFactor1 * Youngs_mod (GPUbounce.cu:1953); the contact
        # forces keep the global Youngs_mod
        p.stiffness1 = np.float32(p.stiffness1 * f[0])
        p.growth_rate = np.float32(p.growth_rate * f[1])
        p.attraction_strength = np.float32(p.attraction_strength * f[2])
        p.division_vol = division_vol_unit(v0)
        # daughters double the local density; lists of 64 keep dense voxels within
        # capacity (the reference's 32-entry lists would silently overflow)
        p.voxel_capacity = 64
        p.div_time_steps = steps
        p.non_div_time_steps = 0
        out.append({"params": p, "X": X, "Y": Y, "Z": Z, "n": int(n), "steps": steps})
    return m, out


def serving_requests(n_requests=64, seed=1, rate=20.0, nmin=8, nmax=4096, hmin=1e3, hmax=1e5, phi=0.30):
    """Poisson arrivals with log-uniform cell counts and horizons."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_requests))).astype(int)
    hor = np.rint(np.exp(rng.uniform(np.log(hmin), np.log(hmax), size=n_requests))).astype(int)
    return [{"arrival": float(a), "n": int(n), "horizon": int(h), "seed": seed * 100000 + i}
            for i, (a, n, h) in enumerate(zip(t, sizes, hor))]
