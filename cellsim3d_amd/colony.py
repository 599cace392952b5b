"""Growing colony: one C180 cell grows by division to ~20,000 cells.

Per-cell growth factors U[0.8, 1.2] scale the pressure increment (a
per-cell multiplier of PressureKernels.cu:36-38's inc, stored in the
GROWTH_SCALE field); division at twice the unit cell's volume; FP32 state.
Reports node-updates/s as the colony grows, the share of the step spent in
the division phase, and an FP64 recomputation of the per-cell volumes of the
final state next to the engine's FP32 values.

    python -m cellsim3d_amd.colony [--cells 20000] [--max-steps 40000]
prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, configs, engine
from . import mesh as meshmod


def volumes_f64(m, X, Y, Z, n):
    """Per-cell volume in float64 with the same face decomposition (fan about
    each face centroid, |p0 . (p1 x a)| / 6, volume.cu:37-130)."""
    faces = list(m["pent"]) + list(m["hexa"])
    x = np.stack([X.reshape(-1, 192)[:n, :180], Y.reshape(-1, 192)[:n, :180], Z.reshape(-1, 192)[:n, :180]],
                 -1).astype(np.float64)
    x = x - x.mean(1, keepdims=True)
    vol = np.zeros(n)
    for ring in faces:
        a = x[:, ring].mean(1)
        for i in range(len(ring)):
            p0, p1 = x[:, ring[i]], x[:, ring[(i + 1) % len(ring)]]
            vol += np.abs(np.sum(p0 * np.cross(p1, a), -1))
    return vol / 6.0


def run(cells=20000, max_steps=60000, chunk=500, grid_interval=10, seed=1, log=None, keep_state=False):
    """Grows the colony from one cell until it has `cells` cells (or max_steps
    steps).  Returns the summary dict; with keep_state, also the final state
    (params, mesh, X, Y, Z, n, growth scales) under "state"."""
    log = log or (lambda msg: None)
    cap = int(cells * 1.25) + 64
    cfg = configs.colony(seed=seed, max_cells=cap)
    m = cfg["mesh"]
    p = cfg["params"]
    p.grid_interval = grid_interval
    # a growing colony compresses its core: up to ~100 cells' boxes meet in one voxel
    p.voxel_capacity = 128
    gm = engine.build_mesh(meshmod.tables(m), configs.rest_positions(m))
    ctx = engine.Context(gm, capacity_cells=cap, voxel_budget=(32 * cap + 4096) * (p.voxel_capacity + 1),
                         bin_budget=(640 * cap + 16384) * abi.BIN_WORDS)
    s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], 1, max_cells=cap, voxels=32 * cap + 4096)
    s.set(abi.F_GROWTH_SCALE, cfg["growth_scale"])
    ctx.set_graph_steps(100)
    ctx.prepare_graph(s)
    ctx.sync()
    samples = []
    t_start = time.perf_counter()
    step = 0
    n = 1
    work0 = 0
    while n < cells and step < max_steps:
        t0 = time.perf_counter()
        s.step(chunk)
        ctx.sync()
        dt = time.perf_counter() - t0
        n_new, step, status = s.info()
        if status & abi.SYS_FATAL_MASK:
            vol = s.get(abi.F_VOLUME)[:n_new]
            pc = s.get(abi.F_PAIRCOUNT).reshape(-1, 192)[:n_new, :180]
            raise SystemExit("colony stopped: status 0x%x at step %d with %d cells; cell volume / V0 in [%.3f, %.3f], "
                             "at most %d contact partners per node" % (status, step, n_new, vol.min() / cfg["v0"],
                                                                       vol.max() / cfg["v0"], int(pc.max())))
        work = int(ctx.work(s.id, 1)[0])  # cell-steps counted on the device
        samples.append({"step": int(step), "cells": int(n_new), "ms_per_step": dt / chunk * 1e3,
                        "node_updates_per_s": (work - work0) * 180.0 / dt})
        work0 = work
        n = n_new
        if len(samples) % 10 == 0:
            log("colony: step %d, %d cells, %.3f ms/step" % (step, n, dt / chunk * 1e3))
    wall = time.perf_counter() - t_start
    node_updates = work0 * 180.0
    # phase split at the final size
    ph = s.ctx.step_timed(20, s.id) / 20
    n_fin, step_fin, _ = s.info()
    s.ctx.sync()
    s.k_com()
    s.k_volumes(step_fin)
    v32 = s.get(abi.F_VOLUME)[:n_fin].astype(np.float64)
    X, Y, Z = (s.get(f) for f in (abi.F_X, abi.F_Y, abi.F_Z))
    v64 = volumes_f64(m, X, Y, Z, n_fin)
    rel = float(np.max(np.abs(v32 - v64) / v64))
    last = samples[-1]
    out = {
        "workload": "colony from 1 unit-radius C180 cell, per-cell growth factor U[0.8, 1.2], division at 2 V0, "
                    "FP32, grid rebuilt every %d steps and after divisions" % grid_interval,
        "final_cells": int(n_fin),
        "steps": int(step_fin),
        "wall_s": wall,
        "node_updates": node_updates,
        "node_updates_per_s": node_updates / wall,
        "final_chunk": {"cells": last["cells"], "ms_per_step": last["ms_per_step"],
                        "node_updates_per_s": last["node_updates_per_s"]},
        "phases_ms_per_step_at_final_size": dict(zip(engine.PHASE_NAMES, (round(float(x), 5) for x in ph))),
        "volume_fp32_vs_fp64_max_rel": rel,
        "samples": samples[::max(1, len(samples) // 40)],
    }
    if keep_state:
        out["state"] = {"params": p, "mesh": m, "X": X, "Y": Y, "Z": Z, "n": int(n_fin),
                        "growth_scale": cfg["growth_scale"]}
    ctx.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=20000)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--grid-interval", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    out = run(args.cells, args.max_steps, args.chunk, args.grid_interval, args.seed, log)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
