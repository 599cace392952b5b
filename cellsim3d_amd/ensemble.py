"""Batched ensembles: many independent tissues stepped together on one GPU.

A Context holds every tissue as a system with its own cell slots, voxel and
bin regions, parameters and RANMAR stream; cs3d_step(ALL) advances all of them
with one launch per kernel phase (captured in a HIP graph), so tissues of a
few cells share the launches of the large ones.  The plain path runs the same
tissues one after another through the per-system drop-in path (cs3d_step on
one system), which is what a caller of the reference's single-tissue loop
does.

The reference has no ensemble mode (GPUbounce.cu:1055-1506 steps one tissue);
results of a tissue are the same alone and in a batch, which
tests/test_gpu_serving.py checks bit for bit.

    python -m cellsim3d_amd.ensemble [--tissues 256] [--steps 5000] ...
prints one JSON line (sims/hour and node-updates/s, batched and plain).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, configs, engine
from . import mesh as meshmod

NODE_BYTES = 4 * 192  # one float per slot


def build(n_tissues=256, seed=1, nmin=1, nmax=512, steps=5000, grid_interval=10, phi=0.30, room=2):
    """Seeded ensemble: (device mesh, tissues).  Each tissue dict has params,
    X, Y, Z, n, max_cells = room * n + 2.  room = 2 lets every cell divide once,
    enough for 1000 steps; a tissue of a few free cells doubles about every 1000
    steps, so longer runs need more room (a tissue that reaches max_cells stops
    with the capacity status)."""
    m, tissues = configs.ensemble(n_tissues=n_tissues, seed=seed, nmin=nmin, nmax=nmax, phi=phi, steps=steps)
    gm = engine.build_mesh(meshmod.tables(m), configs.rest_positions(m))
    for t in tissues:
        t["params"].grid_interval = grid_interval
        t["max_cells"] = room * t["n"] + 2
    return gm, tissues


def node_updates(tissues, steps):
    return float(sum(t["n"] for t in tissues)) * 180.0 * steps


def _context(gm, tissues):
    cap = sum(t["max_cells"] for t in tissues)
    vox = sum(max(4096, 32 * t["max_cells"]) * (1 + t["params"].voxel_capacity) for t in tissues)
    bins = sum(min(max(16384, 640 * t["max_cells"]), 1 << 40) for t in tissues) * abi.BIN_WORDS
    return engine.Context(gm, capacity_cells=cap, voxel_budget=vox, bin_budget=bins)


def run_batched(gm, tissues, steps, wg_per_cu=None, graph_steps=None, warmup=0, ctx=None):
    """All tissues in one context, stepped together.  Returns (seconds, ctx, systems)."""
    own = ctx is None
    if own:
        ctx = _context(gm, tissues)
    if wg_per_cu:
        ctx.set_launch(wg_per_cu)
    if graph_steps:
        ctx.set_graph_steps(graph_steps)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    if warmup:
        ctx.step(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    dt = time.perf_counter() - t0
    for s in systems:
        st = s.info()[2]
        if st & abi.SYS_FATAL_MASK:
            raise RuntimeError("system %d status 0x%x" % (s.id, st))
    return dt, ctx, systems


def run_plain(gm, tissues, steps, sample_steps=None, warmup=10):
    """The tissues one after another, each alone on the drop-in path.  With
    sample_steps, each tissue runs that many timed steps and its time is scaled
    to `steps` (stated as extrapolated by the caller).  Returns seconds."""
    cap = max(t["max_cells"] for t in tissues)
    big = max(tissues, key=lambda t: t["max_cells"])
    ctx = _context(gm, [big])
    total = 0.0
    k = steps if sample_steps is None else min(steps, sample_steps)
    for t in tissues:
        s = ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"])
        if warmup:
            s.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        s.step(k)
        ctx.sync()
        total += (time.perf_counter() - t0) * (steps / k)
        s.remove()
    ctx.close()
    del cap
    return total


def autotune(gm, tissues, probe_steps=100, wg_options=(4, 8, 16, 32, 64), graph_options=(10, 25, 50), log=None):
    """Picks the launch cap (workgroups per CU of the cell kernels, i.e. how
    many cells a workgroup walks through) and the steps per captured graph by
    timing short batched runs.  Returns (wg_per_cu, graph_steps, table)."""
    ctx = _context(gm, tissues)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    table = []
    for wg in wg_options:
        for gs in graph_options:
            ctx.set_launch(wg)
            ctx.set_graph_steps(gs)
            ctx.step(gs)  # capture + warm
            ctx.sync()
            k = max(gs, (probe_steps // gs) * gs)
            t0 = time.perf_counter()
            ctx.step(k)
            ctx.sync()
            ms = (time.perf_counter() - t0) / k * 1e3
            table.append({"wg_per_cu": wg, "graph_steps": gs, "ms_per_step": ms})
            if log:
                log("autotune: wg_per_cu %d graph_steps %d -> %.4f ms/step" % (wg, gs, ms))
    for s in systems:
        s.remove()
    ctx.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return best["wg_per_cu"], best["graph_steps"], table


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--tissues", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--room", type=int, default=2,
                    help="cell slots per initial cell of each tissue (2: every cell divides once)")
    ap.add_argument("--nmax", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--grid-interval", type=int, default=10)
    ap.add_argument("--plain-sample-steps", type=int, default=200,
                    help="timed steps per tissue on the plain path (0 = full run)")
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    gm, tissues = build(args.tissues, args.seed, 1, args.nmax, args.steps, args.grid_interval, room=args.room)
    cells = sum(t["n"] for t in tissues)
    log("ensemble: %d tissues, %d cells, %d steps" % (len(tissues), cells, args.steps))
    wg, gs, table = (16, 10, [])
    if not args.no_tune:
        wg, gs, table = autotune(gm, tissues, log=log)
    dt_b, ctx, _ = run_batched(gm, tissues, args.steps, wg_per_cu=wg, graph_steps=gs, warmup=gs)
    ctx.close()
    nu = node_updates(tissues, args.steps)
    log("batched: %.3f s (%.4g node-updates/s)" % (dt_b, nu / dt_b))
    sample = args.plain_sample_steps or None
    dt_p = run_plain(gm, tissues, args.steps, sample_steps=sample)
    log("plain: %.3f s%s" % (dt_p, " (extrapolated from %d steps per tissue)" % sample if sample else ""))
    out = {
        "workload": "%d independent tissues, N ~ log-U[1, %d] (%d cells), stiffness, growth rate and adhesion "
                    "x U[0.7, 1.3], %d steps, FP32" % (len(tissues), args.nmax, cells, args.steps),
        "batched": {"seconds": dt_b, "sims_per_hour": len(tissues) / dt_b * 3600.0,
                    "node_updates_per_s": nu / dt_b, "wg_per_cu": wg, "graph_steps": gs},
        "plain": {"seconds": dt_p, "sims_per_hour": len(tissues) / dt_p * 3600.0,
                  "node_updates_per_s": nu / dt_p,
                  "extrapolated_from_steps_per_tissue": sample},
        "batched_over_plain": dt_p / dt_b,
        "autotune": table,
    }
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
