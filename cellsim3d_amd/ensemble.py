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
    X, Y, Z, n, max_cells = room * n + 2.  room = 2 lets every cell divide once;
    the runners below move a tissue to a larger room (cs3d_system_grow) whenever
    its cells could otherwise not all divide once more."""
    m, tissues = configs.ensemble(n_tissues=n_tissues, seed=seed, nmin=nmin, nmax=nmax, phi=phi, steps=steps)
    gm = engine.build_mesh(meshmod.tables(m), configs.rest_positions(m))
    for t in tissues:
        t["params"].grid_interval = grid_interval
        t["max_cells"] = room * t["n"] + 2
    return gm, tissues


def node_updates(tissues, steps):
    """Node-updates of the initial cell counts (a lower bound for growing tissues)."""
    return float(sum(t["n"] for t in tissues)) * 180.0 * steps


def grow_rooms(ctx, systems, headroom=2.5):
    """Moves every system whose live cells could not all divide once more to a
    room of headroom * n + 2 cells.  A tissue's cells at most double between two
    calls when the calls are closer together than a cell's regrowth after
    division (about 1000 steps at the default growth rate).  Returns the number
    of systems moved."""
    if not systems:
        return 0
    hi = max(s.id for s in systems) + 1
    n, _, _ = ctx.poll(0, hi)
    moved = 0
    for s in systems:
        k = int(n[s.id])
        if 2 * k + 2 > s.max_cells:
            s.grow(int(headroom * k) + 2)
            moved += 1
    return moved


def run_chunks(ctx, target, steps, systems, chunk=100):
    """steps steps of target in chunks, with rooms grown between chunks.
    Returns the number of room moves."""
    moved = 0
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        ctx.step(k, target)
        done += k
        if done < steps:
            moved += grow_rooms(ctx, systems)
    return moved


def _context(gm, tissues, room=1.0):
    """A context for the tissues with room x their cell slots (growth headroom)."""
    cap = int(room * sum(t["max_cells"] for t in tissues))
    vox = int(room * sum(max(4096, 32 * t["max_cells"]) * (1 + t["params"].voxel_capacity) for t in tissues))
    bins = int(room * sum(min(max(16384, 640 * t["max_cells"]), 1 << 40) for t in tissues)) * abi.BIN_WORDS
    return engine.Context(gm, capacity_cells=cap, voxel_budget=vox, bin_budget=bins)


def run_batched(gm, tissues, steps, wg_per_cu=None, graph_steps=None, ctx=None, chunk=100, room=4.0):
    """All tissues in one context, stepped together through the all-systems graph
    (captured before the timed region), in chunks of `chunk` steps with rooms
    grown between chunks.  Returns (seconds, ctx, systems, info) with info =
    {"cell_steps", "moves", "final_cells"}."""
    own = ctx is None
    if own:
        ctx = _context(gm, tissues, room)
    if wg_per_cu:
        ctx.set_launch(wg_per_cu)
    if graph_steps:
        ctx.set_graph_steps(graph_steps)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    ctx.prepare_graph()
    ctx.sync()
    t0 = time.perf_counter()
    moves = run_chunks(ctx, abi.ALL, steps, systems, chunk)
    ctx.sync()
    dt = time.perf_counter() - t0
    hi = max(s.id for s in systems) + 1
    n, _, status = ctx.poll(0, hi)
    for s in systems:
        if status[s.id] & abi.SYS_FATAL_MASK:
            raise RuntimeError("system %d status 0x%x" % (s.id, status[s.id]))
    work = ctx.work(0, hi)
    info = {"cell_steps": int(sum(int(work[s.id]) for s in systems)), "moves": moves,
            "final_cells": int(sum(int(n[s.id]) for s in systems))}
    return dt, ctx, systems, info


def sample_tissues(tissues, every):
    """Every `every`-th tissue in order of size (a sample stratified by size)."""
    order = sorted(range(len(tissues)), key=lambda i: (tissues[i]["n"], i))
    return [order[k] for k in range(every // 2, len(order), every)]


def run_plain(gm, tissues, steps, indices=None, chunk=100, room=16.0):
    """The tissues one after another, each alone on the drop-in path (its own
    graph), for the full `steps`, with the same room growth as the batched run.
    indices selects a sample (all by default).  Returns (seconds, cell_steps) over
    the tissues run."""
    idx = list(range(len(tissues))) if indices is None else list(indices)
    big = max((tissues[i] for i in idx), key=lambda t: t["max_cells"])
    ctx = _context(gm, [big], room)
    total = 0.0
    cell_steps = 0
    for i in idx:
        t = tissues[i]
        s = ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"])
        ctx.prepare_graph(s)
        ctx.sync()
        t0 = time.perf_counter()
        run_chunks(ctx, s.id, steps, [s], chunk)
        ctx.sync()
        total += time.perf_counter() - t0
        st = s.info()[2]
        if st & abi.SYS_FATAL_MASK:
            raise RuntimeError("tissue %d status 0x%x on the plain path" % (i, st))
        cell_steps += int(ctx.work(s.id, 1)[0])
        s.remove()
    ctx.close()
    return total, cell_steps


def with_skin(tissues, skin):
    """Copies of the tissues with the contact-candidate skin set to `skin`.  The
    skin sets the search tile (bin edge (attraction_range + skin) * 1.001) and how
    often the candidates are searched again; results do not depend on it."""
    if skin is None:
        return tissues
    out = []
    for t in tissues:
        u = dict(t)
        u["params"] = t["params"].copy()
        u["params"].pair_skin = skin
        out.append(u)
    return out


def _probe(ctx, gs, probe_steps):
    ctx.set_graph_steps(gs)
    ctx.prepare_graph()
    ctx.step(gs)  # warm
    ctx.sync()
    k = max(gs, (probe_steps // gs) * gs)
    t0 = time.perf_counter()
    ctx.step(k)
    ctx.sync()
    return (time.perf_counter() - t0) / k * 1e3


def autotune(gm, tissues, probe_steps=100, wg_options=(4, 8, 16, 32, 64), graph_options=(10, 25, 50),
             skin_options=(0.03, 0.05, 0.08), log=None):
    """Picks, by timing short batched runs, the launch cap (workgroups per CU of
    the cell kernels, i.e. how many cells a workgroup walks through) and the
    steps per captured graph over their grid of options, then, at that pair, the
    contact-search tile: the skin, which sets the edge of the search bins and how
    often candidates are searched again.  Returns (wg_per_cu, graph_steps, skin,
    table)."""
    ctx = _context(gm, tissues)
    systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tissues]
    table = []
    for wg in wg_options:
        for gs in graph_options:
            ctx.set_launch(wg)
            ms = _probe(ctx, gs, probe_steps)
            table.append({"wg_per_cu": wg, "graph_steps": gs, "ms_per_step": ms})
            if log:
                log("autotune: wg_per_cu %d graph_steps %d -> %.4f ms/step" % (wg, gs, ms))
    for s in systems:
        s.remove()
    ctx.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    wg, gs = best["wg_per_cu"], best["graph_steps"]
    skins = []
    for skin in skin_options:
        ctx = _context(gm, tissues)
        ctx.set_launch(wg)
        tk = with_skin(tissues, skin)
        systems = [ctx.add_system(t["params"], t["X"], t["Y"], t["Z"], t["n"], max_cells=t["max_cells"]) for t in tk]
        ms = _probe(ctx, gs, probe_steps)
        skins.append({"wg_per_cu": wg, "graph_steps": gs, "pair_skin": skin, "ms_per_step": ms})
        if log:
            log("autotune: pair_skin %.3f -> %.4f ms/step" % (skin, ms))
        ctx.close()
    table += skins
    skin = min(skins, key=lambda r: r["ms_per_step"])["pair_skin"] if skins else None
    return wg, gs, skin, table


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--tissues", type=int, default=256)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--nmax", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--grid-interval", type=int, default=10)
    ap.add_argument("--plain-every", type=int, default=8,
                    help="the plain path runs every k-th tissue by size for the full length (1 = all)")
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    gm, tissues = build(args.tissues, args.seed, 1, args.nmax, args.steps, args.grid_interval)
    cells = sum(t["n"] for t in tissues)
    log("ensemble: %d tissues, %d cells, %d steps" % (len(tissues), cells, args.steps))
    wg, gs, skin, table = (16, 10, None, [])
    if not args.no_tune:
        wg, gs, skin, table = autotune(gm, tissues, log=log)
    out = measure(gm, tissues, args.steps, wg, gs, args.plain_every, log, skin=skin)
    out["autotune"] = table
    print(json.dumps(out), flush=True)
    return out


def measure(gm, tissues, steps, wg, gs, plain_every=8, log=None, skin=None):
    """Batched run of every tissue and plain runs of a size-stratified sample,
    both for the full length with divisions and with the same contact skin;
    node-updates from the cells actually stepped (counted on the device)."""
    tissues = with_skin(tissues, skin)
    dt_b, ctx, _, info = run_batched(gm, tissues, steps, wg_per_cu=wg, graph_steps=gs)
    ctx.close()
    nu = 180.0 * info["cell_steps"]
    if log:
        log("ensemble batched: %.3f s, %d -> %d cells, %d room moves, %.4g node-updates/s" %
            (dt_b, sum(t["n"] for t in tissues), info["final_cells"], info["moves"], nu / dt_b))
    idx = sample_tissues(tissues, plain_every) if plain_every > 1 else None
    dt_s, cs_s = run_plain(gm, tissues, steps, idx)
    ntis = len(idx) if idx is not None else len(tissues)
    # the sample is one tissue in plain_every by size: its time scales by that factor
    scale = len(tissues) / float(ntis)
    dt_p = dt_s * scale
    if log:
        log("ensemble plain: %d tissues in %.3f s (x%.2f -> %.3f s)" % (ntis, dt_s, scale, dt_p))
    return {
        "workload": "%d independent tissues, N ~ log-U[1, 512] (%d cells at the start, %d at the end), stiffness, "
                    "growth rate and adhesion x U[0.7, 1.3], %d steps with growth and division, FP32" % (
                        len(tissues), sum(t["n"] for t in tissues), info["final_cells"], steps),
        "batched_sims_per_hour": round(len(tissues) / dt_b * 3600.0, 1),
        "batched_node_updates_per_s": nu / dt_b,
        "batched_seconds": round(dt_b, 4),
        "room_moves": info["moves"],
        "plain_sims_per_hour": round(len(tissues) / dt_p * 3600.0, 1),
        "plain_node_updates_per_s": nu / dt_p,
        "plain_seconds": round(dt_p, 4),
        "plain_sample": "%d of %d tissues (every %d-th by size), each for the full %d steps with divisions, "
                        "%.3f s, time scaled by %.2f" % (ntis, len(tissues), plain_every, steps, dt_s, scale),
        "node_updates": nu,
        "node_updates_counted": "cells stepped in every step, counted on the device (growing tissues)",
        "batched_over_plain": round(dt_p / dt_b, 2),
        "autotuned": {"wg_per_cu": wg, "graph_steps": gs,
                      "pair_skin": float(skin) if skin is not None else float(tissues[0]["params"].pair_skin)},
    }


if __name__ == "__main__":
    main()
