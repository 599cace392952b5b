#!/usr/bin/env python3
"""Headline benchmark: node-updates/s of the C180 cell-mechanics step at 2000
cells on one MI355X.

Workload (single tissue): 2000 unit-radius C180 cells (360,000 nodes) on a
seeded jittered simple-cubic lattice at volume fraction 0.40, FP32, grid
rebuilt every 10 steps, division off.  A step is one full pass of the step loop
(integrate, conservative + dissipative forces, velocity updates, second
dissipative force, centre of mass, grid rebuild on its schedule) over all
cells; the state is device-resident before the timed region starts.

Prints ONE JSON line.  The captured step graph is instantiated and uploaded
before the warm-up, so the timed region holds graph replays only.  `roofline`
covers the dominant kernel (the fused contact/friction kernel), timed live with
HIP events on the engine's stream; its HBM fraction is reported next to its
FP32 FLOP fraction and the vector-ALU busy fraction measured with rocprofv3
counters (the kernel is limited by vector-ALU issue, not bytes).
`cpu_baseline` is the CPU oracle (a C++/OpenMP restatement of the reference
kernels) timed on this host on a bounded sample of the same workload.
Further objects in the same line:
  search_full_launch_ms  one full contact-candidate search of every cell (the
                   search launch of the steps that search), HIP-event timed
  fp64_error       the FP32 engine against the oracle's FP64 mode on the same
                   2000-cell input (forces, centres of mass, volumes)
  bonded_roofline  the standalone bonded-force kernel on 20,000 cells (3.6 M
                   nodes), springs + pressure (24 B per node) and with bending,
                   HIP-event timed
  ensemble         256 independent tissues for 5000 steps with growth and
                   division, batched in one context, against a size-stratified
                   sample of the same tissues run one after another (plain path)
                   and the oracle on a smaller sample (sims/hour extrapolated)
  serving          256 requests of a Poisson stream through the
                   continuous-batching loop, admitted against a 200 GB budget
  c1_plain_small   8 cells, 10^4 steps with division at 2 V0, on the plain
                   path at full length; the oracle on the run's first steps
  c3_colony        one cell grown to about 20,000 by division; the oracle
                   stepping the colony's final positions
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# algorithmic HBM bytes per node and launch of each phase (FP32 SoA):
#   integrate: reads x, v, F_con, F_dis (48 B), writes x, v_after_integrate (24 B)
#   con_dis:   reads x, v_after_integrate (24 B), writes F_con, v (24 B)
#   dis_post:  reads x, v (24 B), writes F_dis (12 B)
ALG_BYTES = {"integrate": 72.0, "con_dis": 48.0, "dis_post": 36.0}
BONDED_BYTES = 24.0  # x in, F_con out, per node
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak


def con_dis_flops(pc, pin):
    """FP32 operations per node of k_con_dis: springs + pressure (75), bending
    (12 angle terms of about 65), the contact pass over the node's candidates
    (10 per candidate, 20 more per pair inside attraction_range), the first
    friction pass over the accepted pairs (10 per candidate, 17 per pair, 15 for
    bond damping, normal and drag) and the two velocity half-kicks (20)."""
    return 75.0 + 780.0 + (10.0 * pc + 20.0 * pin) + (10.0 * pc + 17.0 * pin + 15.0) + 20.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def graph_steps_for(k):
    """Steps per captured graph for a timed run of k steps: k itself up to 100, else
    the largest divisor of k up to 100 (so that k is a whole number of replays)."""
    if k <= 100:
        return max(1, k)
    return max(d for d in range(1, 101) if k % d == 0)


def load_traffic():
    """HBM bytes per launch measured with rocprofv3 counters (committed summary)."""
    path = os.path.join(ROOT, "profiles", "traffic_single_tissue_2000.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def _errors(a, b):
    """scale-relative max|a-b| / max|b|, and the largest relative error of the
    elements at least 1e-3 of the field's largest magnitude."""
    import numpy as np
    m = max(float(np.max(np.abs(b))), 1e-30)
    big = np.abs(b) >= 1e-3 * m
    el = float(np.max(np.abs(a[big] - b[big]) / np.abs(b[big]))) if big.any() else 0.0
    return float(np.max(np.abs(a - b)) / m), el


def fp64_error(cfg, gm, steps=10):
    """The FP32 engine against the oracle's FP64 mode on the 2000-cell input, with
    growth on (volumes every step) and division out of reach.  same_state: one
    evaluation from identical positions (the start-up forces and centres, and the
    volumes of the input state against a float64 evaluation of the same face
    decomposition); after_steps: both runs after `steps` steps, where the
    trajectories have also drifted apart by their rounding.  Part of the CPU leg:
    the oracle is the checker here."""
    import numpy as np
    from cellsim3d_amd import abi, colony, configs, engine
    from oracle import oracle as O
    p = cfg["params"].copy()
    p.growth_rate = 0.002
    p.division_vol = 1.0e30
    n = cfg["n"]
    node = lambda a: a.reshape(n, 192)[:, :180].ravel()
    sim = O.Sim(p, gm, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=n, budget=1 << 16, precision=1)
    sim.init()
    out = {"workload": "the 2000-cell input, growth on (volumes every step), no division; errors are "
                       "[scale-relative, largest per-element relative over elements >= 1e-3 of the field's "
                       "largest magnitude]"}
    groups = {"force_con": (abi.F_FCX, abi.F_FCY, abi.F_FCZ), "force_dis": (abi.F_FDX, abi.F_FDY, abi.F_FDZ),
              "centre_of_mass": (abi.F_CMX, abi.F_CMY, abi.F_CMZ)}
    with engine.Context(gm, capacity_cells=n, voxel_budget=1 << 24) as ctx:
        s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=n, voxels=1 << 18)
        same = {}
        for name, fields in groups.items():
            errs = [_errors(s.get(f)[:n * 192 if f in abi.NODE_FIELDS else n].astype(np.float64)
                            if f not in abi.NODE_FIELDS else node(s.get(f)[:n * 192].astype(np.float64)),
                            sim.get64(f)[:n] if f not in abi.NODE_FIELDS else node(sim.get64(f)[:n * 192]))
                    for f in fields]
            same[name] = [max(e[0] for e in errs), max(e[1] for e in errs)]
        s.k_volumes(1)
        v64 = colony.volumes_f64(configs.unit_mesh(), cfg["X"], cfg["Y"], cfg["Z"], n)
        same["volume"] = list(_errors(s.get(abi.F_VOLUME)[:n].astype(np.float64), v64))
        out["same_state"] = same
        sim.step(steps)  # (k_volumes wrote volumes, areas and flags only)
        s.step(steps)
        ctx.sync()
        after = {}
        groups["position"] = (abi.F_X, abi.F_Y, abi.F_Z)
        groups["volume"] = (abi.F_VOLUME,)
        for name, fields in groups.items():
            errs = []
            for f in fields:
                if f in abi.NODE_FIELDS:
                    errs.append(_errors(node(s.get(f)[:n * 192].astype(np.float64)), node(sim.get64(f)[:n * 192])))
                else:
                    errs.append(_errors(s.get(f)[:n].astype(np.float64), sim.get64(f)[:n]))
            after[name] = [max(e[0] for e in errs), max(e[1] for e in errs)]
        out["after_steps"] = {"steps": steps, **after}
    return out


def cpu_baseline(cfg, mesh_tables, rest, budget_s=20.0):
    from oracle import oracle as O
    om = O.build_mesh(mesh_tables, rest)
    n = cfg["n"]
    sim = O.Sim(cfg["params"], om, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=n, budget=1 << 16)
    sim.init()
    sim.step(2)  # warm-up
    t0 = time.perf_counter()
    sim.step(1)
    one = time.perf_counter() - t0
    k = max(1, min(200, int(budget_s / max(one, 1e-6))))
    ph0 = sim.phase_seconds()
    t0 = time.perf_counter()
    sim.step(k)
    dt = time.perf_counter() - t0
    ph = (sim.phase_seconds() - ph0) / k
    threads = O.get_threads()
    # one thread, on a short sample (a few seconds) of the same run
    O.set_threads(1)
    try:
        t0 = time.perf_counter()
        sim.step(1)
        one1 = time.perf_counter() - t0
        k1 = max(1, min(50, int(0.25 * budget_s / max(one1, 1e-6))))
        t0 = time.perf_counter()
        sim.step(k1)
        dt1 = time.perf_counter() - t0
    finally:
        O.set_threads(threads)
    return {
        "value": n * 180.0 * k / dt,
        "unit": "node-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle (C++/OpenMP restatement of the reference kernels), same 2000-cell input, "
                  "%d steps after 2 warm-up steps (%.1f s)" % (k, dt),
        "ms_per_step_by_part": {name: round(1e3 * v, 3) for name, v in zip(O.Sim.PHASES, ph)},
        "single_thread": {"value": n * 180.0 * k1 / dt1, "sample": "%d steps (%.1f s)" % (k1, dt1)},
    }


def bonded_roofline(gm, n_cells=20000, log=None):
    """Standalone bonded kernel (cs3d_k_con_force_bonded) on n_cells cells."""
    from cellsim3d_amd import configs, engine
    cfg = configs.single_tissue(n_cells, 0.40, seed=2, grid_interval=10, steps=10)
    res = {}
    with engine.Context(gm, capacity_cells=2 * n_cells + 8, voxel_budget=1 << 26) as ctx:
        for angles in (0, 1):
            p = cfg["params"].copy()
            p.constrain_angles = angles
            s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], n_cells, max_cells=n_cells, voxels=1 << 20)
            best = None
            for wg in (8, 16, 32, 64, 128):  # launch shape: workgroups per CU (cells per workgroup)
                ctx.set_launch(wg)
                ms = s.time_kernel(0, reps=50)
                if best is None or ms < best[1]:
                    best = (wg, ms)
            ctx.set_launch(16)
            wg, ms = best
            gbs = BONDED_BYTES * n_cells * 180 / (ms * 1e-3) / 1e9
            key = "with_bending" if angles else "springs_pressure"
            res[key] = {"avg_launch_ms": round(ms, 5), "achieved_GBs": round(gbs, 1),
                        "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "wg_per_cu": wg}
            if log:
                log("bench: bonded kernel (%s) %d cells: %.4f ms, %.0f GB/s (wg_per_cu %d)" %
                    (key, n_cells, ms, gbs, wg))
            s.remove()
    res["workload"] = "%d cells (%d nodes), %d B per node (x in, F_con out)" % (n_cells, n_cells * 180,
                                                                               BONDED_BYTES)
    return res


def c1_workload(gm, log, cpu=True, cpu_seconds=10.0, chunk=500):
    """The plain small configuration at its stated length (8 cells on a 2x2x2
    lattice, 10^4 steps, division at 2 V0): the GPU's drop-in path (one system, its
    own captured graph) over the whole run, timed per chunk; the oracle on the
    same run from its start for about cpu_seconds.  Both runs are bit-identical,
    so the oracle's node-updates over its prefix are the device-counted ones."""
    from cellsim3d_amd import abi, configs, engine
    cfg = configs.plain_small(seed=1, steps=10000)
    p, n, mc, steps = cfg["params"], cfg["n"], cfg["max_cells"], cfg["steps"]
    marks = []  # (steps, seconds, cell-steps) at chunk ends
    with engine.Context(gm, capacity_cells=mc + 8, voxel_budget=1 << 24) as ctx:
        s = ctx.add_system(p, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, voxels=1 << 16)
        ctx.set_graph_steps(100)
        ctx.prepare_graph(s)
        ctx.sync()
        t = 0.0
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            t0 = time.perf_counter()
            s.step(k)
            ctx.sync()
            t += time.perf_counter() - t0
            done += k
            marks.append((done, t, int(ctx.work(s.id, 1)[0])))
        n_end, _, status = s.info()
        if status & abi.SYS_FATAL_MASK:
            raise SystemExit("C1: system status 0x%x" % status)
    work = marks[-1][2] * 180.0
    out = {"workload": "8 unit-radius C180 cells on a 2x2x2 lattice (spacing 2.2, +-0.01 jitter), 10^4 steps, "
                       "dt 1e-4, FP32, division at 2 V0 (8 -> %d cells)" % n_end,
           "gpu_plain": {"seconds": round(t, 4), "ms_per_step": t / steps * 1e3, "node_updates": work,
                         "node_updates_per_s": work / t, "final_cells": int(n_end)}}
    log("bench: C1 GPU %.3f s for %d steps (%d cells at the end)" % (t, steps, n_end))
    if cpu:
        from oracle import oracle as O
        from cellsim3d_amd import mesh as meshmod
        om = O.build_mesh(meshmod.tables(cfg["mesh"]), configs.rest_positions(cfg["mesh"]))
        sim = O.Sim(p, om, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, budget=1 << 16)
        sim.init()
        tc = 0.0
        done = 0
        while done < steps and tc < cpu_seconds:
            t0 = time.perf_counter()
            sim.step(chunk)
            tc += time.perf_counter() - t0
            done += chunk
        mk = next(m for m in marks if m[0] == done)
        cw = mk[2] * 180.0
        threads = O.get_threads()
        O.set_threads(1)
        try:
            sim1 = O.Sim(p, om, cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=mc, budget=1 << 16)
            sim1.init()
            k1 = min(done, 2000)
            t0 = time.perf_counter()
            sim1.step(k1)
            t1 = time.perf_counter() - t0
        finally:
            O.set_threads(threads)
        w1 = next(m for m in marks if m[0] == k1)[2] * 180.0
        out["cpu_oracle"] = {"cores": threads, "sample": "the first %d steps of the same run (%d cells at its end)"
                                                         % (done, sim.n),
                             "seconds": round(tc, 3), "node_updates_per_s": cw / tc,
                             "gpu_same_steps_seconds": round(mk[1], 4), "gpu_over_cpu_same_steps": round(tc / mk[1], 1),
                             "single_thread": {"sample": "the first %d steps" % k1,
                                               "node_updates_per_s": w1 / t1}}
        log("bench: C1 oracle %d steps in %.2f s" % (done, tc))
    return out


def c3_workload(log, cpu=True, cpu_seconds=10.0):
    """The growing colony (one cell to about 20,000 by division) on the GPU, and
    the oracle stepping the colony's final state for about cpu_seconds: the CPU
    rate at the final size, beside the GPU's over its last chunk."""
    import numpy as np
    from cellsim3d_amd import abi, colony, configs
    from cellsim3d_amd import mesh as meshmod
    res = colony.run(keep_state=cpu, log=log)
    st = res.pop("state", None)
    log("bench: C3 colony %d cells in %d steps, %.1f s" % (res["final_cells"], res["steps"], res["wall_s"]))
    if cpu and st:
        from oracle import oracle as O
        m = st["mesh"]
        om = O.build_mesh(meshmod.tables(m), configs.rest_positions(m))
        n = st["n"]
        cap = n + 64
        sim = O.Sim(st["params"], om, st["X"][:n * 192], st["Y"][:n * 192], st["Z"][:n * 192], n,
                    max_cells=cap, budget=1 << 17)
        sim.set(abi.F_GROWTH_SCALE, st["growth_scale"][:cap].astype(np.float32))
        sim.init()
        sim.step(1)  # warm-up
        t0 = time.perf_counter()
        k = 0
        cell_steps = 0
        while time.perf_counter() - t0 < cpu_seconds:
            c0 = sim.n
            sim.step(1)
            cell_steps += c0
            k += 1
        tc = time.perf_counter() - t0
        threads = O.get_threads()
        O.set_threads(1)
        try:
            c0 = sim.n
            t1 = time.perf_counter()
            sim.step(1)
            t1 = time.perf_counter() - t1
        finally:
            O.set_threads(threads)
        res["cpu_oracle_at_final_size"] = {
            "cores": threads, "sample": "%d steps from the positions of the colony's final state (%d cells; velocities "
                                        "and pressures from rest), after one warm-up step" % (k, n),
            "node_updates_per_s": cell_steps * 180.0 / tc,
            "single_thread": {"sample": "1 step", "node_updates_per_s": c0 * 180.0 / t1}}
        res["gpu_over_cpu_at_final_size"] = round(res["final_chunk"]["node_updates_per_s"] /
                                                  (cell_steps * 180.0 / tc), 1)
        log("bench: C3 oracle %d steps at %d cells in %.2f s" % (k, n, tc))
    return res


def ensemble_cpu(tissues, ens, steps, every=32, seconds=10.0, chunk=25):
    """The oracle on a size-stratified sample of the ensemble's tissues (every
    `every`-th by size), each from its start for an equal share of about `seconds`.
    Each sampled tissue stands for `every` tissues of its size: the CPU time of the
    whole ensemble is `every` x the sum over the sample of (its cells x steps x 180,
    scaled by the ensemble's growth as counted on the GPU) / (its own CPU rate)."""
    from cellsim3d_amd import abi, configs, ensemble
    from cellsim3d_amd import mesh as meshmod
    from oracle import oracle as O
    m = configs.unit_mesh()
    om = O.build_mesh(meshmod.tables(m), configs.rest_positions(m))
    idx = ensemble.sample_tissues(tissues, every)
    share = seconds / len(idx)
    growth = ens["node_updates"] / (180.0 * steps * sum(t["n"] for t in tissues))
    cpu_hours = 0.0
    cell_steps = 0
    t_all = 0.0
    steps_run = []
    for i in idx:
        t = tissues[i]
        # room for several generations: the sample may run thousands of steps
        sim = O.Sim(t["params"], om, t["X"], t["Y"], t["Z"], t["n"], max_cells=16 * t["n"] + 16, budget=1 << 16)
        sim.init()
        k = 0
        cs = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < share:
            c0, s0 = sim.n, sim.stepno
            st = sim.step(chunk)
            done = sim.stepno - s0  # the oracle stops a call at a fatal status
            cs += c0 * done  # cells at the start of the chunk (a lower bound while growing)
            k += done
            if st & abi.SYS_FATAL_MASK:
                break
        dt = time.perf_counter() - t0
        t_all += dt
        cell_steps += cs
        steps_run.append(k)
        rate_i = cs * 180.0 / dt
        cpu_hours += every * growth * t["n"] * steps * 180.0 / rate_i / 3600.0
    rate = cell_steps * 180.0 / t_all
    return {"cores": O.get_threads(), "node_updates_per_s": rate,
            "sims_per_hour_extrapolated": round(len(tissues) / cpu_hours, 1),
            "sample": "%d of %d tissues (every %d-th by size), each from its start for %d-%d steps (%.1f s); "
                      "the full run's CPU time from each sampled tissue's own rate, its size and the "
                      "ensemble's growth (x%.2f)" % (len(idx), len(tissues), every, min(steps_run),
                                                    max(steps_run), t_all, growth),
            "batched_gpu_over_cpu_sims": round(ens["batched_sims_per_hour"] / (len(tissues) / cpu_hours), 1)}


def measure_extras(args, gm, log):
    from cellsim3d_amd import configs, ensemble, serving
    out = {"bonded_roofline": bonded_roofline(gm, log=log)}
    out["c1_plain_small"] = c1_workload(gm, log, cpu=not args.no_cpu)
    out["c3_colony"] = c3_workload(log, cpu=not args.no_cpu)
    g, tissues = ensemble.build(256, 1, 1, 512, args.ensemble_steps, 10)
    wg, gs, skin, _ = ensemble.autotune(g, tissues, probe_steps=50)
    out["ensemble"] = ensemble.measure(g, tissues, args.ensemble_steps, wg, gs, args.plain_every, log, skin=skin)
    log("bench: ensemble x%.2f over the plain loop" % out["ensemble"]["batched_over_plain"])
    if not args.no_cpu:
        out["ensemble"]["cpu_oracle"] = ensemble_cpu(tissues, out["ensemble"], args.ensemble_steps)
        log("bench: ensemble oracle %.3g node-updates/s" % out["ensemble"]["cpu_oracle"]["node_updates_per_s"])
    reqs = configs.serving_requests(args.serving_requests, 1, 20.0, 8, 4096, 1e3, 1e5)
    jobs = serving.make_jobs(reqs)
    srv = serving.Server(g, 200e9, chunk_steps=50)
    wall = srv.run(jobs)
    srv.close()
    out["serving"] = serving.summary(jobs, srv, wall)
    log("bench: serving p50 %.3f s p99 %.3f s rho %.2f" % (out["serving"]["p50_latency_s"],
                                                         out["serving"]["p99_latency_s"], out["serving"]["rho"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cells", type=int, default=2000)
    ap.add_argument("--phi", type=float, default=0.40)
    ap.add_argument("--grid-interval", type=int, default=10)
    ap.add_argument("--timed-steps", type=int, default=100, help="steps with per-phase HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run the known-size copy kernel (for HBM counter scaling)")
    ap.add_argument("--no-extras", action="store_true", help="skip the bonded, ensemble and serving measurements")
    ap.add_argument("--bonded-only", action="store_true",
                    help="only the standalone bonded-kernel sweep (for profiling it); prints its JSON")
    ap.add_argument("--pair-skin", type=float, default=None,
                    help="contact-candidate skin (performance only: results do not depend on it)")
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--plain-every", type=int, default=8,
                    help="plain-path sample of the ensemble: every k-th tissue by size, full length")
    ap.add_argument("--serving-requests", type=int, default=256)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every system of this engine runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cellsim3d_amd import abi, configs, engine, mesh

    cfg = configs.single_tissue(args.cells, args.phi, seed=1, grid_interval=args.grid_interval,
                                steps=args.steps + max(args.warmup, graph_steps_for(args.steps)) + 100)
    if args.pair_skin is not None:
        cfg["params"].pair_skin = args.pair_skin
    tables = mesh.tables(cfg["mesh"])
    rest = configs.rest_positions(cfg["mesh"])
    gm = engine.build_mesh(tables, rest)
    if args.bonded_only:
        print(json.dumps(bonded_roofline(gm, log=log)), flush=True)
        return
    n = cfg["n"]
    log("bench: %d cells, phi %.2f, grid every %d steps" % (n, args.phi, args.grid_interval))
    ctx = engine.Context(gm, capacity_cells=n, voxel_budget=1 << 24)
    s = ctx.add_system(cfg["params"], cfg["X"], cfg["Y"], cfg["Z"], n, max_cells=n, voxels=1 << 18)
    # the timed K steps replay one captured graph of G steps K/G times: G = K when K
    # is at most 100, else the largest divisor of K up to 100.  The graph is
    # captured, instantiated and uploaded before the warm-up, and the warm-up runs
    # at least one replay of it, so that the timed region holds warm replays only
    # (W warm-up steps, rounded up to a multiple of G)
    gs = graph_steps_for(args.steps)
    ctx.set_graph_steps(gs)
    ctx.prepare_graph(s)
    warm = -(-max(args.warmup, gs) // gs) * gs
    ctx.step(warm, s)
    ctx.sync()
    log("bench: graph of %d steps prepared, %d warm-up steps done" % (gs, warm))
    t0 = time.perf_counter()
    ctx.step(args.steps, s)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    n_now, step_now, status = s.info()
    if status & abi.SYS_FATAL_MASK:
        raise SystemExit("system status 0x%x after the timed run" % status)
    value = n * 180.0 * args.steps / elapsed
    log("bench: %d steps in %.4f s -> %.4g node-updates/s" % (args.steps, elapsed, value))
    # contact partners (pairs within attraction_range that pass the candidate rule: the
    # per-node work of the friction passes) and candidates (nodes within attraction_range +
    # pair_skin at the last search: the per-node work of the contact pass) at the end of the
    # timed run
    pc = s.get(abi.F_PAIRCOUNT)[:n * 192].reshape(n, 192)[:, :180]
    cc = s.get(abi.F_CANDCOUNT)[:n * 192].reshape(n, 192)[:, :180]
    contacts = {"partners_per_node_mean": round(float(pc.mean()), 3), "partners_per_node_max": int(pc.max()),
                "candidates_per_node_mean": round(float(cc.mean()), 3), "candidates_per_node_max": int(cc.max())}

    # per-phase kernel time from start/stop HIP events each launch records on the
    # engine's stream (eager steps; launch gaps are not counted)
    ph = ctx.step_timed(max(1, args.timed_steps), s) / max(1, args.timed_steps)
    phases = {name: float(ms) for name, ms in zip(engine.PHASE_NAMES, ph)}
    dom = max(("con_dis", "dis_post", "integrate"), key=lambda k: phases[k])
    dom_ms = phases[dom]
    alg = ALG_BYTES[dom] * n * 180
    achieved = alg / (dom_ms * 1e-3) / 1e9
    traffic = None
    tr = load_traffic()
    if tr and dom in tr.get("bytes_per_launch", {}):
        traffic = tr["bytes_per_launch"][dom]
    roofline = {
        "kernel": {"con_dis": "k_con_dis", "dis_post": "k_dis_post", "integrate": "k_integrate"}[dom],
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "avg_launch_ms": round(dom_ms, 5),
        "alg_bytes_per_launch": alg,
    }
    if dom == "con_dis":
        # the kernel's limit is vector-ALU issue: its FLOP fraction next to the HBM one
        fl = con_dis_flops(contacts["candidates_per_node_mean"], contacts["partners_per_node_mean"])
        tf = fl * n * 180 / (dom_ms * 1e-3) / 1e12
        roofline["limiter"] = "valu"
        roofline["flops_per_node"] = round(fl, 1)
        roofline["achieved_tflops"] = round(tf, 3)
        roofline["valu_frac"] = round(tf / FP32_PEAK_TFLOPS, 5)
        roofline["valu_peak_tflops"] = FP32_PEAK_TFLOPS
        if tr and dom in tr.get("valu_busy", {}):
            # the fraction of vector-ALU issue slots the kernel used, from rocprofv3
            # counters (a run of its own: counters cannot be read in this run)
            roofline["valu_busy_counters"] = tr["valu_busy"][dom]
    log("bench: phases ms/step %s" % json.dumps({k: round(v, 4) for k, v in phases.items()}))
    # one full candidate search of every cell (the search launch of the steps that
    # search; the next step searches again)
    search_ms = s.time_kernel(2, reps=10)
    log("bench: full candidate search %.4f ms" % search_ms)

    calib = None
    if args.calib_copy_mb > 0:
        nbytes = args.calib_copy_mb << 20
        cms = ctx.copy_calibrate(nbytes, reps=5)
        calib = {"bytes_read": nbytes, "bytes_written": nbytes, "avg_ms": cms,
                 "GBs": round(2 * nbytes / (cms * 1e-3) / 1e9, 1)}
        log("bench: copy kernel %d MiB: %.4f ms (%.0f GB/s read+write)" % (args.calib_copy_mb, cms, calib["GBs"]))
    extras = {}
    if not args.no_extras:
        extras = measure_extras(args, gm, log)
    cpu = None
    fp64 = None
    if not args.no_cpu:
        log("bench: CPU oracle baseline (~%.0f s)" % args.cpu_seconds)
        cpu = cpu_baseline(cfg, tables, rest, args.cpu_seconds)
        log("bench: FP64 oracle run for the FP32 error")
        fp64 = fp64_error(cfg, gm)
    ms_per_step = elapsed / args.steps * 1e3
    out = {
        "metric": "node-updates/s at 2k cells",
        "value": value,
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": warm,
        "graph_steps": gs,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single tissue: %d C180 cells (%d nodes), volume fraction %.2f, FP32, "
                        "grid rebuilt every %d steps, division off" % (n, n * 180, args.phi, args.grid_interval),
            "cells": n,
            "nodes": n * 180,
            "grid_interval": args.grid_interval,
        },
        "phases_ms_per_step": {k: round(v, 5) for k, v in phases.items()},
        "search_full_launch_ms": round(search_ms, 5),
        "contacts": contacts,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if fp64:
        out["fp64_error"] = fp64
    out.update(extras)
    if calib:
        out["copy_calibration"] = calib
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
