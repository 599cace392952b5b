"""Serving: a request loop with continuous batching on one GPU.

Simulation requests (a tissue, its parameters and a horizon in steps) arrive
over time.  The server admits a request as soon as its state fits in the
device pool (cell slots, voxel and search-bin regions, which are sized from an
HBM budget), uploads it once, and from then on steps it together with every
other resident tissue: each chunk is one cs3d_step over all systems, and a
system that reaches its horizon stops on the device while the others go on.
Between chunks the server polls all systems with one device read, returns the
final state of the finished ones and frees their slots for the queue.

The reference has no serving mode; a tissue's result is the same whether it
runs alone or in the mix (tests/test_gpu_serving.py checks this bit for bit).

    python -m cellsim3d_amd.serving [--requests 64] [--rate 20] ...
prints one JSON line: p50/p99 completion latency, aggregate node-updates/s and
the offered load (offered over served node-updates/s; above 1 requests queue).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, configs, engine
from . import mesh as meshmod

# device bytes per cell slot (node arrays, contact lists, per-cell scalars),
# and the voxel and search-bin words cs3d_system_add reserves per system:
# max(4096, 32 * cells) voxels of (1 + capacity) words, max(16384, 640 * cells)
# bins of abi.BIN_WORDS words
SLOT_BYTES = (15 + 3) * 192 * 4 + 2 * 40 * 192 * 4 + 2 * 192 * 4 + 16 * 4 + 48 + 1
VOX_WORDS_PER_CELL = 32 * 65
BIN_WORDS_PER_CELL = 640 * abi.BIN_WORDS
VOX_WORDS_MIN = 4096 * 65
BIN_WORDS_MIN = 16384 * abi.BIN_WORDS
SMALL_SYSTEMS = 64  # room for this many systems at the minimum reservations


class Job:
    def __init__(self, jid, arrival, params, X, Y, Z, n, horizon, max_cells=None):
        self.id = jid
        self.arrival = arrival
        self.params = params
        self.X, self.Y, self.Z = X, Y, Z
        self.n = n
        self.horizon = horizon
        self.max_cells = max_cells or (n + max(2, n // 8))
        self.admitted = None
        self.finished = None
        self.result = None

    def bytes(self):
        return self.max_cells * SLOT_BYTES + 4 * (max(VOX_WORDS_MIN, self.max_cells * VOX_WORDS_PER_CELL) +
                                                  max(BIN_WORDS_MIN, self.max_cells * BIN_WORDS_PER_CELL))


def make_jobs(requests, grid_interval=10, phi=0.30):
    """Jobs from configs.serving_requests entries (seeded tissues)."""
    jobs = []
    for r in requests:
        _, X, Y, Z, info = configs.lattice_tissue(r["n"], phi, seed=r["seed"])
        p = configs.base_params(1 + r["seed"] % 300)
        # fixed cell counts: horizons reach 1e5 steps, over which growing cells would
        # divide many times; growth_rate = 0 turns volumes and division off
        # (GPUbounce.cu:1168)
        p.growth_rate = 0.0
        p.division_vol = configs.division_vol_unit(info["v0"])
        p.div_time_steps = r["horizon"]
        p.non_div_time_steps = 0
        p.grid_interval = grid_interval
        p.voxel_capacity = 64
        jobs.append(Job(len(jobs), r["arrival"], p, X, Y, Z, r["n"], r["horizon"]))
    return jobs


class Server:
    """policy "least_work": before each chunk the resident jobs are ordered by their
    remaining work (cells x steps still to go) and the chunk steps the shortest
    ones, up to active_cells cells (at least one job); the others are paused on the
    device (cs3d_systems_set_horizon) and stay resident.  A job that arrives with
    little work therefore finishes without waiting behind long ones, while the batch
    stays large enough to fill the GPU.  policy "all": every resident job steps in
    every chunk (processor sharing by steps).  Results do not depend on the policy:
    pausing leaves a system's state as it is."""

    def __init__(self, gm, hbm_budget_bytes, max_cells_total=None, chunk_steps=50, graph_steps=10,
                 wg_per_cu=16, fields=(abi.F_X, abi.F_Y, abi.F_Z), policy="least_work", active_cells=16384):
        if policy not in ("least_work", "all"):
            raise ValueError("policy must be 'least_work' or 'all'")
        self.policy = policy
        self.active_cells = int(active_cells)
        self.pauses = 0  # jobs paused by the scheduler (a job counts once per pause)
        per_cell = SLOT_BYTES + 4 * (VOX_WORDS_PER_CELL + BIN_WORDS_PER_CELL)
        cap = int(hbm_budget_bytes // per_cell)
        if max_cells_total:
            cap = min(cap, int(max_cells_total))
        self.capacity = cap
        self.budget = hbm_budget_bytes
        self.peak_jobs = 0
        self.ctx = engine.Context(gm, capacity_cells=cap,
                                  voxel_budget=cap * VOX_WORDS_PER_CELL + SMALL_SYSTEMS * VOX_WORDS_MIN,
                                  bin_budget=cap * BIN_WORDS_PER_CELL + SMALL_SYSTEMS * BIN_WORDS_MIN)
        self.ctx.set_graph_steps(graph_steps)
        self.ctx.set_launch(wg_per_cu)
        self.chunk = chunk_steps
        self.fields = fields
        self.node_updates = 0.0

    def close(self):
        self.ctx.close()

    def _admit(self, job, now):
        try:
            s = self.ctx.add_system(job.params, job.X, job.Y, job.Z, job.n, max_cells=job.max_cells,
                                    voxels=max(4096, 32 * job.max_cells))
        except engine.EngineError as e:
            if e.status == abi.CAPACITY:
                return None
            raise
        s.set_horizon(job.horizon)
        job.admitted = now
        return s

    def _finish(self, job, s, now):
        n = s.n
        job.result = {f: s.get(f)[:n * 192].copy() for f in self.fields}
        job.result["n"] = n
        job.result["status"] = s.info()[2]
        job.finished = now
        s.remove()

    def run(self, jobs, time_scale=1.0, log=None):
        """Serves jobs (arrival times in seconds, scaled by time_scale) and returns
        the wall time.  Each job gets admitted/finished times and a result."""
        queue = sorted(jobs, key=lambda j: j.arrival)
        active = {}
        steps_now = {}  # system id -> steps done at the last poll
        running = None  # system ids the last horizon update let step
        paused = set()  # system ids it paused
        t0 = time.perf_counter()
        qi = 0
        while qi < len(queue) or active:
            now = time.perf_counter() - t0
            while qi < len(queue) and queue[qi].arrival * time_scale <= now:
                s = self._admit(queue[qi], now)
                if s is None:
                    if not active:
                        raise RuntimeError("job %d (%d cells) does not fit an empty pool of %d cells" %
                                           (queue[qi].id, queue[qi].n, self.capacity))
                    break  # first in line waits for room (arrival order)
                active[s.id] = (queue[qi], s)
                steps_now[s.id] = 0
                running = None  # the new job takes part in the next choice
                qi += 1
            if not active:
                if qi < len(queue):
                    time.sleep(max(0.0, queue[qi].arrival * time_scale - (time.perf_counter() - t0)))
                continue
            self.peak_jobs = max(self.peak_jobs, len(active))
            hi = max(active) + 1
            if self.policy == "least_work":
                order = sorted(active, key=lambda sid: (active[sid][0].n * (active[sid][0].horizon - steps_now[sid]),
                                                        active[sid][0].arrival, sid))
                chosen, cells = set(), 0
                for sid in order:
                    chosen.add(sid)
                    cells += active[sid][0].n
                    if cells >= self.active_cells:
                        break
                if chosen != running:
                    hz = np.full(hi, -1, dtype=np.int64)
                    for sid, (job, s) in active.items():
                        hz[sid] = job.horizon if sid in chosen else steps_now[sid]
                    held = set(active) - chosen
                    self.pauses += len(held - paused)
                    paused = held
                    self.ctx.set_horizons(0, hz)
                    running = chosen
            self.ctx.step(self.chunk)
            ncell, steps, status = self.ctx.poll(0, hi)
            now = time.perf_counter() - t0
            for sid in list(active):
                job, s = active[sid]
                steps_now[sid] = int(steps[sid])
                done = steps[sid] >= job.horizon or (status[sid] & abi.SYS_FATAL_MASK)
                if done:
                    self.node_updates += float(job.n) * 180.0 * float(steps[sid])
                    self._finish(job, s, now)
                    del active[sid]
                    del steps_now[sid]
                    running = None
                    if log:
                        log("job %d done: %d cells, %d steps, latency %.3f s" %
                            (job.id, job.n, job.horizon, now - job.arrival * time_scale))
        return time.perf_counter() - t0


def summary(jobs, srv, wall):
    """p50/p99 completion latency (result on the host minus arrival), aggregate
    node-updates/s and the utilisation rho = offered node-updates/s over served
    node-updates/s (above 1 the requests queue, as streams of the configured size do)."""
    lat = np.array([j.finished - j.arrival for j in jobs])
    span = max(j.arrival for j in jobs) if jobs else 0.0
    offered = sum(j.n * 180.0 * j.horizon for j in jobs)
    served = srv.node_updates / wall
    waited = [j.admitted - j.arrival for j in jobs]
    return {
        "workload": "%d requests, Poisson arrivals over %.1f s, N ~ log-U[%d, %d] (mean %.0f), horizon ~ "
                    "log-U[%.0e, %.0e] steps (mean %.0f), FP32, growth off, admitted against a %.0f GB HBM budget "
                    "(pool of %d cells)" % (
                        len(jobs), span, min(j.n for j in jobs), max(j.n for j in jobs),
                        float(np.mean([j.n for j in jobs])), min(j.horizon for j in jobs),
                        max(j.horizon for j in jobs), float(np.mean([j.horizon for j in jobs])),
                        srv.budget / 1e9, srv.capacity),
        "requests": len(jobs),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
        "p50_queue_wait_s": round(float(np.percentile(waited, 50)), 4),
        "wall_s": round(wall, 3),
        "node_updates_per_s": served,
        "offered_node_updates_per_s": offered / span if span > 0 else None,
        "rho": round((offered / span) / served, 3) if span > 0 and served > 0 else None,
        "peak_resident_jobs": srv.peak_jobs,
        "policy": srv.policy if srv.policy == "all" else "%s (the jobs with the least work left step, until %d "
                                                          "cells)" % (srv.policy, srv.active_cells),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--nmax", type=int, default=4096)
    ap.add_argument("--hmin", type=float, default=1e3)
    ap.add_argument("--hmax", type=float, default=1e5)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-cells", type=int, default=0, help="cap of the device pool in cells (0: the budget decides)")
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--policy", choices=("least_work", "all"), default="least_work")
    ap.add_argument("--active-cells", type=int, default=16384,
                    help="least_work: cells of the jobs that step in a chunk (the least work left first)")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    m = configs.unit_mesh()
    gm = engine.build_mesh(meshmod.tables(m), configs.rest_positions(m))
    reqs = configs.serving_requests(args.requests, args.seed, args.rate, 8, args.nmax, args.hmin, args.hmax)
    jobs = make_jobs(reqs)
    srv = Server(gm, args.budget_gb * 1e9, max_cells_total=args.max_cells or None, chunk_steps=args.chunk,
                 policy=args.policy, active_cells=args.active_cells)
    log("serving: %d requests, %d cells in total, pool of %d cells" % (len(jobs), sum(j.n for j in jobs),
                                                                      srv.capacity))
    wall = srv.run(jobs, log=log)
    srv.close()
    out = summary(jobs, srv, wall)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
