"""CPU legs of bench.py, run on small inputs: the oracle's samples for the
ensemble and for the colony's final state, and the tissue copies with another
contact skin.  These are the parts of the bench line that run on the host; the
GPU legs are exercised by the GPU tests and by bench.py itself."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _tissues(k=16):
    from cellsim3d_amd import configs
    _, tissues = configs.ensemble(n_tissues=k, seed=2, nmin=1, nmax=12, phi=0.30, steps=100)
    for t in tissues:
        t["params"].grid_interval = 10
        t["max_cells"] = 2 * t["n"] + 2
    return tissues


def test_with_skin_copies_parameters():
    from cellsim3d_amd import ensemble
    tissues = _tissues(4)
    before = [float(t["params"].pair_skin) for t in tissues]
    out = ensemble.with_skin(tissues, 0.08)
    assert [float(t["params"].pair_skin) for t in tissues] == before  # the originals are untouched
    assert all(abs(float(t["params"].pair_skin) - 0.08) < 1e-7 for t in out)
    assert all(a["X"] is b["X"] and a["n"] == b["n"] for a, b in zip(tissues, out))
    assert ensemble.with_skin(tissues, None) is tissues


def test_ensemble_cpu_sample_and_extrapolation():
    import bench
    tissues = _tissues(16)
    cells = sum(t["n"] for t in tissues)
    ens = {"node_updates": 1.5 * 180.0 * 100 * cells, "batched_node_updates_per_s": 1e9,
           "batched_sims_per_hour": 1e5}
    r = bench.ensemble_cpu(tissues, ens, 100, every=8, seconds=0.6)
    assert r["node_updates_per_s"] > 0 and r["sims_per_hour_extrapolated"] > 0
    assert r["cores"] >= 1
    assert "2 of 16 tissues" in r["sample"] and "x1.50" in r["sample"]
    expect = ens["batched_sims_per_hour"] / r["sims_per_hour_extrapolated"]
    assert abs(r["batched_gpu_over_cpu_sims"] - round(expect, 1)) <= 0.1


def test_colony_cpu_leg_on_a_given_state(monkeypatch):
    import bench
    from cellsim3d_amd import colony, configs
    m, X, Y, Z, _ = configs.lattice_tissue(20, 0.35, seed=4)
    p = configs.colony(seed=1, max_cells=200)["params"]
    p.voxel_capacity = 128
    gs = np.random.default_rng(1).uniform(0.8, 1.2, size=200).astype(np.float32)

    def fake_run(keep_state=False, log=None):
        out = {"final_cells": 20, "steps": 1, "wall_s": 1.0,
               "final_chunk": {"cells": 20, "ms_per_step": 1.0, "node_updates_per_s": 1e8}}
        if keep_state:
            out["state"] = {"params": p, "mesh": m, "X": X, "Y": Y, "Z": Z, "n": 20, "growth_scale": gs}
        return out

    monkeypatch.setattr(colony, "run", fake_run)
    r = bench.c3_workload(lambda msg: None, cpu=True, cpu_seconds=0.3)
    assert "state" not in r
    c = r["cpu_oracle_at_final_size"]
    assert c["node_updates_per_s"] > 0 and c["single_thread"]["node_updates_per_s"] > 0
    assert r["gpu_over_cpu_at_final_size"] == round(1e8 / c["node_updates_per_s"], 1)
    r2 = bench.c3_workload(lambda msg: None, cpu=False)
    assert "cpu_oracle_at_final_size" not in r2
