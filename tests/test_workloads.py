"""Workload generators of the ensemble and serving runners, on the CPU.

The batched ensemble draws 256 tissues with cell counts log-uniform in [1, 512]
and bond stiffness, growth rate and adhesion each within +-30 % of the defaults;
the serving workload draws Poisson(20/s) arrivals with cell counts log-uniform in
[8, 4096] and horizons log-uniform in [1e3, 1e5] steps.  The draws are seeded, so
the totals quoted with the measurements (20,232 cells for the ensemble) are fixed.
"""
import numpy as np


def test_ensemble_tissues_follow_the_stated_distributions():
    from cellsim3d_amd import configs
    base = configs.base_params(1)
    _, tissues = configs.ensemble(n_tissues=256, seed=1, nmin=1, nmax=512, steps=1000)
    sizes = np.array([t["n"] for t in tissues])
    assert len(tissues) == 256 and sizes.min() >= 1 and sizes.max() <= 512
    assert sizes.sum() == 20232
    # log-uniform: the median is near sqrt(512) = 22.6, far below the uniform median of 256
    assert 10 <= np.median(sizes) <= 50
    for key in ("stiffness1", "growth_rate", "attraction_strength"):
        f = np.array([getattr(t["params"], key) for t in tissues]) / getattr(base, key)
        assert 0.7 <= f.min() and f.max() <= 1.3, key
        assert f.max() - f.min() > 0.4, key  # the factors spread over the range
    for t in tissues:
        assert t["params"].div_time_steps == 1000 and t["X"].size >= t["n"] * 192


def test_ensemble_room_sets_the_cell_slots_of_each_tissue():
    from cellsim3d_amd import ensemble
    for room in (2, 5):
        _, tissues = ensemble.build(n_tissues=8, seed=2, nmin=1, nmax=40, steps=100, grid_interval=7, room=room)
        for t in tissues:
            assert t["max_cells"] == room * t["n"] + 2
            assert t["params"].grid_interval == 7


def test_serving_requests_follow_the_stated_distributions():
    from cellsim3d_amd import configs, serving
    req = configs.serving_requests(n_requests=400, seed=3, rate=20.0)
    t = np.array([r["arrival"] for r in req])
    n = np.array([r["n"] for r in req])
    h = np.array([r["horizon"] for r in req])
    assert np.all(np.diff(t) > 0)
    # Poisson arrivals at 20/s: mean gap 0.05 s (400 gaps: standard error 0.0025 s)
    assert abs(np.diff(t).mean() - 0.05) < 0.01
    assert n.min() >= 8 and n.max() <= 4096 and h.min() >= 1000 and h.max() <= 100000
    jobs = serving.make_jobs(req[:3])
    for j, r in zip(jobs, req[:3]):
        assert (j.n, j.horizon) == (r["n"], r["horizon"])
        assert j.params.growth_rate == 0.0 and j.max_cells >= j.n + 2
        assert j.bytes() > j.max_cells * 180 * 4
