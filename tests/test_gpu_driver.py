"""The celldiv host driver end to end on the GPU: mitotic-index file
(GPUbounce.cu:1288-1302, 1515-1547), soft cells chosen when growth ends
(:1340-1429) and counting of internal cells only (:2537-2565).  The cell counts
the driver books are checked against the counts in its own trajectory frames."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import test_driver as D

pytestmark = pytest.mark.gpu

CELLDIV = os.path.join(ROOT, "cellsim3d_amd", "celldiv")


def _run(tmp_path, edit, n):
    from cellsim3d_amd import configs
    v0 = configs.mesh_volume(configs.unit_mesh())

    def full_edit(d):
        d["core"].update(division_Vol=float(2.0 * v0), div_time_steps=600, non_div_time_steps=50, trajWriteInt=2,
                         MaxNoofC180s=256)
        d["counting"].update({"cell_count_int": 100, "overwrite_mit_ind_file?": 1})
        edit(d)
    d = json.loads(D._strip_comments(D.INP))
    full_edit(d)
    (tmp_path / "inp.json").write_text(json.dumps(d))
    from cellsim3d_amd import mesh as M
    M.write_reference_files(str(tmp_path), configs.unit_mesh())
    print("celldiv %d cells in %s" % (n, tmp_path), flush=True)
    r = subprocess.run([CELLDIV, str(n), "inp.json", "0"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    hdr, frames = D._read_frames(tmp_path / "t.xyz")
    counts = {f[0]: f[2] for f in frames}
    mi = [float(x) for x in (tmp_path / "m.dat").read_text().split()]
    return d, frames, counts, mi, r.stdout


def _expected_mi(counts, steps=600, every=100):
    out = []
    for k in range(1, steps // every + 1):
        new = counts[k * every] - counts[(k - 1) * every]
        tot = counts[k * every] - new
        out.append(float("%f" % (np.float32(new) / np.float32(tot))) if new and tot else 0.0)
    return out + [0.0]


def test_mitotic_index_and_soft_cells_after_growth(tmp_path):
    def edit(d):
        d["stiffnessParams"].update(useDifferentStiffnesses=1, duringGrowth=0, closenessToCenter=0.01,
                                    numberOfSofterCells=2, fractionOfSofterCells=0.0, chooseRandomCellIndices=0,
                                    daughtSameStiffness=1)
    d, frames, counts, mi, out = _run(tmp_path, edit, 4)
    assert counts[600] > counts[0], "no division: the test does not exercise the counting"
    assert mi == _expected_mi(counts)
    # no soft cell at set-up (every centre is at the largest distance); at step 602
    # the first two cells within (1 - 0.01) of the largest distance about the stale
    # system centre (1, 1, 1) become soft
    by_step = {f[0]: f for f in frames}
    assert by_step[0][6].sum() == 0 and by_step[600][6].sum() == 0
    from cellsim3d_amd import engine
    step, _, n, X, Y, Z, _ = by_step[602]
    cm = engine.cell_centres(X.ravel(), Y.ravel(), Z.ravel(), n)
    r2max = np.float32(0)
    for c in range(n):
        dx, dy, dz = (np.float32(cm[a][c] - np.float32(1)) for a in range(3))
        r2max = max(r2max, np.float32(np.float32(dx * dx + dy * dy) + dz * dz))
    sc = [np.float32(0)] * 3
    for a in range(3):
        for c in range(n):
            sc[a] = np.float32(sc[a] + cm[a][c])
        sc[a] = np.float32(sc[a] / np.float32(n))
    f = np.float32(1) - np.float32(0.01)
    want = np.zeros(n, np.int32)
    for c in range(n):
        dx, dy, dz = (np.float32(cm[a][c] - sc[a]) for a in range(3))
        if np.float32(np.float32(dx * dx + dy * dy) + dz * dz) <= np.float32(f * f) * r2max and want.sum() < 2:
            want[c] = 1
    assert want.sum() == 2
    assert np.array_equal(by_step[604][6], want) and np.array_equal(frames[-1][6], want)
    assert "Cell growth halted" in out


def test_counting_only_internal_cells(tmp_path):
    """With count_only_internal_cells? a new cell counts only if its centre is
    within radius_cutoff of the largest distance; a daughter's centre is 0 in
    the step that made it, so in a colony far from the origin every daughter is
    far and the index stays 0, while plain counting sees the divisions."""
    d, frames, counts, mi, _ = _run(tmp_path, lambda d: None, 64)
    assert counts[600] > counts[0] >= 50
    assert mi == _expected_mi(counts) and max(mi) > 0
    d, frames, counts2, mi2, _ = _run(tmp_path, lambda d: d["counting"].update(
        {"count_only_internal_cells?": 1, "radius_cutoff": 0.6}), 64)
    assert counts2 == counts
    assert mi2 == [0.0] * 7


@pytest.mark.parametrize("binary", [1, 0])
def test_output_thread_writes_the_same_bytes(tmp_path, binary):
    """Frames written by the writer thread (the default) and in place
    (--sync-output) are byte-identical: trajectory (binary or ASCII) and forces
    CSV, frames in order."""
    from cellsim3d_amd import configs
    from cellsim3d_amd import mesh as M
    v0 = configs.mesh_volume(configs.unit_mesh())
    d = json.loads(D._strip_comments(D.INP))
    d["core"].update(division_Vol=float(1.05 * v0), div_time_steps=200, non_div_time_steps=0, trajWriteInt=20,
                     write_cont_force=1, binaryOutput=binary)
    outs = []
    for flag in ([], ["--sync-output"]):
        wd = tmp_path / ("sync" if flag else "async")
        wd.mkdir()
        (wd / "inp.json").write_text(json.dumps(d))
        M.write_reference_files(str(wd), configs.unit_mesh())
        print("celldiv 9 cells %s" % (flag or "async"), flush=True)
        r = subprocess.run([CELLDIV, "9", "inp.json", "0"] + flag, cwd=str(wd), capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:]
        outs.append(((wd / "t.xyz").read_bytes(), (wd / "f.csv").read_bytes()))
    assert outs[0] == outs[1]
    lines = outs[0][1].decode().splitlines()
    steps = sorted({int(l.split(",")[0]) for l in lines[1:]})
    assert steps == list(range(0, 201, 20))
    if binary:
        _, frames = D._read_frames(tmp_path / "async" / "t.xyz")
        assert [f[0] for f in frames] == list(range(0, 201, 20))


def test_soft_cells_at_setup_enter_the_start_up_forces(tmp_path):
    """Soft cells chosen at set-up are on the device before the start-up forces: the
    reference uploads d_Youngs_mod (GPUbounce.cu:691) before its initial
    CalculateConForce (:978-1004).  The driver's state after 4 steps (--save-state)
    equals, bit for bit, an engine run from the driver's own parameters, mesh and frame 0
    (--dump-setup) with the stiffness set before the start-up (System.startup).  The same
    run with the start-up forces taken at stiffness1 ends with other velocities, so the
    comparison sees the difference."""
    from cellsim3d_amd import abi, configs, engine
    from cellsim3d_amd import mesh as M
    v0 = configs.mesh_volume(configs.unit_mesh())
    d = json.loads(D._strip_comments(D.INP))
    d["core"].update(division_Vol=float(2.0 * v0), div_time_steps=4, non_div_time_steps=0, trajWriteInt=1)
    d["counting"]["cell_count_int"] = 2
    D._soft({"fractionOfSofterCells": 0.3})(d)
    (tmp_path / "inp.json").write_text(json.dumps(d))
    M.write_reference_files(str(tmp_path), configs.unit_mesh())
    print("celldiv 16 cells, 5 soft, 4 steps", flush=True)
    r = subprocess.run([CELLDIV, "16", "inp.json", "0", "--dump-setup", "setup", "--save-state", "end.cs3d"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = abi.Params.from_buffer_copy((tmp_path / "setup.params").read_bytes())
    mesh = abi.Mesh.from_buffer_copy((tmp_path / "setup.mesh").read_bytes())
    _, frames = D._read_frames(tmp_path / "t.xyz")
    step0, _, n, X, Y, Z, types = frames[0]
    assert step0 == 0 and n == 16 and types.sum() == 5
    mc = d["core"]["MaxNoofC180s"]
    Ecell = np.where(types == 1, p.stiffness2, p.stiffness1).astype(np.float32)
    ctx = engine.Context(mesh, capacity_cells=3 * mc + 16)
    drv = ctx.load_system(str(tmp_path / "end.cs3d"))
    runs = []
    for startup in (True, False):
        s = ctx.add_system(p, X.ravel(), Y.ravel(), Z.ravel(), n, max_cells=mc)
        s.set(abi.F_STIFFNESS, Ecell)
        if startup:
            s.startup()
        s.step(4)
        runs.append(s)
    ctx.sync()
    assert drv.info()[:2] == (n, 4) and runs[0].info()[:2] == (n, 4)
    for f in (abi.F_X, abi.F_Y, abi.F_Z, abi.F_VX, abi.F_VY, abi.F_VZ, abi.F_FCX, abi.F_FDX, abi.F_STIFFNESS):
        a, b = drv.get(f), runs[0].get(f)
        k = n * 192 if f in abi.NODE_FIELDS else n
        assert np.array_equal(a[:k], b[:k]), "field %d differs from the engine run" % f
    assert not all(np.array_equal(drv.get(f)[:n * 192], runs[1].get(f)[:n * 192])
                   for f in (abi.F_VX, abi.F_VY, abi.F_VZ))
    ctx.close()


def _ascii_frames(path):
    """Frames of an ASCII trajectory (GPUbounce.cu:2412-2444): (step, n, text of
    each cell's 192 position lines)."""
    lines = path.read_text().splitlines()
    out, i = [], lines.index("Header End") + 1  # the header (GPUbounce.cu:1019-1032) comes first
    while i < len(lines):
        n = int(lines[i]) // 192
        step = int(lines[i + 1].split()[1])
        cells = []
        i += 2
        for c in range(n):
            cells.append(lines[i + 1:i + 1 + 192])
            i += 193
        out.append((step, n, cells))
    return out


@pytest.mark.parametrize("binary", [1, 0])
def test_outputs_match_oracle(tmp_path, binary):
    """The driver's output files against the CPU oracle run from the driver's own
    parameters, mesh and frame 0 (--dump-setup): every trajectory frame (binary
    floats bit for bit, or the ASCII text line for line), every forces-CSV row
    (GPUbounce.cu:2567-2598: the total conservative force, its magnitude, the
    velocity and its magnitude, position, pressure and volume, as text) and the
    mitotic-index file (GPUbounce.cu:1515-1547) from the oracle's cell counts."""
    from cellsim3d_amd import abi, configs
    from cellsim3d_amd import mesh as M
    from oracle import oracle as O
    v0 = configs.mesh_volume(configs.unit_mesh())
    d = json.loads(D._strip_comments(D.INP))
    d["core"].update(division_Vol=float(1.05 * v0), div_time_steps=300, non_div_time_steps=0, trajWriteInt=50,
                     write_cont_force=1, binaryOutput=binary, MaxNoofC180s=64)
    d["counting"].update({"cell_count_int": 100, "overwrite_mit_ind_file?": 1})
    (tmp_path / "inp.json").write_text(json.dumps(d))
    M.write_reference_files(str(tmp_path), configs.unit_mesh())
    print("celldiv 9 cells, 300 steps, binary %d" % binary, flush=True)
    r = subprocess.run([CELLDIV, "9", "inp.json", "0", "--dump-setup", "setup"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = abi.Params.from_buffer_copy((tmp_path / "setup.params").read_bytes())
    mesh = abi.Mesh.from_buffer_copy((tmp_path / "setup.mesh").read_bytes())
    mc = d["core"]["MaxNoofC180s"]
    X, Y, Z = (np.fromfile(str(tmp_path / ("setup." + a)), dtype=np.float32) for a in "xyz")
    n = len(X) // 192
    frames = D._read_frames(tmp_path / "t.xyz")[1] if binary else _ascii_frames(tmp_path / "t.xyz")
    assert frames[0][0] == 0 and n == 9
    vcells = max(1 << 18, 32 * mc)
    sim = O.Sim(p, mesh, X, Y, Z, n, max_cells=mc, budget=vcells)
    sim.init()
    csv = (tmp_path / "f.csv").read_text().splitlines()
    assert csv[0].startswith("step,num_cells,cell_ind")
    rows = {}
    for line in csv[1:]:
        rows.setdefault(int(line.split(",", 1)[0]), []).append(line)
    counts = {}
    done = 0
    for k, fr in enumerate(frames):
        step = fr[0]
        if step > done:
            sim.step(step - done)
            done = step
        ns = sim.n
        counts[step] = ns
        Xo, Yo, Zo = (sim.get(f)[:ns * 192] for f in (abi.F_X, abi.F_Y, abi.F_Z))
        if binary:
            _, _, nf, Xf, Yf, Zf, _ = fr
            assert nf == ns
            for a, b in ((Xf, Xo), (Yf, Yo), (Zf, Zo)):
                assert np.array_equal(a.ravel(), b), "frame at step %d differs from the oracle" % step
        else:
            _, nf, cells = fr
            assert nf == ns
            want = ["%.7f,  %.7f,  %.7f" % (Xo[c * 192 + q], Yo[c * 192 + q], Zo[c * 192 + q])
                    for c in range(ns) for q in range(192)]
            assert [l for c in cells for l in c] == want, "ASCII frame at step %d differs" % step
        F = [sim.get(f)[:ns * 192].reshape(ns, 192) for f in (abi.F_FCX, abi.F_FCY, abi.F_FCZ, abi.F_VX, abi.F_VY,
                                                                 abi.F_VZ, abi.F_X, abi.F_Y, abi.F_Z)]
        P, V = sim.get(abi.F_PRESSURE)[:ns], sim.get(abi.F_VOLUME)[:ns]
        want = []
        for c in range(ns):
            for q in range(180):
                fx, fy, fz, vx, vy, vz, x, y, z = (a[c, q] for a in F)
                fm = np.sqrt(fx * fx + fy * fy + fz * fz)
                vm = np.sqrt(vx * vx + vy * vy + vz * vz)
                want.append("%d,%d,%d,%d,%d,%s" % (step, ns, c, q, c * 192 + q, ",".join(
                    "%.8f" % v for v in (fx, fy, fz, fm, vx, vy, vz, vm, x, y, z, P[c], V[c]))))
        assert rows[step] == want, "forces CSV rows of step %d differ from the oracle" % step
    assert counts[300] > counts[0], "no division: the test does not exercise the counting"
    mi = [float(x) for x in (tmp_path / "m.dat").read_text().split()]
    assert mi == _expected_mi(counts, steps=300, every=100)
