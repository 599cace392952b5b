"""The celldiv host driver on the CPU (--setup-only: parameters, mesh files,
placement and trajectory frame 0, no device): inp.json parsing with comments,
read_json_params' validation and exit codes, initialize_C180s' placements."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import REF_ROOT, ROOT

from cellsim3d_amd import configs
from cellsim3d_amd import mesh as M
from oracle import oracle as O

CELLDIV = os.path.join(ROOT, "cellsim3d_amd", "celldiv")

INP = """{
    // core parameters (comments as jsoncpp accepts them)
    "core": {
        "MaxNoofC180s": 100, "particle_mass": 0.04, "repulsion_range": 0.2, "attraction_range": 0.3,
        "repulsion_strength": 100.0, "attraction_strength": 0.5, "Youngs_mod": 1000, "stiffFactor1": 1.0,
        "viscotic_damping": 1, "internal_damping": 100.0, "gamma_visc": 1.0, "division_Vol": 2.9,
        "random_z_offset?": 0, "z_offset": 0, "div_time_steps": 200, "time_interval": 0.0001,
        "Restart": 0, "trajWriteInt": 100, "non_div_time_steps": 100, "trajFileName": "t.xyz",
        "binaryOutput": 1, "maxPressure": 65.0, "minPressure": 50.0, "growth_rate": 0.002,
        "checkSphericity": 1, "constrainAngles": 1, "doAdaptive_dt": 0, "dt_max": 1e-4, "dt_tol": 1e-8,
        "phase_count": 2000000000, "write_cont_force": 0, "forces_file": "f.csv", "correct_com": 0
    },
    "counting": {"countCells": 0, "mit-index_file_name": "m.dat", "count_only_internal_cells?": 0,
                 "radius_cutoff": 0.6, "overwrite_mit_ind_file?": 0, "cell_count_int": 100},
    "population": {"doPopModel": 0, "totalFood": 5000.0},
    "walls": { /* deprecated */ "useWalls": 0, "perpAxis": "z", "threshDist": 0.01},
    "divParams": {"useDivPlaneBasis": 0, "divPlaneBasisX": 0, "divPlaneBasisY": 0, "divPlaneBasisZ": 1},
    "stiffnessParams": {"useDifferentStiffnesses": 0, "softStiffFactor": 0.8, "numberOfSofterCells": 0,
                        "fractionOfSofterCells": 0.3, "daughtSameStiffness": 1, "closenessToCenter": 0,
                        "recalc_r0": 0},
    "boxParams": {"useRigidSimulationBox": 1, "usePBCs": 0, "box_len_x": 60.0, "box_len_y": 60.0,
                  "box_len_z": 4.0, "flatbox": 1, "rand_pos": 0},
    "rand_params": {"add_rands": 0, "rand_seed": 7, "rand_dist": 0, "rand_scale_factor": 0.001}
}
"""


def _strip_comments(text):
    out, i = [], 0
    while i < len(text):
        if text.startswith("//", i):
            i = text.find("\n", i)
            i = len(text) if i < 0 else i
        elif text.startswith("/*", i):
            i = text.find("*/", i) + 2
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def _setup(tmp_path, edit=None, n=9, mesh=None):
    d = json.loads(_strip_comments(INP))
    if edit:
        edit(d)
    text = INP
    if edit:  # re-serialise the edited parameters (comments are covered by the unedited cases)
        text = json.dumps(d)
    (tmp_path / "inp.json").write_text(text)
    M.write_reference_files(str(tmp_path), mesh or configs.unit_mesh())
    r = subprocess.run([CELLDIV, str(n), "inp.json", "0", "--setup-only"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=60)
    return r, d


def _read_frames(path):
    """All frames of a binary trajectory (GPUbounce.cu:1007-1018, 2446-2477):
    header, then per frame (step, frame, n, X, Y, Z, types or None)."""
    b = open(path, "rb").read()
    hdr = struct.unpack_from("<3i", b, 0)
    var = hdr[1]
    off = 12
    frames = []
    while off < len(b):
        step, frame, n = struct.unpack_from("<3i", b, off)
        off += 12
        X, Y, Z = (np.zeros((n, 192), np.float32) for _ in range(3))
        types = np.zeros(n, np.int32) if var else None
        for c in range(n):
            (idx,) = struct.unpack_from("<i", b, off)
            assert idx == c
            off += 4
            for A in (X, Y, Z):
                A[c] = np.frombuffer(b, dtype="<f4", count=192, offset=off)
                off += 768
            if var:
                (types[c],) = struct.unpack_from("<i", b, off)
                off += 4
        frames.append((step, frame, n, X, Y, Z, types))
    assert off == len(b)
    return hdr, frames


def _read_traj(path):
    hdr, frames = _read_frames(path)
    assert len(frames) == 1
    step, frame, n, X, Y, Z, _ = frames[0]
    return hdr, (step, frame, n), X, Y, Z


pytestmark = pytest.mark.skipif(not os.path.exists(CELLDIV), reason="celldiv not built")


def _centred_mesh():
    x = configs.unit_mesh()["x"].astype(np.float32)
    s = [np.float32(0)] * 3
    for a in range(3):
        acc = np.float32(0)
        for v in x[:, a]:
            acc = np.float32(acc + v)
        s[a] = np.float32(np.float64(acc) / 180.0)
    return x - np.array(s, np.float32)


def test_lattice_placement_in_the_rigid_box(tmp_path):
    r, d = _setup(tmp_path, n=9)
    assert r.returncode == 0, r.stdout + r.stderr
    hdr, (step, frame, n), X, Y, Z = _read_traj(tmp_path / "t.xyz")
    assert hdr == (9, 0, (200 + 100 + 1) // 100) and (step, frame, n) == (0, 1, 9)
    base = _centred_mesh()
    for c in range(9):
        ey, ex = c % 3, c // 3
        shift = np.array([3.0 * ex + 1.5 + 30.0, 3.0 * ey + 1.5 + 30.0, 2.0])
        got = np.stack([X[c, :180], Y[c, :180], Z[c, :180]], 1)
        assert np.allclose(got, base + shift, atol=4e-6)  # a translated copy of the centred mesh
        assert np.all(X[c, 180:] == 0) and np.all(Y[c, 180:] == 0) and np.all(Z[c, 180:] == 0)
    for f in ("psfil", "velocity2.xyz"):
        assert (tmp_path / f).exists() and (tmp_path / f).stat().st_size == 0


def test_random_placement_follows_the_ranmar_stream(tmp_path):
    def edit(d):
        d["boxParams"]["rand_pos"] = 1
        d["boxParams"]["flatbox"] = 0
        d["boxParams"]["box_len_z"] = 12.0
        d["boxParams"]["box_len_x"] = 12.0
        d["boxParams"]["box_len_y"] = 12.0
    r, d = _setup(tmp_path, edit, n=12)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, X, Y, Z = _read_traj(tmp_path / "t.xyz")
    # centres drawn from RANMAR(7, 700) with rejection (GPUbounce.cu:1672-1712)
    rng = O.Ranmar(7, 700)
    rcheck = np.float32(np.power(np.float32(0.75 * (1.0 / np.float32(3.14159)) * 0.786), np.float32(1 / 3)))
    bm = np.float32(12.0)
    cms = []
    while len(cms) < 12:
        u = rng.draw(3)
        cm = np.array([u[0] * (bm - 1) + 1, u[1] * (bm - 1) + 1, u[2] * (bm - 1) + 1], np.float32)
        ok = not np.any(cm + rcheck > bm)
        ok = ok and all(np.sqrt(np.sum((c - cm) ** 2)) >= 2 * rcheck for c in cms)
        if ok:
            cms.append(cm)
    base = _centred_mesh()
    for c in range(12):
        got = np.stack([X[c, :180], Y[c, :180], Z[c, :180]], 1)
        assert np.allclose(got, base + cms[c], atol=4e-6)
    cm = np.array(cms)
    dist = np.linalg.norm(cm[:, None] - cm[None], axis=-1) + np.eye(12) * 10
    assert dist.min() >= 2 * rcheck * (1 - 1e-6)


@pytest.mark.parametrize("edit,code", [
    (lambda d: d["core"].__setitem__("trajWriteInt", 70), 255),        # div_time_steps % trajWriteInt != 0
    (lambda d: d["core"].__setitem__("trajWriteInt", 1000), 255),      # interval longer than the run
    (lambda d: d["boxParams"].update(box_len_x=1.5, box_len_y=1.5, box_len_z=1.5), 27),  # floor(V/0.786) < n
    (lambda d: d["stiffnessParams"].__setitem__("closenessToCenter", 2.0), 255),
    (lambda d: d["walls"].__setitem__("useWalls", 1), 255),            # walls and rigid box together
    (lambda d: d["rand_params"].__setitem__("rand_dist", 1), 255),     # only the uniform default exists
    (lambda d: d["core"].__setitem__("doAdaptive_dt", 1), 255),        # not provided
    (lambda d: d.pop("divParams"), 255),
    (lambda d: d["core"].update(dt_tol=1.0), 196),                     # exit(-60)
])
def test_validation_exit_codes(tmp_path, edit, code):
    r, _ = _setup(tmp_path, edit)
    assert r.returncode == code, r.stdout


def test_too_many_cells_prints_usage(tmp_path):
    r, _ = _setup(tmp_path, n=101)
    assert r.returncode == 0 and "Usage" in r.stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_ROOT, "inp.json")), reason="reference inp.json absent")
def test_reads_the_reference_parameter_file(tmp_path):
    """The reference's own inp.json (with // and /* */ comments) parses."""
    import shutil
    shutil.copy(os.path.join(REF_ROOT, "inp.json"), tmp_path / "inp.json")
    M.write_reference_files(str(tmp_path), configs.unit_mesh())
    r = subprocess.run([CELLDIV, "16", "inp.json", "0", "--setup-only"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert "parsed successfully" in r.stdout and "16 cells placed" in r.stdout


# ------------------------------------------------------------------ soft cells, correct_com, walls
def _lattice_centres(n, side, box=60.0):
    c = np.array([[3.0 * (k // side) + 1.5 + box / 2, 3.0 * (k % side) + 1.5 + box / 2] for k in range(n)])
    return c


def _soft(edit_stf, rigid=1):
    def edit(d):
        d["stiffnessParams"].update(useDifferentStiffnesses=1, **edit_stf)
        d["boxParams"]["useRigidSimulationBox"] = rigid
    return edit


def test_soft_cells_within_a_radius_of_the_centre(tmp_path):
    """GPUbounce.cu:623-659: cells whose centre lies within closenessToCenter of the
    largest centre distance (about the mean centre) become soft; the frame carries
    the cell types."""
    r, _ = _setup(tmp_path, _soft({"closenessToCenter": 0.5}), n=16)
    assert r.returncode == 0, r.stdout
    hdr, frames = _read_frames(tmp_path / "t.xyz")
    assert hdr[1] == 1
    types = frames[0][6]
    c = _lattice_centres(16, 4)
    d = np.linalg.norm(c - c.mean(0), axis=1)
    assert np.array_equal(types, (d / d.max() <= 0.5).astype(np.int32))
    assert types.sum() == 4  # the inner 2 x 2 block of the 4 x 4 lattice


def test_soft_cells_first_ones_in_a_rigid_box(tmp_path):
    """GPUbounce.cu:661-681: round(fraction * n) cells; the first ones in a rigid box."""
    r, _ = _setup(tmp_path, _soft({"fractionOfSofterCells": 0.3}), n=16)
    assert r.returncode == 0, r.stdout
    _, frames = _read_frames(tmp_path / "t.xyz")
    assert list(frames[0][6]) == [1] * 5 + [0] * 11


def test_soft_cells_drawn_from_the_ranmar_stream(tmp_path):
    """GPUbounce.cu:682-689: without a rigid box each soft cell index is
    round(u * n) for the next RANMAR number (repeats allowed)."""
    def edit(d):
        _soft({"fractionOfSofterCells": 0.0, "numberOfSofterCells": 6}, rigid=0)(d)
    r, _ = _setup(tmp_path, edit, n=16)
    assert r.returncode == 0, r.stdout
    _, frames = _read_frames(tmp_path / "t.xyz")
    rng = O.Ranmar(7, 700)
    want = np.zeros(16, np.int32)
    for u in rng.draw(6):
        i = int(np.round(np.float32(u) * np.float32(16)))
        if i < 16:
            want[i] = 1
    assert np.array_equal(frames[0][6], want)


def test_too_many_soft_cells_is_an_error(tmp_path):
    r, _ = _setup(tmp_path, _soft({"fractionOfSofterCells": 0.0, "numberOfSofterCells": 17}), n=16)
    assert r.returncode == 12516 % 256 and "Too many softer cells" in r.stdout


def test_correct_com_centres_the_start_positions(tmp_path):
    """GPUbounce.cu:785-806: the start positions minus the mean of the cell
    centres (float sums in cell order); mesh nodes only."""
    r0, _ = _setup(tmp_path, n=9)
    _, _, X0, Y0, Z0 = _read_traj(tmp_path / "t.xyz")
    r1, _ = _setup(tmp_path, lambda d: d["core"].__setitem__("correct_com", 1), n=9)
    assert r0.returncode == 0 and r1.returncode == 0, r1.stdout
    _, _, X1, Y1, Z1 = _read_traj(tmp_path / "t.xyz")
    from cellsim3d_amd import engine
    cm = engine.cell_centres(X0.ravel(), Y0.ravel(), Z0.ravel(), 9)
    for A0, A1, c in zip((X0, Y0, Z0), (X1, Y1, Z1), cm):
        acc = np.float32(0)
        for v in c:
            acc = np.float32(acc + v)
        sc = np.float32(acc / np.float32(9))
        assert np.array_equal(A1[:, :180], (A0[:, :180] - sc).astype(np.float32))
        assert np.all(A1[:, 180:] == 0)
        assert abs(np.asarray(A1[:, :180], np.float64).mean()) < 1e-5


@pytest.mark.parametrize("walls,code", [
    ({"useWalls": 1, "perpAxis": "x", "dAxis": 2.0}, 255),   # only z walls exist
    ({"useWalls": 1, "perpAxis": "z", "dAxis": 0.01}, 255),  # the gap is narrower than the cells
    ({"useWalls": 1, "perpAxis": "Z", "dAxis": 5.0}, 0),
])
def test_wall_setup_checks(tmp_path, walls, code):
    """GPUbounce.cu:808-861 (walls need a free box: the rigid box and walls exclude
    each other)."""
    def edit(d):
        d["walls"].update(walls)
        d["boxParams"]["useRigidSimulationBox"] = 0
    r, _ = _setup(tmp_path, edit, n=9)
    assert r.returncode == code, r.stdout


def test_dump_setup_writes_the_engine_inputs(tmp_path):
    """--dump-setup writes the parameter block and mesh tables handed to the engine
    (the bytes of cs3d_params and cs3d_mesh), also with --setup-only."""
    import ctypes as C
    from cellsim3d_amd import abi
    d = json.loads(_strip_comments(INP))
    _soft({"fractionOfSofterCells": 0.3})(d)
    (tmp_path / "inp.json").write_text(json.dumps(d))
    M.write_reference_files(str(tmp_path), configs.unit_mesh())
    r = subprocess.run([CELLDIV, "9", "inp.json", "0", "--setup-only", "--dump-setup", "s"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    pb, mb = (tmp_path / "s.params").read_bytes(), (tmp_path / "s.mesh").read_bytes()
    assert len(pb) == C.sizeof(abi.Params) and len(mb) == C.sizeof(abi.Mesh)
    p = abi.Params.from_buffer_copy(pb)
    assert p.use_different_stiffnesses == 1 and p.use_rigid_box == 1
    assert p.stiffness2 == np.float32(0.8 * 1000.0) and p.stiffness1 == np.float32(1000.0)
    assert p.ranmar_ij == 7 and p.ranmar_kl == 700
    m = abi.Mesh.from_buffer_copy(mb)
    nn = np.ctypeslib.as_array(m.nn)
    assert nn.min() >= 0 and nn.max() < 180
