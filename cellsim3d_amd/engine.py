"""Python binding of libcs3d.so (the C ABI in include/cs3d.h).

This is the host-side mirror of the engine for Python callers and tests: a
Context owns one GPU pool; Systems are tissues inside it.  The binding loads
the in-tree library built by `make -C cellsim3d_amd/csrc` and raises if it is
missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcs3d.so")

_lib = None

vp = C.c_void_p
PP = C.POINTER(abi.Params)
MP = C.POINTER(abi.Mesh)
RP = C.POINTER(abi.RanmarState)
st_t = C.c_int
PHASES = 5
PHASE_NAMES = ("integrate", "search", "con_dis", "dis_post", "divide")

# every symbol declared in include/cs3d.h with its (argtypes, restype)
SIGNATURES = {
    "cs3d_status_str": ([st_t], C.c_char_p),
    "cs3d_last_error": ([], C.c_char_p),
    "cs3d_version": ([], C.c_char_p),
    "cs3d_params_default": ([PP], st_t),
    "cs3d_params_check": ([PP], st_t),
    "cs3d_mesh_build": ([vp, vp, vp, vp, vp, vp, MP], st_t),
    "cs3d_cell_centres": ([vp, vp, vp, C.c_int32, vp, vp, vp], st_t),
    "cs3d_ranmar_init": ([RP, C.c_int32, C.c_int32], st_t),
    "cs3d_ranmar": ([RP, vp, C.c_int32], st_t),
    "cs3d_ctx_create": ([C.c_int, MP, C.c_int64, C.c_int64, C.c_int64, C.POINTER(vp)], st_t),
    "cs3d_ctx_destroy": ([vp], st_t),
    "cs3d_sync": ([vp], st_t),
    "cs3d_system_add": ([vp, PP, C.c_int32, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_int32)],
                        st_t),
    "cs3d_system_add_raw": ([vp, PP, C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_int32)], st_t),
    "cs3d_system_remove": ([vp, C.c_int32], st_t),
    "cs3d_system_grow": ([vp, C.c_int32, C.c_int32, C.c_int64], st_t),
    "cs3d_system_startup": ([vp, C.c_int32], st_t),
    "cs3d_system_set_horizon": ([vp, C.c_int32, C.c_int64], st_t),
    "cs3d_system_info": ([vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint32)], st_t),
    "cs3d_system_set_counts": ([vp, C.c_int32, C.c_int32, C.c_int64], st_t),
    "cs3d_system_save": ([vp, C.c_int32, C.c_char_p], st_t),
    "cs3d_system_load": ([vp, C.c_char_p, C.c_int32, C.c_int64, C.POINTER(C.c_int32)], st_t),
    "cs3d_get": ([vp, C.c_int32, C.c_int, vp, C.c_size_t], st_t),
    "cs3d_set": ([vp, C.c_int32, C.c_int, vp, C.c_size_t], st_t),
    "cs3d_field_bytes": ([vp, C.c_int32, C.c_int], C.c_size_t),
    "cs3d_step": ([vp, C.c_int32, C.c_int32], st_t),
    "cs3d_set_graph": ([vp, C.c_int32], st_t),
    "cs3d_set_graph_steps": ([vp, C.c_int32], st_t),
    "cs3d_set_launch": ([vp, C.c_int32], st_t),
    "cs3d_systems_poll": ([vp, C.c_int32, C.c_int32, vp, vp, vp], st_t),
    "cs3d_systems_work": ([vp, C.c_int32, C.c_int32, vp], st_t),
    "cs3d_systems_set_horizon": ([vp, C.c_int32, C.c_int32, vp], st_t),
    "cs3d_graph_prepare": ([vp, C.c_int32], st_t),
    "cs3d_step_timed": ([vp, C.c_int32, C.c_int32, vp], st_t),
    "cs3d_time_kernel": ([vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)], st_t),
    "cs3d_k_integrate": ([vp, C.c_int32], st_t),
    "cs3d_k_pressure_update": ([vp, C.c_int32, C.c_int64], st_t),
    "cs3d_k_con_force": ([vp, C.c_int32], st_t),
    "cs3d_k_con_force_bonded": ([vp, C.c_int32], st_t),
    "cs3d_k_dis_force": ([vp, C.c_int32, C.c_int32], st_t),
    "cs3d_k_velocity_ab": ([vp, C.c_int32], st_t),
    "cs3d_k_velocity_a": ([vp, C.c_int32], st_t),
    "cs3d_k_velocity_b": ([vp, C.c_int32], st_t),
    "cs3d_k_pressure_reset": ([vp, C.c_int32, vp, C.c_int32, C.c_float], st_t),
    "cs3d_k_pressure_reset_noalloc": ([vp, C.c_int32, vp, C.c_int32, C.c_float], st_t),
    "cs3d_k_bboxes": ([vp, C.c_int32], st_t),
    "cs3d_k_extent": ([vp, C.c_int32, vp], st_t),
    "cs3d_get_state": ([vp, C.c_int32, vp, vp, vp, vp, vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)], st_t),
    "cs3d_k_com": ([vp, C.c_int32], st_t),
    "cs3d_k_volumes": ([vp, C.c_int32, C.c_int64], st_t),
    "cs3d_k_grid": ([vp, C.c_int32], st_t),
    "cs3d_k_divide": ([vp, C.c_int32, C.POINTER(C.c_int32)], st_t),
    "cs3d_copy_calibrate": ([vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)], st_t),
    "cs3d_check_sqrt": ([vp, C.POINTER(C.c_uint64)], st_t),
    "cs3d_check_div": ([vp, C.c_int32, C.c_int64, C.POINTER(C.c_uint64)], st_t),
    "cs3d_device_acos": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64], st_t),
    "cs3d_traj_open": ([C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)], st_t),
    "cs3d_traj_append": ([vp, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp], st_t),
    "cs3d_traj_close": ([vp, C.c_int32], st_t),
}


class EngineError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        msg = lib().cs3d_last_error().decode(errors="replace")
        super().__init__("%s failed: %s (%s)" % (what, lib().cs3d_status_str(status).decode(), msg))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcs3d.so is not built (run __graft_entry__.build() or make -C cellsim3d_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status, what):
    if status != abi.OK:
        raise EngineError(status, what)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(vp)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def default_params():
    p = abi.Params()
    check(lib().cs3d_params_default(C.byref(p)), "cs3d_params_default")
    return p


def build_mesh(tables, rest):
    """abi.Mesh from (nn, sign, faces) device tables and 180 rest positions."""
    nn, sign, faces = (np.ascontiguousarray(t, dtype=np.int32) for t in tables)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float32) for a in rest)
    m = abi.Mesh()
    check(lib().cs3d_mesh_build(_ptr(nn), _ptr(sign), _ptr(faces), _ptr(x), _ptr(y), _ptr(z), C.byref(m)),
          "cs3d_mesh_build")
    return m


def cell_centres(X, Y, Z, n):
    """Centres of n cells on the host, bit-identical to the device CenterOfMass."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, Z))
    if min(len(x), len(y), len(z)) < n * abi.SLOTS:
        raise ValueError("cell_centres: arrays shorter than n * 192")
    out = [np.zeros(max(n, 1), np.float32) for _ in range(3)]
    check(lib().cs3d_cell_centres(_ptr(x), _ptr(y), _ptr(z), int(n), *(_ptr(o) for o in out)), "cs3d_cell_centres")
    return [o[:n] for o in out]


class Context:
    def __init__(self, mesh, capacity_cells, voxel_budget=None, device=0, bin_budget=None):
        if voxel_budget is None:
            voxel_budget = max(1 << 20, capacity_cells * 32 * 40)
        if bin_budget is None:
            bin_budget = max(1 << 22, capacity_cells * 640 * abi.BIN_WORDS)
        h = vp()
        check(lib().cs3d_ctx_create(device, C.byref(mesh), int(capacity_cells), int(voxel_budget), int(bin_budget),
                                    C.byref(h)), "cs3d_ctx_create")
        self.h = h
        self.mesh = mesh
        self.systems = {}

    def close(self):
        if getattr(self, "h", None):
            lib().cs3d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_system(self, params, X, Y, Z, n, max_cells=None, voxels=0, V=None):
        out = C.c_int32()
        X, Y, Z = _f32(X), _f32(Y), _f32(Z)
        vx = vy = vz = None
        if V is not None:
            vx, vy, vz = (_f32(a) for a in V)
        check(lib().cs3d_system_add(self.h, C.byref(params), int(n), int(max_cells or n), int(voxels), _ptr(X),
                                    _ptr(Y), _ptr(Z), _ptr(vx), _ptr(vy), _ptr(vz), C.byref(out)), "cs3d_system_add")
        s = System(self, out.value, params, int(max_cells or n))
        self.systems[s.id] = s
        return s

    def add_system_raw(self, params, n, max_cells=None, voxels=0):
        out = C.c_int32()
        check(lib().cs3d_system_add_raw(self.h, C.byref(params), int(n), int(max_cells or n), int(voxels),
                                        C.byref(out)), "cs3d_system_add_raw")
        s = System(self, out.value, params, int(max_cells or n))
        self.systems[s.id] = s
        return s

    def load_system(self, path, max_cells=0, voxels=0):
        """A system restored from a checkpoint written by System.save."""
        out = C.c_int32()
        check(lib().cs3d_system_load(self.h, str(path).encode(), int(max_cells), int(voxels), C.byref(out)),
              "cs3d_system_load")
        s = System(self, out.value, None, 0)
        s.max_cells = lib().cs3d_field_bytes(self.h, out.value, abi.F_PRESSURE) // 4
        self.systems[s.id] = s
        return s

    def step(self, n_steps, target=abi.ALL):
        sid = target.id if isinstance(target, System) else target
        check(lib().cs3d_step(self.h, sid, int(n_steps)), "cs3d_step")

    def sync(self):
        check(lib().cs3d_sync(self.h), "cs3d_sync")

    def set_graph(self, enable):
        check(lib().cs3d_set_graph(self.h, 1 if enable else 0), "cs3d_set_graph")

    def set_graph_steps(self, steps):
        check(lib().cs3d_set_graph_steps(self.h, int(steps)), "cs3d_set_graph_steps")

    def set_launch(self, wg_per_cu):
        check(lib().cs3d_set_launch(self.h, int(wg_per_cu)), "cs3d_set_launch")

    def poll(self, first, count):
        """(n_cells, steps, status) arrays of systems first .. first+count-1."""
        n = np.zeros(count, dtype=np.int32)
        s = np.zeros(count, dtype=np.int64)
        st = np.zeros(count, dtype=np.uint32)
        check(lib().cs3d_systems_poll(self.h, int(first), int(count), n.ctypes.data_as(vp), s.ctypes.data_as(vp),
                                      st.ctypes.data_as(vp)), "cs3d_systems_poll")
        return n, s, st

    def work(self, first, count):
        """Cell-steps (node-updates / 180) done by systems first .. first+count-1."""
        w = np.zeros(count, dtype=np.int64)
        check(lib().cs3d_systems_work(self.h, int(first), int(count), w.ctypes.data_as(vp)), "cs3d_systems_work")
        return w

    def set_horizons(self, first, horizons):
        """Horizons of systems first .. first+len(horizons)-1 (ids not in use are
        ignored); a horizon at or below a system's step pauses it."""
        h = np.ascontiguousarray(horizons, dtype=np.int64)
        check(lib().cs3d_systems_set_horizon(self.h, int(first), int(h.size), h.ctypes.data_as(vp)),
              "cs3d_systems_set_horizon")

    def prepare_graph(self, target=abi.ALL):
        """Captures and uploads the graph that step() replays for target, without stepping."""
        sid = target.id if isinstance(target, System) else target
        check(lib().cs3d_graph_prepare(self.h, sid), "cs3d_graph_prepare")

    def check_sqrt(self):
        """Differing results of the kernels' float square root against sqrtf over all 2^32 inputs."""
        out = C.c_uint64(0)
        check(lib().cs3d_check_sqrt(self.h, C.byref(out)), "cs3d_check_sqrt")
        return int(out.value)

    def check_div(self, a_significands=1 << 23, scaled_pairs=1 << 30):
        """The kernels' reciprocal and quotients against IEEE division, as six counts:
        reciprocal differing / compared (every float with |b| in [2^-40, 2^40]);
        quotient differing / control over every b in [1, 2) times a_significands
        values of a in [1, 2); the same over scaled_pairs random pairs across the
        admitted range.  A control counts the pairs where the uncorrected product
        RN(a * RN(1/b)) differs from a / b."""
        out = (C.c_uint64 * 6)()
        check(lib().cs3d_check_div(self.h, int(a_significands), int(scaled_pairs), out), "cs3d_check_div")
        return tuple(int(v) for v in out)

    def device_acos(self, x):
        """The kernels' double acos (the angle force's) of each element of x, computed on the device."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        dp = C.POINTER(C.c_double)
        check(lib().cs3d_device_acos(self.h, x.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size),
              "cs3d_device_acos")
        return y

    def copy_calibrate(self, nbytes, reps=10):
        """Average ms of a plain copy kernel over nbytes (reads nbytes, writes nbytes)."""
        out = C.c_double()
        check(lib().cs3d_copy_calibrate(self.h, int(nbytes), int(reps), C.byref(out)), "cs3d_copy_calibrate")
        return out.value

    def step_timed(self, n_steps, target=abi.ALL):
        """Eager steps with HIP events around each phase; returns total ms per phase
        (integrate, search, con_dis, dis_post, divide)."""
        sid = target.id if isinstance(target, System) else target
        out = np.zeros(PHASES, dtype=np.float64)
        check(lib().cs3d_step_timed(self.h, sid, int(n_steps), out.ctypes.data_as(vp)), "cs3d_step_timed")
        return out


class System:
    def __init__(self, ctx, sid, params, max_cells):
        self.ctx = ctx
        self.id = sid
        self.params = params
        self.max_cells = max_cells

    def save(self, path):
        """Checkpoint of the complete state (continues bit-identically after load)."""
        check(lib().cs3d_system_save(self.ctx.h, self.id, str(path).encode()), "cs3d_system_save")

    def startup(self):
        """Start-up forces again from the current state (after per-cell fields were set)."""
        check(lib().cs3d_system_startup(self.ctx.h, self.id), "cs3d_system_startup")

    def remove(self):
        check(lib().cs3d_system_remove(self.ctx.h, self.id), "cs3d_system_remove")
        self.ctx.systems.pop(self.id, None)

    def grow(self, max_cells, voxels=0):
        """Moves the system to a room of max_cells cells (results unchanged)."""
        check(lib().cs3d_system_grow(self.ctx.h, self.id, int(max_cells), int(voxels)), "cs3d_system_grow")
        self.max_cells = int(max_cells)

    def info(self):
        n, st, status = C.c_int32(), C.c_int64(), C.c_uint32()
        check(lib().cs3d_system_info(self.ctx.h, self.id, C.byref(n), C.byref(st), C.byref(status)),
              "cs3d_system_info")
        return n.value, st.value, status.value

    @property
    def n(self):
        return self.info()[0]

    def set_horizon(self, h):
        check(lib().cs3d_system_set_horizon(self.ctx.h, self.id, int(h)), "cs3d_system_set_horizon")

    def set_counts(self, n, step):
        check(lib().cs3d_system_set_counts(self.ctx.h, self.id, int(n), int(step)), "cs3d_system_set_counts")

    def _dtype(self, f):
        if f == abi.F_DIVFLAG:
            return np.int8
        if f in (abi.F_GRID, abi.F_VOXCOUNT, abi.F_VOXLIST, abi.F_PAIRCOUNT, abi.F_PAIRS, abi.F_CANDCOUNT):
            return np.int32
        return np.float32

    def get(self, f):
        if f == abi.F_RANMAR:
            st = abi.RanmarState()
            check(lib().cs3d_get(self.ctx.h, self.id, f, C.addressof(st), C.sizeof(st)), "cs3d_get")
            return st
        nbytes = lib().cs3d_field_bytes(self.ctx.h, self.id, f)
        dt = np.dtype(self._dtype(f))
        out = np.zeros(nbytes // dt.itemsize, dtype=dt)
        check(lib().cs3d_get(self.ctx.h, self.id, f, out.ctypes.data_as(vp), out.nbytes), "cs3d_get")
        return out

    def set(self, f, arr):
        if f == abi.F_RANMAR:
            check(lib().cs3d_set(self.ctx.h, self.id, f, C.addressof(arr), C.sizeof(arr)), "cs3d_set")
            return
        a = np.ascontiguousarray(arr, dtype=self._dtype(f))
        check(lib().cs3d_set(self.ctx.h, self.id, f, a.ctypes.data_as(vp), a.nbytes), "cs3d_set")

    def step(self, n_steps):
        self.ctx.step(n_steps, self.id)

    def time_kernel(self, which, reps=20):
        """Average ms of one launch: which 0 = bonded force, 1 = centre of mass, 2 = a full
        contact-candidate search (the next step searches again)."""
        out = C.c_double()
        check(lib().cs3d_time_kernel(self.ctx.h, self.id, int(which), int(reps), C.byref(out)), "cs3d_time_kernel")
        return out.value

    # per-kernel entry points
    def k_integrate(self):
        check(lib().cs3d_k_integrate(self.ctx.h, self.id), "cs3d_k_integrate")

    def k_pressure_update(self, step):
        check(lib().cs3d_k_pressure_update(self.ctx.h, self.id, int(step)), "cs3d_k_pressure_update")

    def k_con_force(self):
        check(lib().cs3d_k_con_force(self.ctx.h, self.id), "cs3d_k_con_force")

    def k_con_force_bonded(self):
        check(lib().cs3d_k_con_force_bonded(self.ctx.h, self.id), "cs3d_k_con_force_bonded")

    def k_dis_force(self, which=0):
        check(lib().cs3d_k_dis_force(self.ctx.h, self.id, int(which)), "cs3d_k_dis_force")

    def k_velocity_ab(self):
        check(lib().cs3d_k_velocity_ab(self.ctx.h, self.id), "cs3d_k_velocity_ab")

    def k_velocity_a(self):
        check(lib().cs3d_k_velocity_a(self.ctx.h, self.id), "cs3d_k_velocity_a")

    def k_velocity_b(self):
        check(lib().cs3d_k_velocity_b(self.ctx.h, self.id), "cs3d_k_velocity_b")

    def k_pressure_reset(self, cells, value):
        idx = np.ascontiguousarray(cells, dtype=np.int32)
        check(lib().cs3d_k_pressure_reset(self.ctx.h, self.id, idx.ctypes.data_as(vp), len(idx), float(value)),
              "cs3d_k_pressure_reset")

    def k_bboxes(self):
        check(lib().cs3d_k_bboxes(self.ctx.h, self.id), "cs3d_k_bboxes")

    def k_extent(self):
        """{minx, maxx, miny, maxy, minz, maxz} of the cell boxes (minmaxpre/post)."""
        out = np.zeros(6, dtype=np.float32)
        check(lib().cs3d_k_extent(self.ctx.h, self.id, out.ctypes.data_as(vp)), "cs3d_k_extent")
        return out

    def get_state(self):
        """Positions, velocities (each (3, n*192)), pressure, centres (3, n), volumes, n and status."""
        n = self.info()[0]
        xyz = np.zeros((3, n * 192), np.float32)
        v = np.zeros((3, n * 192), np.float32)
        P, V = np.zeros(n, np.float32), np.zeros(n, np.float32)
        CM = np.zeros((3, n), np.float32)
        nc, st = C.c_int32(), C.c_uint32()
        check(lib().cs3d_get_state(self.ctx.h, self.id, xyz.ctypes.data_as(vp), v.ctypes.data_as(vp),
                                   P.ctypes.data_as(vp), CM.ctypes.data_as(vp), V.ctypes.data_as(vp),
                                   C.byref(nc), C.byref(st)), "cs3d_get_state")
        return {"xyz": xyz, "v": v, "P": P, "CM": CM, "V": V, "n": nc.value, "status": st.value}

    def k_com(self):
        check(lib().cs3d_k_com(self.ctx.h, self.id), "cs3d_k_com")

    def k_volumes(self, step):
        check(lib().cs3d_k_volumes(self.ctx.h, self.id, int(step)), "cs3d_k_volumes")

    def k_grid(self):
        check(lib().cs3d_k_grid(self.ctx.h, self.id), "cs3d_k_grid")

    def k_divide(self):
        n = C.c_int32()
        check(lib().cs3d_k_divide(self.ctx.h, self.id, C.byref(n)), "cs3d_k_divide")
        return n.value


class Trajectory:
    """Binary trajectory writer (reference format, read by scripts/celldiv.py)."""

    def __init__(self, path, max_cells, variable_stiffness, max_frames):
        h = vp()
        check(lib().cs3d_traj_open(path.encode(), int(max_cells), int(variable_stiffness), int(max_frames),
                                   C.byref(h)), "cs3d_traj_open")
        self.h = h

    def append(self, step, frame, n_cells, X, Y, Z, types=None):
        X, Y, Z = _f32(X), _f32(Y), _f32(Z)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        check(lib().cs3d_traj_append(self.h, int(step), int(frame), int(n_cells), _ptr(X), _ptr(Y), _ptr(Z),
                                     _ptr(t)), "cs3d_traj_append")

    def close(self, final_cells):
        if self.h:
            check(lib().cs3d_traj_close(self.h, int(final_cells)), "cs3d_traj_close")
            self.h = None
