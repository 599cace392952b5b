"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.  The engine never
imports this module.
"""
import ctypes as C
import os

import numpy as np

from cellsim3d_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# CS3D_ORACLE_LIB points the binding at another build of the oracle (oracle/mutants.py uses it to
# load copies with planted mistakes without touching the real library).
LIB_PATH = os.environ.get("CS3D_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_ranmar.so")

_lib = None
_ref = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
vp = C.c_void_p
PP = C.POINTER(abi.Params)
MP = C.POINTER(abi.Mesh)
RP = C.POINTER(abi.RanmarState)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.or_acos.argtypes = [f64p, f64p, C.c_int]
        L.or_ranmar_init.argtypes = [RP, C.c_int, C.c_int]
        L.or_ranmar_init.restype = C.c_int
        L.or_ranmar.argtypes = [RP, f32p, C.c_int]
        L.or_random_vector.argtypes = [RP, f32p]
        L.or_random_vector_basis.argtypes = [RP, f32p, f32p]
        L.or_mesh_rest.argtypes = [i32p, f32p, f32p, f32p, f32p, f32p]
        L.or_com.argtypes = [C.c_int, f32p, f32p, f32p, f32p, f32p, f32p]
        L.or_volumes.argtypes = [C.c_int, PP, i32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, C.c_int,
                                 f32p, f32p, i8p]
        L.or_volumes.restype = C.c_uint32
        L.or_bboxes.argtypes = [C.c_int, f32p, f32p, f32p, f32p]
        L.or_extent.argtypes = [C.c_int, f32p, f32p]
        L.or_grid_build.argtypes = [C.c_int, PP, f32p, C.c_int32, i32p, i32p, i32p]
        L.or_grid_build.restype = C.c_uint32
        L.or_con_force.argtypes = [C.c_int, PP, MP, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                   i32p, i32p, i32p, C.c_int, f32p, f32p, f32p]
        L.or_con_force.restype = C.c_uint32
        L.or_dis_force.argtypes = [C.c_int, PP, MP, f32p, f32p, f32p, f32p, f32p, f32p, f32p, i32p, i32p,
                                   i32p, f32p, f32p, f32p]
        L.or_dis_force.restype = C.c_uint32
        L.or_integrate.argtypes = [C.c_int, PP] + [f32p] * 12
        L.or_velocity_ab.argtypes = [C.c_int, PP] + [f32p] * 9
        L.or_rgrowth.argtypes = [PP, C.c_int]
        L.or_rgrowth.restype = C.c_float
        L.or_pressure_update.argtypes = [C.c_int, PP, f32p, f32p, vp, C.c_int, C.c_float]
        L.or_sim_create.argtypes = [PP, MP, C.c_int, C.c_int, C.c_int32, vp, vp, vp, C.c_int]
        L.or_sim_create.restype = vp
        L.or_sim_destroy.argtypes = [vp]
        L.or_sim_init.argtypes = [vp]
        L.or_sim_init.restype = C.c_uint32
        L.or_sim_step.argtypes = [vp, C.c_int]
        L.or_sim_step.restype = C.c_uint32
        L.or_sim_phase_seconds.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
        L.or_sim_ncells.argtypes = [vp]
        L.or_sim_stepno.argtypes = [vp]
        L.or_sim_stepno.restype = C.c_int64
        L.or_sim_get.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.or_sim_get.restype = C.c_size_t
        L.or_sim_set.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.or_sim_set.restype = C.c_size_t
        L.or_sim_get_f64.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.or_sim_get_f64.restype = C.c_size_t
        L.or_sim_set_counts.argtypes = [vp, C.c_int, C.c_int64]
        L.or_philox.argtypes = [np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), C.c_uint32, C.c_uint32]
        L.or_ln_unit.argtypes = [C.c_float]
        L.or_ln_unit.restype = C.c_float
        L.or_normals3.argtypes = [C.c_uint32, C.c_uint32, C.c_int64, f32p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference's own RANMAR / GetRandomVector code, compiled from the
    reference tree into oracle/_ref (None when it was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        L.ref_rmarin.argtypes = [C.c_int, C.c_int]
        L.ref_ranmar.argtypes = [f32p, C.c_int]
        L.ref_random_vector.argtypes = [f32p]
        L.ref_random_vector_basis.argtypes = [f32p, f32p]
        _ref = L
    return _ref


class Ranmar:
    def __init__(self, ij, kl):
        self.st = abi.RanmarState()
        if lib().or_ranmar_init(C.byref(self.st), ij, kl) != 0:
            raise ValueError("RANMAR seeds out of range")

    def draw(self, n):
        out = np.zeros(n, dtype=np.float32)
        lib().or_ranmar(C.byref(self.st), out, n)
        return out

    def vector(self):
        out = np.zeros(3, dtype=np.float32)
        lib().or_random_vector(C.byref(self.st), out)
        return out

    def vector_basis(self, basis):
        out = np.zeros(3, dtype=np.float32)
        lib().or_random_vector_basis(C.byref(self.st), out, np.asarray(basis, dtype=np.float32))
        return out


def acos(x):
    """The angle force's double acos (cellsim3d_amd/csrc/cs3d_acos.h) of each element."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().or_acos(x, y, x.size)
    return y


def mesh_rest(nn, x, y, z):
    R0 = np.zeros(3 * abi.SLOTS, dtype=np.float32)
    th = np.zeros(3 * abi.SLOTS, dtype=np.float32)
    lib().or_mesh_rest(np.ascontiguousarray(nn, np.int32), *(np.ascontiguousarray(a, np.float32) for a in (x, y, z)),
                       R0, th)
    return R0, th


def _fptr(a):
    return a.ctypes.data_as(vp) if a is not None else None


class Sim:
    """Full reference step loop on the CPU (float reference-exact or double)."""

    def __init__(self, params, mesh, X, Y, Z, n, max_cells=None, budget=1 << 16, precision=0):
        self.params = params
        self.mesh = mesh
        self.max_cells = max_cells or n
        self.budget = budget
        X, Y, Z = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, Z))
        self.h = lib().or_sim_create(C.byref(params), C.byref(mesh), n, self.max_cells, budget, _fptr(X),
                                     _fptr(Y), _fptr(Z), precision)
        if not self.h:
            raise ValueError("or_sim_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_destroy(self.h)
            self.h = None

    def init(self):
        return lib().or_sim_init(self.h)

    def step(self, k=1):
        return lib().or_sim_step(self.h, k)

    PHASES = ("integrate_pressure", "con_force", "dis_force", "velocity", "com_volumes_division", "grid")

    def phase_seconds(self):
        """Wall seconds per part of the step since the simulation was created (see PHASES)."""
        out = np.zeros(6, dtype=np.float64)
        lib().or_sim_phase_seconds(self.h, out)
        return out

    @property
    def n(self):
        return lib().or_sim_ncells(self.h)

    @property
    def stepno(self):
        return lib().or_sim_stepno(self.h)

    def set_counts(self, n, step):
        lib().or_sim_set_counts(self.h, n, step)

    def field_size(self, f):
        if f in abi.NODE_FIELDS:
            return self.max_cells * abi.SLOTS, np.float32
        if f in abi.CELL_FIELDS:
            return self.max_cells, np.float32
        if f == abi.F_BBOX:
            return self.max_cells * 6, np.float32
        if f == abi.F_DIVFLAG:
            return self.max_cells, np.int8
        if f == abi.F_GRID:
            return 8, np.int32
        if f == abi.F_SYSCM:
            return 3, np.float32
        if f == abi.F_VOXCOUNT:
            return self.budget, np.int32
        if f == abi.F_VOXLIST:
            return self.budget * self.params.voxel_capacity, np.int32
        raise ValueError(f)

    def get(self, f):
        if f == abi.F_RANMAR:
            st = abi.RanmarState()
            lib().or_sim_get(self.h, f, C.addressof(st), C.sizeof(st))
            return st
        cnt, dt = self.field_size(f)
        out = np.zeros(cnt, dtype=dt)
        lib().or_sim_get(self.h, f, out.ctypes.data_as(vp), out.nbytes)
        return out

    def get64(self, f):
        cnt, _ = self.field_size(f)
        out = np.zeros(cnt, dtype=np.float64)
        lib().or_sim_get_f64(self.h, f, out.ctypes.data_as(vp), cnt)
        return out

    def set(self, f, arr):
        if f == abi.F_RANMAR:
            lib().or_sim_set(self.h, f, C.addressof(arr), C.sizeof(arr))
            return
        a = np.ascontiguousarray(arr)
        lib().or_sim_set(self.h, f, a.ctypes.data_as(vp), a.nbytes)


def philox(ctr, key):
    """Philox4x32-10 of a 4-word counter and 2-word key."""
    c = np.array(ctr, dtype=np.uint32)
    lib().or_philox(c, int(key[0]), int(key[1]))
    return c


def normals3(seed, node, step):
    z = np.zeros(3, dtype=np.float32)
    lib().or_normals3(int(seed), int(node), int(step), z)
    return z


def set_threads(n):
    lib().or_set_threads(n)


def get_threads():
    return lib().or_get_threads()


def build_mesh(tables, rest):
    """abi.Mesh from (nn, sign, faces) tables and rest positions (x, y, z of 180 nodes)."""
    nn, sign, faces = tables
    m = abi.Mesh()
    np.ctypeslib.as_array(m.nn)[:] = nn
    np.ctypeslib.as_array(m.sign)[:] = sign
    np.ctypeslib.as_array(m.faces)[:] = faces
    R0, th = mesh_rest(nn, *rest)
    np.ctypeslib.as_array(m.R0)[:] = R0
    np.ctypeslib.as_array(m.theta0)[:] = th
    return m
