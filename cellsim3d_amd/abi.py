"""ctypes mirrors of the structs and constants declared in include/cs3d.h.

These are plain data layouts shared by the engine binding (cellsim3d_amd.engine)
and the CPU oracle's test binding; nothing here computes.
"""
import ctypes as C

NODES = 180
SLOTS = 192
BIN_CAP = 32  # CS3D_BIN_CAP
BIN_WORDS = 1 + 4 * BIN_CAP  # int32 words per search bin
FACES = 92
ALL = -1

# cs3d_status
OK, INVALID_ARG, OUT_OF_MEMORY, CAPACITY, NAN_DETECTED, LIST_OVERFLOW, HIP_ERROR, NO_DEVICE, UNSUPPORTED = range(9)

# sticky per-system status bits
SYS_NAN = 1 << 0
SYS_VOXEL_OVERFLOW = 1 << 1
SYS_GRID_TOO_LARGE = 1 << 2
SYS_CAPACITY = 1 << 3
SYS_ANGLE_NONFINITE = 1 << 4
SYS_VOLUME_NONFINITE = 1 << 5
SYS_BAD_NORMAL = 1 << 6
SYS_NN_OVER_10 = 1 << 7
SYS_CAND_OVERFLOW = 1 << 8
SYS_BIN_OVERFLOW = 1 << 9
SYS_FATAL_MASK = 0x37F

# cs3d_field
(F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_FCX, F_FCY, F_FCZ, F_FDX, F_FDY, F_FDZ, F_VAX, F_VAY, F_VAZ,
 F_CMX, F_CMY, F_CMZ, F_BBOX, F_PRESSURE, F_STIFFNESS, F_VOLUME, F_AREA, F_DIVFLAG, F_GRID,
 F_VOXCOUNT, F_VOXLIST, F_RANMAR, F_PAIRCOUNT, F_PAIRS, F_GROWTH_SCALE, F_FRX, F_FRY, F_FRZ, F_SYSCM, F_CANDCOUNT,
 F__COUNT) = range(37)

NODE_FIELDS = {F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_FCX, F_FCY, F_FCZ, F_FDX, F_FDY, F_FDZ, F_VAX, F_VAY, F_VAZ,
               F_FRX, F_FRY, F_FRZ}
CELL_FIELDS = {F_CMX, F_CMY, F_CMZ, F_PRESSURE, F_STIFFNESS, F_VOLUME, F_AREA, F_GROWTH_SCALE}


class Mesh(C.Structure):
    _fields_ = [
        ("nn", C.c_int32 * (3 * SLOTS)),
        ("sign", C.c_int32 * SLOTS),
        ("faces", C.c_int32 * (FACES * 7)),
        ("R0", C.c_float * (3 * SLOTS)),
        ("theta0", C.c_float * (3 * SLOTS)),
    ]


class Params(C.Structure):
    _fields_ = [
        ("mass", C.c_float),
        ("repulsion_range", C.c_float),
        ("attraction_range", C.c_float),
        ("repulsion_strength", C.c_float),
        ("attraction_strength", C.c_float),
        ("youngs_mod", C.c_float),
        ("stiffness1", C.c_float),
        ("stiffness2", C.c_float),
        ("viscotic_damping", C.c_float),
        ("internal_damping", C.c_float),
        ("gamma_visc", C.c_float),
        ("division_vol", C.c_float),
        ("dt", C.c_float),
        ("max_pressure", C.c_float),
        ("min_pressure", C.c_float),
        ("growth_rate", C.c_float),
        ("check_sphericity", C.c_int32),
        ("constrain_angles", C.c_int32),
        ("div_time_steps", C.c_int32),
        ("non_div_time_steps", C.c_int32),
        ("phase_count", C.c_int32),
        ("use_rigid_box", C.c_int32),
        ("box_max", C.c_float * 3),
        ("thresh_dist", C.c_float),
        ("use_different_stiffnesses", C.c_int32),
        ("recalc_r0", C.c_int32),
        ("daught_same_stiffness", C.c_int32),
        ("use_div_plane_basis", C.c_int32),
        ("div_plane_basis", C.c_float * 3),
        ("ranmar_ij", C.c_int32),
        ("ranmar_kl", C.c_int32),
        ("do_pop_model", C.c_int32),
        ("max_pop", C.c_float),
        ("add_rands", C.c_int32),
        ("angle_k", C.c_float),
        ("grid_dl", C.c_float),
        ("voxel_capacity", C.c_int32),
        ("grid_interval", C.c_int32),
        ("daughter_com_fix", C.c_int32),
        ("pair_skin", C.c_float),
        ("rand_scale_factor", C.c_float),
        ("noise_seed", C.c_uint32),
        ("correct_com", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]

    def copy(self):
        other = Params()
        C.memmove(C.byref(other), C.byref(self), C.sizeof(Params))
        return other

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = list(v)
            out[name] = v
        return out


class RanmarState(C.Structure):
    _fields_ = [
        ("u", C.c_float * 98),
        ("c", C.c_float),
        ("cd", C.c_float),
        ("cm", C.c_float),
        ("i97", C.c_int32),
        ("j97", C.c_int32),
        ("ready", C.c_int32),
    ]


def default_params():
    """inp.json defaults (reference inp.json:1-108), with the engine's extra fields
    at their reference values: angle_k 1000 (propagate.cu:335), DL 2.9
    (GPUbounce.cu:756), voxel capacity 32 (propagatebound.cu:259), grid rebuilt
    every step."""
    p = Params()
    p.mass = 0.04
    p.repulsion_range = 0.2
    p.attraction_range = 0.3
    p.repulsion_strength = 100.0
    p.attraction_strength = 0.5
    p.youngs_mod = 1000.0
    p.stiffness1 = 1.0 * 1000.0
    p.stiffness2 = 0.8 * 1000.0
    p.viscotic_damping = 1.0
    p.internal_damping = 100.0
    p.gamma_visc = 1.0
    p.division_vol = 2.9
    p.dt = 0.0001
    p.max_pressure = 65.0
    p.min_pressure = 50.0
    p.growth_rate = 0.002
    p.check_sphericity = 1
    p.constrain_angles = 1
    p.div_time_steps = 50000
    p.non_div_time_steps = 50000
    p.phase_count = 2000000000
    p.use_rigid_box = 1
    p.box_max[:] = [500.0, 500.0, 2.0]
    p.thresh_dist = 0.01
    p.use_different_stiffnesses = 0
    p.recalc_r0 = 0
    p.daught_same_stiffness = 1
    p.use_div_plane_basis = 0
    p.div_plane_basis[:] = [0.0, 0.0, 1.0]
    p.ranmar_ij = 1
    p.ranmar_kl = 100
    p.do_pop_model = 0
    p.max_pop = 0.0
    p.add_rands = 0
    p.angle_k = 1000.0
    p.grid_dl = 2.9
    p.voxel_capacity = 32
    p.grid_interval = 1
    p.daughter_com_fix = 0
    p.pair_skin = 0.05
    p.rand_scale_factor = 0.001
    p.noise_seed = 1
    return p
