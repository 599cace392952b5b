/*
 * cs3d.h - C ABI of the MI355X-native CellSim3D engine (libcs3d.so).
 *
 * The reference (SoftSimu/CellSim3D) has no library API: its hot path is a
 * set of CUDA kernels launched from main() in src/GPUbounce.cu with raw
 * device pointers.  Each entry point below names the reference code it
 * replaces.  Conventions:
 *   - plain C, no exceptions cross the boundary; every call returns a
 *     cs3d_status and cs3d_last_error() holds a message for the calling thread;
 *   - node arrays are SoA float32 with a 192-slot stride per cell
 *     (x[cell*192 + node], nodes 180..191 are padding), exactly the layout of
 *     d_X/d_Y/d_Z in src/GPUbounce.cu:365-373;
 *   - the context owns all device memory; caller buffers are host memory and
 *     are copied, never retained;
 *   - a context is used from one host thread and runs on one HIP stream.
 */
#ifndef CS3D_H
#define CS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS3D_NODES 180   /* real nodes per C180 cell */
#define CS3D_SLOTS 192   /* node slots per cell (stride), GPUbounce.cu:365 */
#define CS3D_BIN_CAP 32  /* nodes per contact-search bin; a bin takes 1 + 4 * CS3D_BIN_CAP int32 words of the
                            context's bin budget */
#define CS3D_FACES 92    /* 12 pentagons + 80 hexagons, GPUbounce.cu:1858-1886 */
#define CS3D_ALL (-1)    /* cs3d_step target: every active system in one batch */

typedef enum cs3d_status {
    CS3D_OK = 0,
    CS3D_INVALID_ARG = 1,
    CS3D_OUT_OF_MEMORY = 2,
    CS3D_CAPACITY = 3,       /* population or pool capacity exceeded (GPUbounce.cu:1264-1270 returns -69) */
    CS3D_NAN_DETECTED = 4,   /* reference traps: propagate.cu:241-248, volume.cu:134-139 */
    CS3D_LIST_OVERFLOW = 5,  /* voxel list / candidate list capacity exceeded */
    CS3D_HIP_ERROR = 6,
    CS3D_NO_DEVICE = 7,
    CS3D_UNSUPPORTED = 8     /* e.g. add_rands = 1 (cuRAND stream, not reproduced) */
} cs3d_status;

/* Sticky per-system status bits (replace the reference's device traps). */
#define CS3D_SYS_NAN              (1u << 0)  /* NaN node position (propagate.cu:241-248) */
#define CS3D_SYS_VOXEL_OVERFLOW   (1u << 1)  /* > voxel_capacity cells in a voxel (propagatebound.cu:247-259 is unchecked) */
#define CS3D_SYS_GRID_TOO_LARGE   (1u << 2)  /* Xdiv*Ydiv*Zdiv beyond the system's voxel budget */
#define CS3D_SYS_CAPACITY         (1u << 3)  /* division past max_cells (GPUbounce.cu:1264-1270) */
#define CS3D_SYS_ANGLE_NONFINITE  (1u << 4)  /* propagate.cu:48-52 */
#define CS3D_SYS_VOLUME_NONFINITE (1u << 5)  /* volume.cu:134-139 */
#define CS3D_SYS_BAD_NORMAL       (1u << 6)  /* postscriptinit.cu:215-219 */
#define CS3D_SYS_NN_OVER_10       (1u << 7)  /* diagnostic: > 10 candidate cells for a node; the reference
                                                writes past localNNs[10] there (propagate.cu:222,398-403) */
#define CS3D_SYS_CAND_OVERFLOW    (1u << 8)  /* a node has more contact candidates (nodes of other cells within
                                                attraction_range + pair_skin) than its list holds (40) */
#define CS3D_SYS_BIN_OVERFLOW     (1u << 9)  /* more nodes in one contact-search bin than it holds (CS3D_BIN_CAP) */
#define CS3D_SYS_FATAL_MASK       (0x37Fu)   /* every bit except the NN_OVER_10 diagnostic */

/* Mesh topology and rest state shared by all cells.
 * nn[k*192+i] (k=0..2) and sign[i]: C180NN, 0-based (GPUbounce.cu:1826-1833).
 * faces[f*7+j]: C180_pentahexa rows; pentagons padded N1..N5,N1,N1, hexagons
 *   N1..N6,N1 (GPUbounce.cu:1861-1886).
 * R0[k*192+i]: rest length of bond (i, nn[k][i]) (GPUbounce.cu:1892-1908).
 * theta0[3*i+{0,1,2}]: rest angles aij, ajk, aik of node i (GPUbounce.cu:909-938,
 *   struct angles3 in inc/VectorFunctions.hpp:10-12). */
typedef struct cs3d_mesh {
    int32_t nn[3 * CS3D_SLOTS];
    int32_t sign[CS3D_SLOTS];
    int32_t faces[CS3D_FACES * 7];
    float R0[3 * CS3D_SLOTS];
    float theta0[3 * CS3D_SLOTS];
} cs3d_mesh;

/* Physical and numerical parameters of one system (tissue).  Names follow the
 * inp.json keys read by read_json_params (GPUbounce.cu:1914-2247). */
typedef struct cs3d_params {
    float mass;                 /* core.particle_mass */
    float repulsion_range;      /* core.repulsion_range */
    float attraction_range;     /* core.attraction_range */
    float repulsion_strength;   /* core.repulsion_strength */
    float attraction_strength;  /* core.attraction_strength */
    float youngs_mod;           /* core.Youngs_mod (contact and wall forces use this global value) */
    float stiffness1;           /* core.stiffFactor1 * Youngs_mod (GPUbounce.cu:1953) */
    float stiffness2;           /* stiffnessParams.softStiffFactor * Youngs_mod (GPUbounce.cu:2048) */
    float viscotic_damping;     /* gamma_ext, inter-cell friction */
    float internal_damping;     /* gamma_int, bond damping */
    float gamma_visc;           /* gamma_o, drag */
    float division_vol;         /* core.division_Vol */
    float dt;                   /* core.time_interval */
    float max_pressure;         /* core.maxPressure */
    float min_pressure;         /* core.minPressure */
    float growth_rate;          /* core.growth_rate (rMax) */
    int32_t check_sphericity;   /* core.checkSphericity */
    int32_t constrain_angles;   /* core.constrainAngles */
    int32_t div_time_steps;     /* core.div_time_steps (Time_steps) */
    int32_t non_div_time_steps; /* core.non_div_time_steps */
    int32_t phase_count;        /* core.phase_count */
    int32_t use_rigid_box;      /* boxParams.useRigidSimulationBox */
    float box_max[3];           /* boxParams.box_len_x/y/z */
    float thresh_dist;          /* walls.threshDist */
    int32_t use_different_stiffnesses; /* stiffnessParams.useDifferentStiffnesses */
    int32_t recalc_r0;          /* stiffnessParams.recalc_r0 */
    int32_t daught_same_stiffness; /* stiffnessParams.daughtSameStiffness */
    int32_t use_div_plane_basis;   /* divParams.useDivPlaneBasis */
    float div_plane_basis[3];      /* divParams.divPlaneBasisX/Y/Z */
    int32_t ranmar_ij, ranmar_kl;  /* RANMAR seeds (GPUbounce.cu:1777-1810: ij = rand_seed, kl = 100*rand_seed) */
    int32_t do_pop_model;          /* population.doPopModel */
    float max_pop;                 /* population.max_pop */
    int32_t add_rands;             /* rand_params.add_rands: random force (propagate.cu:667-676, 681-725) */
    float angle_k;                 /* bending constant, 1000 in the reference (propagate.cu:335) */
    float grid_dl;                 /* voxel edge DL, 2.9 in the reference (GPUbounce.cu:756) */
    int32_t voxel_capacity;        /* cells per voxel list, 32 in the reference (propagatebound.cu:259) */
    int32_t grid_interval;         /* rebuild bbox + voxel grid every k steps (1 = reference) */
    int32_t daughter_com_fix;      /* 0 = reference: a new daughter's centre of mass stays 0 until the next
                                      CenterOfMass; 1 = set it from the daughter's nodes at division */
    float pair_skin;               /* engine: contact candidates are searched within attraction_range +
                                      pair_skin and re-searched once a node has moved pair_skin/2; every step
                                      applies the exact reference tests, so results do not depend on it */
    float rand_scale_factor;       /* rand_params.rand_scale_factor: scale of the normal random force */
    uint32_t noise_seed;           /* key of the random-force stream.  The reference draws from cuRAND
                                      (closed, not reproduced); this engine draws standard normals with the
                                      Marsaglia polar method from Philox4x32-10 (Salmon et al., SC'11),
                                      counter (cell * 192 + node, step), reproducible per tissue, node and step */
    int32_t correct_com;           /* core.correct_com: at the end of every step the system centre (mean of
                                      the cell centres, summed in cell order) is subtracted from every node
                                      position (GPUbounce.cu:1431-1441, propagate.cu:766-776) */
    int32_t reserved[3];
} cs3d_params;

/* Fields readable/writable with cs3d_get / cs3d_set.  Sizes are per system
 * with n = max_cells of the system (node fields: n*192 floats). */
typedef enum cs3d_field {
    CS3D_F_X = 0, CS3D_F_Y, CS3D_F_Z,          /* positions d_X/d_Y/d_Z */
    CS3D_F_VX, CS3D_F_VY, CS3D_F_VZ,           /* velocities d_velListX/Y/Z (end of step) */
    CS3D_F_FCX, CS3D_F_FCY, CS3D_F_FCZ,        /* conservative force d_fConList */
    CS3D_F_FDX, CS3D_F_FDY, CS3D_F_FDZ,        /* dissipative force d_fDisList */
    CS3D_F_VAX, CS3D_F_VAY, CS3D_F_VAZ,        /* velocity after Integrate (read by the first DisForce) */
    CS3D_F_CMX, CS3D_F_CMY, CS3D_F_CMZ,        /* per-cell centre of mass d_CMx/y/z */
    CS3D_F_BBOX,                               /* 6 floats per cell: minx maxx miny maxy minz maxz */
    CS3D_F_PRESSURE,                           /* d_pressList */
    CS3D_F_STIFFNESS,                          /* d_Youngs_mod (per-cell spring constant) */
    CS3D_F_VOLUME, CS3D_F_AREA,                /* d_volume, d_area */
    CS3D_F_DIVFLAG,                            /* int8 per cell, d_cell_div */
    CS3D_F_GRID,                               /* 8 x int32: float minx, miny, minz (bit patterns), xdiv, ydiv, zdiv,
                                                  voxel budget, voxel capacity */
    CS3D_F_VOXCOUNT,                           /* int32 per voxel (budget entries), d_NoofNNlist */
    CS3D_F_VOXLIST,                            /* int32 budget*capacity, d_NNlist (each list ascending) */
    CS3D_F_RANMAR,                             /* cs3d_ranmar_state */
    CS3D_F_PAIRCOUNT,                          /* int32 per node: stored contact partners (engine-specific) */
    CS3D_F_PAIRS,                              /* int32 40 x 192 per cell: partner keys cell*192 + node,
                                                  ascending, k-major (engine-specific) */
    CS3D_F_GROWTH_SCALE,                       /* per-cell multiplier of the pressure increment (default 1) */
    CS3D_F_FRX, CS3D_F_FRY, CS3D_F_FRZ,        /* random force of the latest step, d_fRanList (add_rands) */
    CS3D_F_SYSCM,                              /* 3 floats, read-only: the system centre subtracted at the end
                                                  of the latest step (sysCMx/y/z; correct_com), else 0 */
    CS3D_F_CANDCOUNT,                          /* int32 per node, read-only: contact candidates kept at
                                                  the last search (engine-specific) */
    CS3D_F__COUNT
} cs3d_field;

/* RANMAR generator state (src/marsaglia.h:50-52). */
typedef struct cs3d_ranmar_state {
    float u[98];
    float c, cd, cm;
    int32_t i97, j97;
    int32_t ready;
} cs3d_ranmar_state;

typedef struct cs3d_ctx cs3d_ctx;
typedef int32_t cs3d_sys;

/* ---- diagnostics -------------------------------------------------------- */
const char* cs3d_status_str(cs3d_status s);
const char* cs3d_last_error(void);
const char* cs3d_version(void);

/* ---- host-side helpers (no device needed) -------------------------------- */
/* Defaults of inp.json (the file shipped with the reference, inp.json:1-108). */
cs3d_status cs3d_params_default(cs3d_params* p);
/* Validation rules of read_json_params (GPUbounce.cu:2168-2244) that apply to the kernel path. */
cs3d_status cs3d_params_check(const cs3d_params* p);
/* Builds R0 and theta0 from the 180 rest positions x/y/z (one cell, as placed),
 * replacing read_fullerene_nn's R0 loop (GPUbounce.cu:1892-1908) and the theta0
 * loop (GPUbounce.cu:909-938). nn is C180_nn (3*192, 0-based), sign 180 ints,
 * faces 92*7 ints. */
cs3d_status cs3d_mesh_build(const int32_t* nn, const int32_t* sign, const int32_t* faces,
                            const float* x, const float* y, const float* z, cs3d_mesh* out);
/* CenterOfMass (centermass.cu:8-50) of n_cells cells on the host: the same
 * 256-slot summation tree as the device kernel, so the results are bit-identical
 * to CS3D_F_CMX/Y/Z after a device CenterOfMass.  x/y/z hold 192 slots per cell.
 * Used by the host driver where the reference copies d_CMx to the host at set-up
 * (soft-cell selection, GPUbounce.cu:623-640; correct_com, :785-792). */
cs3d_status cs3d_cell_centres(const float* x, const float* y, const float* z, int32_t n_cells, float* cmx,
                              float* cmy, float* cmz);
/* RANMAR: rmarin (marsaglia.h:54-113) and ranmar (marsaglia.h:115-144). */
cs3d_status cs3d_ranmar_init(cs3d_ranmar_state* st, int32_t ij, int32_t kl);
cs3d_status cs3d_ranmar(cs3d_ranmar_state* st, float* out, int32_t n);

/* ---- context ------------------------------------------------------------ */
/* Allocates a pool of capacity_cells cell slots, a voxel pool of voxel_budget
 * int32 words (voxel counts and lists of all systems) and a search-bin pool of
 * bin_budget int32 words on `device`.  Replaces the cudaMalloc block of main
 * (GPUbounce.cu:358-615). */
cs3d_status cs3d_ctx_create(int device, const cs3d_mesh* mesh, int64_t capacity_cells,
                            int64_t voxel_budget, int64_t bin_budget, cs3d_ctx** out);
cs3d_status cs3d_ctx_destroy(cs3d_ctx* ctx);
cs3d_status cs3d_sync(cs3d_ctx* ctx);

/* ---- systems ------------------------------------------------------------ */
/* Adds a system of n_cells cells (room for max_cells), uploads positions
 * (n_cells*192 per component; velocities nullable = 0), sets pressures to
 * min_pressure and stiffness to stiffness1, then runs the reference's start-up
 * sequence: CenterOfMass, bounding_boxes, minmaxpre/post, makeNNlist
 * (GPUbounce.cu:739-763), initial CalculateConForce and CalculateDisForce
 * (GPUbounce.cu:978-1004).  voxels = voxel budget of this system (0 = auto). */
cs3d_status cs3d_system_add(cs3d_ctx* ctx, const cs3d_params* p, int32_t n_cells, int32_t max_cells,
                            int64_t voxels, const float* x, const float* y, const float* z,
                            const float* vx, const float* vy, const float* vz, cs3d_sys* out);
/* Adds a system without running the start-up sequence (state injection for
 * per-kernel parity); fields are then set with cs3d_set. */
cs3d_status cs3d_system_add_raw(cs3d_ctx* ctx, const cs3d_params* p, int32_t n_cells, int32_t max_cells,
                                int64_t voxels, cs3d_sys* out);
cs3d_status cs3d_system_remove(cs3d_ctx* ctx, cs3d_sys sys);
/* Moves a system to a larger room of max_cells cell slots (voxels: its voxel
 * budget, 0 = auto, never below the current one) and frees the old slots.  Cell
 * indices, results and the system id are unchanged.  For batched ensembles and
 * served jobs whose tissues outgrow their room between steps; the reference
 * sizes every array for MaxNoofC180s up front (GPUbounce.cu:358-415) and has no
 * counterpart.  CAPACITY when the pools cannot hold the new room (the system is
 * then left as it was).  Synchronous. */
cs3d_status cs3d_system_grow(cs3d_ctx* ctx, cs3d_sys sys, int32_t max_cells, int64_t voxels);
/* Runs the start-up sequence of cs3d_system_add again on the system's current
 * state: CenterOfMass, bounding boxes, grid, initial CalculateConForce and
 * CalculateDisForce (GPUbounce.cu:739-763, 978-1004).  For callers that change
 * per-cell state between cs3d_system_add (or _raw plus cs3d_set) and the first
 * step: the reference uploads the per-cell stiffness d_Youngs_mod
 * (GPUbounce.cu:691) before its initial forces, so soft cells chosen at set-up
 * enter the start-up forces.  Synchronous. */
cs3d_status cs3d_system_startup(cs3d_ctx* ctx, cs3d_sys sys);
/* Limits a system to `horizon` total steps (stepping stops there; -1 = none). */
cs3d_status cs3d_system_set_horizon(cs3d_ctx* ctx, cs3d_sys sys, int64_t horizon);
/* Sets the horizons of systems first .. first+count-1 at once (entries of ids not in
 * use are ignored).  A horizon at or below a system's current step pauses it: the
 * step kernels skip it until a later horizon resumes it, and its state stays on the
 * device unchanged, so a paused and resumed system ends where an uninterrupted run
 * ends.  For schedulers that choose which resident systems advance (the serving
 * loop steps the jobs with the least remaining work).  No counterpart in the
 * reference, which steps one tissue.  Synchronous. */
cs3d_status cs3d_systems_set_horizon(cs3d_ctx* ctx, int32_t first, int32_t count, const int64_t* horizons);
cs3d_status cs3d_system_info(cs3d_ctx* ctx, cs3d_sys sys, int32_t* n_cells, int64_t* step, uint32_t* status);
/* Sets the live cell count and step counter (state injection). */
cs3d_status cs3d_system_set_counts(cs3d_ctx* ctx, cs3d_sys sys, int32_t n_cells, int64_t step);
/* Checkpoint / restart of one system's complete state (positions, velocities,
 * forces, per-cell state, stale centres of mass, the contact-rule grid,
 * RANMAR stream, counters).  A system loaded into a context with the same mesh
 * continues bit-identically to the saved one.  The reference reads a Restart
 * key but never implements it (GPUbounce.cu:1963); this has no counterpart.
 * load: max_cells <= 0 keeps the saved capacity; voxels as cs3d_system_add. */
cs3d_status cs3d_system_save(cs3d_ctx* ctx, cs3d_sys sys, const char* path);
cs3d_status cs3d_system_load(cs3d_ctx* ctx, const char* path, int32_t max_cells, int64_t voxels, cs3d_sys* out);
cs3d_status cs3d_get(cs3d_ctx* ctx, cs3d_sys sys, cs3d_field f, void* host, size_t bytes);
cs3d_status cs3d_set(cs3d_ctx* ctx, cs3d_sys sys, cs3d_field f, const void* host, size_t bytes);
size_t cs3d_field_bytes(cs3d_ctx* ctx, cs3d_sys sys, cs3d_field f);
/* The state a caller of the reference reads back after a run, in one call:
 * positions xyz and velocities v as three SoA blocks of n_cells*192 floats
 * (x block, then y, then z), pressure P[n_cells], centres of mass CM as three
 * blocks of n_cells floats, volumes V[n_cells], the live cell count and the
 * sticky status.  Every pointer may be NULL; the buffers are sized for the
 * system's current cell count (cs3d_system_info).  Replaces the cudaMemcpy
 * block before the writers (GPUbounce.cu:1007-1048). */
cs3d_status cs3d_get_state(cs3d_ctx* ctx, cs3d_sys sys, float* xyz, float* v, float* P, float* CM, float* V,
                           int32_t* n_cells, uint32_t* status);

/* ---- stepping ----------------------------------------------------------- */
/* Runs n_steps of the reference step loop (GPUbounce.cu:1055-1506) on one
 * system (plain per-system path) or on every active system (sys = CS3D_ALL,
 * batched path: one launch per kernel for all systems, captured in a HIP
 * graph).  Asynchronous: returns after enqueueing; cs3d_sync waits. */
cs3d_status cs3d_step(cs3d_ctx* ctx, cs3d_sys sys, int32_t n_steps);
/* Enables (1) or disables (0) HIP graph capture for cs3d_step. */
cs3d_status cs3d_set_graph(cs3d_ctx* ctx, int32_t enable);
/* Number of steps captured per HIP graph (default 10). */
cs3d_status cs3d_set_graph_steps(cs3d_ctx* ctx, int32_t steps);
/* Launch shape of the cell kernels: at most 256 x wg_per_cu workgroups (one
 * cell each; beyond that a workgroup strides over several cells).  Default 16.
 * No counterpart in the reference (one block per cell, GPUbounce.cu:1059-1164);
 * set by the ensemble autotuner. */
cs3d_status cs3d_set_launch(cs3d_ctx* ctx, int32_t wg_per_cu);
/* Cell counts, completed steps and status bits of systems first .. first+count-1
 * in one device read (-1 / 0 for unused ids); for the serving loop, which has
 * no counterpart in the reference. */
cs3d_status cs3d_systems_poll(cs3d_ctx* ctx, int32_t first, int32_t count, int32_t* n_cells, int64_t* steps,
                              uint32_t* status);
/* Work done by systems first .. first+count-1: for each, the sum over its
 * completed steps of the cells it stepped (node-updates / 180; -1 for unused
 * ids), counted on the device at the end of every step, so growing systems are
 * counted with their actual cell counts.  No counterpart in the reference. */
cs3d_status cs3d_systems_work(cs3d_ctx* ctx, int32_t first, int32_t count, int64_t* cell_steps);
/* Captures, instantiates and uploads the HIP graph that cs3d_step replays for
 * `sys` (or CS3D_ALL) without running it, so that a timed region holds replays
 * only.  The all-systems graph reads the systems in use on the device and stays
 * valid when systems are added or removed.  No counterpart in the reference. */
cs3d_status cs3d_graph_prepare(cs3d_ctx* ctx, cs3d_sys sys);

/* Phases of one step, in launch order (one kernel launch each; the division phase adds
 * CorrectCoMMotion for systems with correct_com). */
#define CS3D_PHASES 5
#define CS3D_PH_INTEGRATE 0  /* Integrate + ForwardTime + PressureUpdate; adoption of the bboxes and
                                voxel counts of a grid rebuilt at the end of the previous step;
                                search-bin insertion on steps that search */
#define CS3D_PH_SEARCH 1     /* contact-candidate lists (only on steps that search) */
#define CS3D_PH_CON_DIS 2    /* CalculateConForce + first CalculateDisForce + VelocityUpdateA/B */
#define CS3D_PH_DIS_POST 3   /* second CalculateDisForce + CenterOfMass + volumes + bounding_boxes */
#define CS3D_PH_DIVIDE 4     /* division phase, end of step, extent + grid geometry on rebuild steps */
/* Runs n_steps eagerly and adds each phase's kernel time in milliseconds to
 * phase_ms[CS3D_PHASES]: every launch records start/stop HIP events on the
 * context's stream (hipExtLaunchKernelGGL), so launch gaps are not counted (no
 * counterpart in the reference, which has no timers; it replaces profiling by
 * hand). Synchronous. */
cs3d_status cs3d_step_timed(cs3d_ctx* ctx, cs3d_sys sys, int32_t n_steps, double* phase_ms);
/* Average milliseconds of one launch of a kernel on the system's state, from
 * per-launch start/stop HIP events over reps launches on the context's stream: which = 0 the
 * bonded force (cs3d_k_con_force_bonded), 1 CenterOfMass (cs3d_k_com), 2 a full
 * contact-candidate search of every cell (with the search bins filled from the current
 * positions before and cleared after; the next step searches again).  For roofline
 * measurements; no counterpart in the reference. */
cs3d_status cs3d_time_kernel(cs3d_ctx* ctx, cs3d_sys sys, int32_t which, int32_t reps, double* avg_ms);

/* ---- per-kernel entry points (state injection / parity) ------------------
 * Each runs one reference kernel (or fixed kernel group) on the device state
 * of one system and returns after it completes. */
/* Integrate + ForwardTime (propagate.cu:681-712, 738-763). */
cs3d_status cs3d_k_integrate(cs3d_ctx* ctx, cs3d_sys sys);
/* PressureUpdate (PressureKernels.cu:4-41) for step `step` with rGrowth from
 * the population model (GPUbounce.cu:1077-1088). */
cs3d_status cs3d_k_pressure_update(cs3d_ctx* ctx, cs3d_sys sys, int64_t step);
/* CalculateConForce (propagate.cu:183-520) -> FC. */
cs3d_status cs3d_k_con_force(cs3d_ctx* ctx, cs3d_sys sys);
/* Spring + pressure (+ angle) part of CalculateConForce only (propagate.cu:253-337) -> FC. */
cs3d_status cs3d_k_con_force_bonded(cs3d_ctx* ctx, cs3d_sys sys);
/* CalculateDisForce (propagate.cu:523-665) with velocities `which` (0 = V, 1 = VA) -> FD. */
cs3d_status cs3d_k_dis_force(cs3d_ctx* ctx, cs3d_sys sys, int32_t which);
/* VelocityUpdateA then VelocityUpdateB (propagate.cu:714-736) on V. */
cs3d_status cs3d_k_velocity_ab(cs3d_ctx* ctx, cs3d_sys sys);
/* VelocityUpdateA alone (propagate.cu:714-725): V += 0.5 (dt F_con + sqrt(dt) F_ran) / m, in double. */
cs3d_status cs3d_k_velocity_a(cs3d_ctx* ctx, cs3d_sys sys);
/* VelocityUpdateB alone (propagate.cu:727-736): V += 0.5 dt F_dis / m, in double. */
cs3d_status cs3d_k_velocity_b(cs3d_ctx* ctx, cs3d_sys sys);
/* PressureReset (PressureKernels.cu:45-65): the pressure of cells[0 .. count) (local
 * indices, each below the live cell count; INVALID_ARG otherwise) set to value.
 * The reference also overwrites its index list with -1; the host list here is
 * only read. */
cs3d_status cs3d_k_pressure_reset(cs3d_ctx* ctx, cs3d_sys sys, const int32_t* cells, int32_t count,
                                  float value);
/* PressureReset as cs3d_k_pressure_reset, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The cell list is
 * staged in a fixed device array of 65536 entries that the library holds on
 * each GPU until the process ends (cs3d_k_pressure_reset grows a context buffer
 * with hipMalloc instead); longer lists are refused with INVALID_ARG.  The call
 * allocates nothing and returns after the reset completes.  Calls that overlap
 * on one GPU (contexts on other host threads) share the array and race. */
cs3d_status cs3d_k_pressure_reset_noalloc(cs3d_ctx* ctx, cs3d_sys sys, const int32_t* cells, int32_t count,
                                          float value);
/* CenterOfMass (centermass.cu:8-50). */
cs3d_status cs3d_k_com(cs3d_ctx* ctx, cs3d_sys sys);
/* volumes (volume.cu:5-163) for step `step`: VOLUME, AREA, DIVFLAG. */
cs3d_status cs3d_k_volumes(cs3d_ctx* ctx, cs3d_sys sys, int64_t step);
/* bounding_boxes alone (propagatebound.cu:10-141): the box of each live cell's
 * nodes {minx, maxx, miny, maxy, minz, maxz} -> CS3D_F_BBOX.  The voxel ranges of
 * the contact rule follow the boxes only at the next cs3d_k_grid (as makeNNlist
 * follows bounding_boxes in the reference step). */
cs3d_status cs3d_k_bboxes(cs3d_ctx* ctx, cs3d_sys sys);
/* minmaxpre + minmaxpost (propagatebound.cu:145-204, 266-323) over the boxes in
 * CS3D_F_BBOX: out6 = {minx, maxx, miny, maxy, minz, maxz} from the sentinels
 * +-1e8 (GPUbounce.cu:1331 copies the same six floats to the host). */
cs3d_status cs3d_k_extent(cs3d_ctx* ctx, cs3d_sys sys, float* out6);
/* bounding_boxes + minmaxpre/post + grid dims + makeNNlist
 * (propagatebound.cu:10-323, GPUbounce.cu:1312-1338) with ascending voxel lists. */
cs3d_status cs3d_k_grid(cs3d_ctx* ctx, cs3d_sys sys);
/* Division phase: count_and_get_div, RANMAR normals, cell_division,
 * PressureReset and stiffness bookkeeping (GPUbounce.cu:1222-1286,
 * postscriptinit.cu:148-256, PressureKernels.cu:45-65). Returns events via
 * n_events (nullable). */
cs3d_status cs3d_k_divide(cs3d_ctx* ctx, cs3d_sys sys, int32_t* n_events);

/* ---- measurement helpers (no counterpart in the reference) ---------------
 * Copies `bytes` bytes between two scratch buffers of that size with a plain
 * float4 copy kernel (k_copy_calibrate), `reps` times, and returns the average
 * launch time in ms: the known-size kernel used to scale HBM byte counters.
 * Allocates and frees its buffers. */
cs3d_status cs3d_copy_calibrate(cs3d_ctx* ctx, int64_t bytes, int32_t reps, double* avg_ms);

/* ---- arithmetic self-checks (no counterpart in the reference; for tests) --
 * cs3d_check_sqrt: the kernels' correctly rounded float square root (sqrt_cr)
 * against sqrtf over all 2^32 inputs on the context's device; *mismatches is
 * the number of differing results (NaN equals NaN).
 * cs3d_device_acos: the kernels' double acos of the angle force
 * (cellsim3d_amd/csrc/cs3d_acos.h, propagate.cu:36) of n host arguments, into
 * host y.
 * cs3d_check_div: the kernels' quotients from a reciprocal
 * (cellsim3d_amd/csrc/cs3d_div.h) against IEEE division, as (differing,
 * control) pairs.  out[0..1]: the reciprocal against 1.0f/b over every float b
 * with |b| in [2^-40, 2^40], and the number of such b compared; out[2..3]: the
 * quotient for every b in [1, 2) and the first a_significands significands of a
 * in [1, 2) (2^23 covers all 2^46 pairs); out[4..5]: scaled_pairs pseudo-random
 * pairs over the whole admitted range.  The control of a quotient check counts
 * the pairs where the uncorrected product RN(a * RN(1/b)) differs from a/b.
 * All three allocate and free device buffers. */
cs3d_status cs3d_check_sqrt(cs3d_ctx* ctx, uint64_t* mismatches);
cs3d_status cs3d_check_div(cs3d_ctx* ctx, int32_t a_significands, int64_t scaled_pairs, uint64_t out[6]);
cs3d_status cs3d_device_acos(cs3d_ctx* ctx, const double* x, double* y, int64_t n);

/* ---- trajectory (binary format of GPUbounce.cu:1007-1018, 2446-2477) ----- */
typedef struct cs3d_traj cs3d_traj;
cs3d_status cs3d_traj_open(const char* path, int32_t max_cells, int32_t variable_stiffness,
                           int32_t max_frames, cs3d_traj** out);
/* Appends one frame of n_cells cells (x/y/z with 192 stride; types nullable). */
cs3d_status cs3d_traj_append(cs3d_traj* t, int32_t step, int32_t frame, int32_t n_cells,
                             const float* x, const float* y, const float* z, const int32_t* types);
/* Rewrites the header's first int with final_cells (GPUbounce.cu:1508-1511) and closes. */
cs3d_status cs3d_traj_close(cs3d_traj* t, int32_t final_cells);

#ifdef __cplusplus
}
#endif
#endif /* CS3D_H */
