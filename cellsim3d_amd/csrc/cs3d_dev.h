// cs3d_dev.h - data layout shared by the HIP kernels and the host side of libcs3d.
//
// HBM layout (one context = one pool on one GPU):
//   node arrays  float[cap_cells * 192]  x, y, z, v, v_after_integrate, F_con, F_dis, F_ran
//                (SoA, 192-slot stride per cell as the reference's d_X, GPUbounce.cu:365-373)
//   cell arrays  float[cap_cells]        CM x/y/z, pressure, stiffness, volume, area, growth scale
//                float[cap_cells * 6]    bbox used by the contact candidate test (rebuilt with the grid)
//                float[cap_cells * 6]    bbox of the positions at the end of the latest step
//                int32[cap_cells * 6]    voxel range of the contact-rule bbox (kept with it)
//   pair lists   int32[cap_cells * kPMax * 192]  contact partners of each node, (cell, node)
//                ascending, k-major per cell so lanes access consecutive words
//                int32[cap_cells * 192]  partner count per node
//                int32[cap_cells * kPMax * 192]  contact candidates (within attraction_range + skin
//                at the last search), same layout, + count per node
//                float[cap_cells * 192] x3  node positions at the last search
//   voxel pool   int32                   per system: counts[budget] then lists[budget * capacity]
//                                        (the reference's 2.9-wide cell grid)
//   bin pool     int32                   per system: counts[bbudget] then entries[bbudget * kBCap]
//                                        float4 (x, y, z, cell*192 + node) (node bins of edge >=
//                                        attraction_range for the pair search)
//   slot table   int32[cap_cells]        system owning each cell slot (-1 = free)
// A system (tissue) owns a contiguous range of cell slots [base, base + cap), a voxel
// region and a bin region; its live cells are base .. base + n - 1 (local 0 .. n - 1).
#pragma once
#include <stdint.h>
#include "../../include/cs3d.h"

namespace cs3d {

constexpr int kSlots = CS3D_SLOTS;   // 192
constexpr int kNodes = CS3D_NODES;   // 180
constexpr int kPMax = 40;            // contact candidates / partners stored per node
constexpr int kBCap = CS3D_BIN_CAP;  // nodes per search bin
constexpr int kMaxSys = 4096;        // systems per context
constexpr int kMaxPartial = 192;     // listed cells beyond which the next search is a full one (at most the
                                     // 192 lanes of a search workgroup, which test one listed cell each)
static_assert(kMaxPartial <= kSlots, "the partial search tests one listed cell per lane");

// Per-system state kept on the device.
struct SysMeta {
    int32_t n;          // live cells
    int32_t cap;        // cell slots owned
    int32_t base;       // first cell slot
    int32_t active;     // 1 = stepped by cs3d_step
    int64_t step;       // completed steps
    int64_t horizon;    // stop after this many steps (-1 = none)
    uint32_t status;    // CS3D_SYS_* bits (sticky)
    int32_t rebuild;    // fused step: 1 = the grid was rebuilt at the end of the last step and its
                        // per-cell part (bbox adoption, voxel counts) is still to be done by the
                        // next step; explicit grid kernels: 1 = rebuild now
    int64_t vbase;      // offset of counts[] in the voxel pool
    int64_t lbase;      // offset of lists[] in the voxel pool
    int32_t vbudget;    // voxels available
    int32_t vcap;       // entries per voxel list
    float gmin[3];      // voxel grid origin (minx, miny, minz)
    int32_t dims[3];    // Xdiv, Ydiv, Zdiv
    int32_t ext[6];     // extent accumulators (order-preserving int encoding)
    int32_t n_events;   // divisions in the last division phase
    int32_t bbudget;    // search bins available
    int64_t bbase;      // offset of bin counts in the bin pool
    int64_t ebase;      // offset of bin entries in the bin pool
    float bmin[3];      // bin grid origin
    float bh;           // bin edge
    int32_t bdims[3];   // bins per axis
    int32_t need_search;  // search contact candidates this step: 0 no, 1 every cell, 2 only the
                          // cells near the cells listed in RSL (partial search)
    int32_t search_next;  // set during a step when the next step's drift takes a node pair_skin/2 away
    int32_t n_reset;      // cells listed in RSL for the next partial search (moved past the skin or
                          // divided); counts past kMaxPartial mean a full search
    float sys_cm[3];      // system centre subtracted at the end of the latest step (correct_com)
    int32_t shift_pending;  // 1 = the step that just ended computed sys_cm and its nodes are still to be
                            // shifted (set by the division phase, consumed by the shift kernel)
    int64_t cell_steps;   // sum over the completed steps of the cells stepped (node-updates / 180)
    // the velocity kicks' double divisions by the particle mass as products with
    // double-double constants, verified for every float significand (kick_consts):
    // kick_a* = 1 / (2 m) for 0.5 * s / m, kick_b* = dt / (2 m) for 0.5 * dt * f / m
    double kick_ah, kick_al, kick_bh, kick_bl;
    int32_t kick_fast;    // 1: the constants above reproduce the divisions exactly
    cs3d_ranmar_state rng;
};

// Fills the kick constants of a system with particle mass m and time step dt and
// checks them exhaustively (on the host; results are cached per (m, dt)).
void kick_consts(float m, float dt, SysMeta& M);

struct MeshDev {
    int32_t nn[3 * kSlots];
    int32_t sign[kSlots];
    int32_t faces[CS3D_FACES * 7];
    float R0[3 * kSlots];
    float R0inv[3 * kSlots];  // RN(1/R0) where quot_ok(R0), NaN elsewhere
    float theta0[3 * kSlots];
    // angle gather table: for node p, six offsets (in floats) into the per-cell
    // angle-force array, in the summation order of CalculateAngleForce
    // (propagate.cu:124-178)
    int32_t gather[kNodes * 6];
};

struct Pool {
    float *X, *Y, *Z;
    float *VX, *VY, *VZ;     // velocity at end of step
    float *VAX, *VAY, *VAZ;  // velocity after Integrate (read by the first DisForce)
    float *FCX, *FCY, *FCZ;
    float *FDX, *FDY, *FDZ;
    float *FRX, *FRY, *FRZ;  // random force drawn in the latest step (d_fRanList)
    float *CMX, *CMY, *CMZ;
    float *P, *E, *VOL, *AREA, *GS;
    int8_t* FLAG;
    float* BBOX;
    float* BBOXN;
    int32_t* VRANGE;  // makeNNlist's voxel range of BBOX (sx, ex, sy, ey, sz, ez) per cell
    int32_t* PAIRS;
    int32_t* NPAIR;
    int32_t* CAND;
    int32_t* NCAND;
    float *XBX, *XBY, *XBZ;
    float* SBOX;     // float[cap_cells * 6]: box of each cell's search positions (written by full
                     // searches, and by k_integrate for RESET cells before a partial search)
    int8_t* RESET;   // int8[cap_cells]: 1 = the cell takes its current positions as search positions at
                     // the next partial search (a node left the skin, or the cell divided)
    int32_t* RSL;    // int32[cap_cells]: per system, the local indices of its RESET cells (n_reset)
    int32_t* LIVE;   // int32[cap_cells]: live cell slots of the stepped systems (k_live)
    int32_t* n_live; // int32[1]: entries of LIVE
    int32_t* EVL;
    int32_t* slot_sys;
    int32_t* nsys;   // int32[1]: system ids in use are below this (kernels over all systems read it, so a
                     // captured all-systems step stays valid when systems are added or removed)
    SysMeta* meta;
    const cs3d_params* prm;
    int32_t* vox;
    int32_t* bins;
    const MeshDev* mesh;
    int32_t wg_per_cu;  // host side: cap of the grid of the cell kernels, workgroups per CU
};

// What a launch works on: cell slots [slot_lo, slot_hi) and systems [sys_lo, sys_hi).
struct Range {
    int32_t slot_lo, slot_hi;
    int32_t sys_lo, sys_hi;
    int32_t one_sys;  // 1: the range is exactly the slot range of system sys_lo
    int32_t live;     // 1: the cell kernels walk Pool::LIVE (built by k_live) instead of the slots
};

// kernel modes
enum : int32_t {
    M_STEP = 0,        // fused step path (state flags decided on the device)
    M_CON = 1,         // CalculateConForce only -> FC
    M_BONDED = 2,      // spring + pressure + angle only -> FC
    M_DIS_V = 3,       // CalculateDisForce with V -> FD
    M_DIS_VA = 4,      // CalculateDisForce with VA -> FD
    M_COM = 5,         // CenterOfMass only
    M_VOLUMES = 6,     // volumes only (explicit step)
    M_INIT = 7,        // start-up: forces from the injected state
    M_PAIRS = 8,       // contact-partner lists only (for a following CalculateDisForce)
};

// k_integrate flags
enum : int32_t {
    I_KICK = 1,       // Integrate + ForwardTime
    I_PRESSURE = 2,   // PressureUpdate
    I_BINS = 4,       // insert the nodes into the search bins
    I_CLEAR = 8,      // clear the search-bin counters of the nodes
    I_IFSEARCH = 32,  // I_BINS only on systems that search this step
    I_GRIDCOUNT = 64, // systems whose grid was rebuilt at the end of the last step: adopt the new
                      // bboxes and count the voxel occupancy (overflow check)
    I_WRITEV = 128,   // with I_KICK: also store the new velocity in V (standalone Integrate)
};

}  // namespace cs3d
