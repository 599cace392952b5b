// cs3d_dev.h - data layout shared by the HIP kernels and the host side of libcs3d.
//
// HBM layout (one context = one pool on one GPU):
//   node arrays  float[cap_cells * 192]  x, y, z, v, v_after_integrate, F_con, F_dis
//                (SoA, 192-slot stride per cell as the reference's d_X, GPUbounce.cu:365-373)
//   cell arrays  float[cap_cells]        CM x/y/z, pressure, stiffness, volume, area, growth scale
//                float[cap_cells * 6]    bbox used by the contact candidate test (rebuilt with the grid)
//                float[cap_cells * 72]   12 patch boxes of 16 consecutive nodes (rebuilt every step)
//                int32[cap_cells * NCMAX] per-cell list of candidate neighbour cells (ascending)
//   voxel pool   int32                   per system: counts[budget] then lists[budget * capacity]
//   slot table   int32[cap_cells]        system owning each cell slot (-1 = free)
// A system (tissue) owns a contiguous range of cell slots [base, base + cap) and a
// voxel region; its live cells are base .. base + n - 1 (local index 0 .. n - 1).
#pragma once
#include <stdint.h>
#include "../../include/cs3d.h"

namespace cs3d {

constexpr int kSlots = CS3D_SLOTS;   // 192
constexpr int kNodes = CS3D_NODES;   // 180
constexpr int kPatch = 16;           // nodes per contact patch
constexpr int kNPatch = 12;          // ceil(180 / 16)
constexpr int kPBox = kNPatch * 6;   // floats of patch boxes per cell
constexpr int kNCMax = 128;          // candidate neighbour cells per cell
constexpr int kNVMax = 64;           // voxels overlapped by one cell's box
constexpr int kMaxEnt = 2048;        // voxel-list entries gathered per cell
constexpr int kMaxSys = 4096;        // systems per context

// Per-system state kept on the device.
struct SysMeta {
    int32_t n;          // live cells
    int32_t cap;        // cell slots owned
    int32_t base;       // first cell slot
    int32_t active;     // 1 = stepped by cs3d_step
    int64_t step;       // completed steps
    int64_t horizon;    // stop after this many steps (-1 = none)
    uint32_t status;    // CS3D_SYS_* bits (sticky)
    int32_t rebuild;    // 1 = rebuild bbox + grid at the end of the current step
    int64_t vbase;      // offset of counts[] in the voxel pool
    int64_t lbase;      // offset of lists[] in the voxel pool
    int32_t vbudget;    // voxels available
    int32_t vcap;       // entries per voxel list
    float gmin[3];      // grid origin (minx, miny, minz)
    int32_t dims[3];    // Xdiv, Ydiv, Zdiv
    int32_t ext[6];     // extent accumulators (order-preserving int encoding)
    int32_t n_events;   // divisions in the last division phase
    int32_t pad0;
    cs3d_ranmar_state rng;
};

struct MeshDev {
    int32_t nn[3 * kSlots];
    int32_t sign[kSlots];
    int32_t faces[CS3D_FACES * 7];
    float R0[3 * kSlots];
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
    float *CMX, *CMY, *CMZ;
    float *P, *E, *VOL, *AREA, *GS;
    int8_t* FLAG;
    float* BBOX;
    float* PBOX;
    int32_t* NCL;
    int32_t* NCN;
    int32_t* EVL;
    int32_t* slot_sys;
    SysMeta* meta;
    const cs3d_params* prm;
    int32_t* vox;
    const MeshDev* mesh;
};

// What a launch works on: cell slots [slot_lo, slot_hi) and systems [sys_lo, sys_hi).
struct Range {
    int32_t slot_lo, slot_hi;
    int32_t sys_lo, sys_hi;
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
    M_INIT = 7,        // start-up: no pressure update, forces from state
};

}  // namespace cs3d
