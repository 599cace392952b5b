// cs3d_kernels.hip - CDNA4 (gfx950) kernels of the CellSim3D hot path.
//
// Cell kernels run one cell per 192-lane workgroup (3 x wave64; lanes 180..191
// are the padding slots) and grid-stride over cell slots, so a fixed launch
// covers a dynamic cell count.  Every kernel reproduces the floating-point
// operation order of the reference kernel it replaces (cited per function), so
// results match the reference arithmetic; the library is compiled with
// -ffp-contract=off and correctly rounded fp32 division and square root.
//
// Contact pairs.  The reference scans, per node, the cells listed in its
// voxel (2.9-wide grid), keeps those whose bbox lies within attraction_range of
// the node, and tests all 180 nodes of each kept cell (propagate.cu:375-442,
// 603-655).  Here each node keeps a sorted list of candidate partners: the nodes
// of other cells within attraction_range + pair_skin, found through periodic
// search bins of that edge (find_candidates).  The list is rebuilt only on steps where some
// node has moved pair_skin/2 since the last search, so it always contains every
// node now within attraction_range; after divisions only the lists near the divided
// cells are rebuilt (k_search's partial search).  Every step then applies the reference's own
// tests to the candidates in (cell, node) order -- the order of the reference's
// loops: the distance test, voxel-list membership and the bbox test
// (accept_forces).  Voxel-list membership is evaluated arithmetically: makeNNlist
// puts cell b into voxel v exactly when v lies in the voxel range of b's bbox
// widened by attraction_range (propagatebound.cu:224-259), so the same range test
// on the same bbox and grid answers "b in list(v)".  The accepted pairs are
// stored for the second friction pass.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <utility>
#include <vector>

#include "cs3d_acos.h"
#include "cs3d_div.h"
#include "cs3d_dev.h"

namespace cs3d {

// The system owning a cell slot.  A launch over one system's own slot range needs
// no table read: every slot of the range belongs to it.
__device__ __forceinline__ int sys_of(const Pool& p, const Range& r, int slot) {
    return r.one_sys ? r.sys_lo : p.slot_sys[slot];
}

template <class MP> __device__ __forceinline__ bool sys_live(MP m) {
    return m->active && (m->horizon < 0 || m->step < m->horizon) && !(m->status & CS3D_SYS_FATAL_MASK);
}

__device__ __forceinline__ void flag(SysMeta* m, uint32_t bits) {
    if (bits) atomicOr(&m->status, bits);
}

// Loads and stores at a uniform base pointer plus a 32-bit element index: the
// address is an SGPR pair plus one VGPR byte offset (the saddr form of the global
// memory instructions), so per-lane addressing costs one 32-bit register that all
// arrays of a cell share instead of a 64-bit pointer per array.  The byte offset
// must stay below 4 GiB: a cell's node offset (< 768 B), a list entry (< 31 KB),
// or a node key of a system (a context holds at most 2^30 / 192 cells, so every
// key * 4 < 2^32).
template <class T> __device__ __forceinline__ T ldu(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}
template <class T> __device__ __forceinline__ void stu(T* base, uint32_t i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T)) = v;
}

// Correctly rounded float square root: the hardware estimate and one correction
// step (the neighbours one ulp either side of the estimate, chosen by the signs of
// the exact fma residuals x - s' s).  Inputs below 2^-96, where the residuals
// would lose bits, are scaled by 2^32 and the result by 2^-16 (exact).  This is
// the library's sequence without its fix-ups for 0 and infinity, which the
// correction step already returns unchanged (as it does NaN; a negative input
// gives NaN).  Equal to sqrtf bit for bit over all 2^32 inputs (cs3d_check_sqrt).
__device__ __forceinline__ float sqrt_cr(float x) {
    const bool tiny = x < 0x1p-96f;
    const float xs = tiny ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = fmaf(-sm, s, xs), rp = fmaf(-sp, s, xs);
    float t = rm <= 0.f ? sm : s;
    t = rp > 0.f ? sp : t;
    return tiny ? t * 0x1p-16f : t;
}

// RN(1/b) for quot_ok(b): the hardware reciprocal estimate and one Newton step
// with an fma residual.  Equal to 1.0f / b for every float b with quot_ok(b)
// (cs3d_check_div).
__device__ __forceinline__ float rcp_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, y0, 1.0f);
    return fmaf(e, y0, y0);
}

// Marks the division path of a quotient as a branch: the empty volatile asm
// keeps the compiler from evaluating both paths and selecting (which it does for
// a short division otherwise).
__device__ __forceinline__ void slow_path() { asm volatile(""); }

// 1.0f / b: from rcp_rn where quot_ok(b), by division otherwise (the same bits)
__device__ __forceinline__ float inv_rn(float b) {
    float y = rcp_rn(b);
    if (!quot_ok(b)) {
        slow_path();
        y = 1.0f / b;
    }
    return y;
}

// One bond's spring term stiffness*(R-R0)/R0*d/R for d = dx, dy, dz
// (propagate.cu:311-313).  spring_rcp takes the quotients from the reciprocals of
// R and R0 (iR0 = RN(1/R0), NaN where R0 is outside quot_ok's range) and returns
// whether every operand was in range (quot_ok, quot_num_ok); where it was, the
// terms equal spring_div's, the reference's divisions, bit for bit.  The callers
// sum a node's three bonds with spring_rcp and, where some operand was out of
// range, sum them again with spring_div (one branch per node).
__device__ __forceinline__ bool spring_rcp(float stiffness, float R, float R0, float iR0, float dx, float dy, float dz,
                                           float& tx, float& ty, float& tz) {
    const float a = stiffness * (R - R0);
    const float c = quot_rn(a, R0, iR0);
    const float yR = rcp_rn(R);
    const float nx = c * dx, ny = c * dy, nz = c * dz;
    tx = quot_rn(nx, R, yR);
    ty = quot_rn(ny, R, yR);
    tz = quot_rn(nz, R, yR);
    bool ok = quot_ok(R);  // every compare evaluated, no short-circuit branches
    ok &= quot_num_ok(a);
    ok &= quot_num_ok(nx);
    ok &= quot_num_ok(ny);
    ok &= quot_num_ok(nz);
    return ok;
}
__device__ __forceinline__ void spring_div(float stiffness, float R, float R0, float dx, float dy, float dz, float& tx,
                                           float& ty, float& tz) {
    tx = stiffness * (R - R0) / R0 * dx / R;
    ty = stiffness * (R - R0) / R0 * dy / R;
    tz = stiffness * (R - R0) / R0 * dz / R;
}

// rGrowth of the population model, GPUbounce.cu:1077-1088
template <class PP> __device__ __forceinline__ float rgrowth(PP P, int n) {
    if (P->do_pop_model == 1) {
        float r = (float)((double)P->growth_rate * (1 - (n * 1.0 / (double)P->max_pop)));
        if (r < 0) r = 0;
        return r;
    }
    return P->growth_rate;
}

// voxel coordinate of one axis, propagate.cu:360-370
__device__ __forceinline__ int vox1(float x, float mn, float DL, int div) {
    int s = (int)((x - mn) / DL);
    if (s < 0) s = 0;
    if (s >= div) s = div - 1;
    return s;
}

// Read-only views of a system's parameters and metadata in the constant address
// space: a kernel reads them with scalar loads into scalar registers (the values
// are the same for every lane of a workgroup).  Only for data no kernel writes
// while the reader runs; the scalar cache is only read through.
typedef __attribute__((address_space(4))) const cs3d_params* CParams;
typedef __attribute__((address_space(4))) const SysMeta* CMeta;
__device__ __forceinline__ CParams cparams(const cs3d_params* q) { return (CParams)q; }
__device__ __forceinline__ CMeta cmeta(const SysMeta* q) { return (CMeta)q; }
template <class T> __device__ __forceinline__ T cld(const T* q) {
    return *(__attribute__((address_space(4))) const T*)q;
}

// The kernel's Pool argument (the first kernel parameter) read again where it is
// used: scalar loads from the kernel-argument segment behind an empty asm
// statement, so that the compiler neither hoists them nor keeps the pointers live
// (in scalar registers, or spilled to vector lanes) across the contact loop.
typedef __attribute__((address_space(4))) const Pool* CPool;
__device__ __forceinline__ CPool pool_late() {
    CPool q = (CPool)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(q));
    return q;
}

// The cell slot of loop iteration `it` of a cell kernel, made wave-uniform with
// readfirstlane: the per-cell base pointers derived from it stay in scalar
// registers instead of becoming per-lane 64-bit induction variables.
template <int kLive>
__device__ __forceinline__ int cell_slot(const Pool& p, const Range& r, int it) {
    return __builtin_amdgcn_readfirstlane(kLive ? p.LIVE[it] : r.slot_lo + it);
}

// The lane's node index as a value the compiler cannot see through (an empty asm
// statement per loop iteration): otherwise it combines "array + lane" into one
// loop-invariant 64-bit pointer per array and lane, which costs two VGPRs per
// array for the whole kernel.  With it, every access is the cell's uniform base
// plus this one 32-bit offset.
__device__ __forceinline__ uint32_t lane_index() {
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// A value read from LDS that the compiler may not replace by a register copy of
// what was stored there: reading a node's own position again where it is needed
// ends the live range of the register that held it.
__device__ __forceinline__ float lds_reread(const float* a, int i) {
    float v = a[i];
    asm volatile("" : "+v"(v));
    return v;
}

// XCD-aware first cell of a workgroup.  The dispatcher hands workgroup v to XCD
// v mod 8; this maps the 8 XCDs to 8 contiguous runs of workgroup indices, so the
// cells one XCD works on are neighbours in index and, for tissues placed in
// order, in space: their partner gathers hit the XCD's own L2.  A bijection on
// [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x, v = blockIdx.x;
    const int q = G / kXcds, rem = G % kXcds;
    const int k = v % kXcds, j = v / kXcds;
    return k * q + min(k, rem) + j;
}

// Search bins are periodic: bin coordinate floor((x - origin) / h) modulo the bin
// count of the axis.  Regions farther apart than the bin grid alias into the same
// bins, which only adds entries the distance test rejects, so the bin edge never
// has to grow with the system's extent (outlying cells cannot coarsen the bins of
// a dense tissue).
__device__ __forceinline__ int bin1(float x, float mn, float h, int div) {
    float q = floorf((x - mn) / h);
    q = q == q ? fminf(fmaxf(q, -1.0e9f), 1.0e9f) : 0.0f;  // NaN positions go to bin 0 (flagged elsewhere)
    int s = (int)q % div;
    return s < 0 ? s + div : s;
}
__device__ __forceinline__ int wrap1(int s, int div) { return s < 0 ? s + div : (s >= div ? s - div : s); }

template <class MP> __device__ __forceinline__ int bin_of(MP M, float x, float y, float z) {
    const int bx = bin1(x, M->bmin[0], M->bh, M->bdims[0]);
    const int by = bin1(y, M->bmin[1], M->bh, M->bdims[1]);
    const int bz = bin1(z, M->bmin[2], M->bh, M->bdims[2]);
    return (bz * M->bdims[1] + by) * M->bdims[0] + bx;
}

// order-preserving int encoding of floats for atomic min/max
__device__ __forceinline__ int f2o(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

// 0.5 * (double)s / (double)m of the velocity kicks (propagate.cu:698-705,
// 721-723) for a float s.  On systems whose constants passed the host's
// exhaustive check (kick_consts): s times the double-double 1 / (2 m), one
// multiply and one fma with the division's exact result; a zero keeps its sign.
__device__ __forceinline__ double kick_a(float s, float m, CMeta M) {
    const double x = (double)s;
    if (M->kick_fast) {
        const double q = fma(x, M->kick_ah, x * M->kick_al);
        return s == 0.0f ? x : q;
    }
    return 0.5 * x / (double)m;
}
// 0.5 * (double)dt * (double)f / (double)m of VelocityUpdateB (propagate.cu:732-734)
__device__ __forceinline__ double kick_b(float f, float dt, float m, CMeta M) {
    const double x = (double)f;
    if (M->kick_fast) {
        const double q = fma(x, M->kick_bh, x * M->kick_bl);
        return f == 0.0f ? x : q;
    }
    return 0.5 * (double)dt * x / (double)m;
}

// Integrate's velocity half-kick of one component (propagate.cu:698-705): the
// sum in float, the rest in double (0.5 is a double literal), rounded to float.
// fr is the random force drawn in the previous step (0 without add_rands).
__device__ __forceinline__ float half_kick(float v, float fc, float fd, float fr, float dt, float rdt, float m,
                                           CMeta M) {
    return (float)((double)v + kick_a(dt * fc + dt * fd + rdt * fr, m, M));
}

// ---- random force (CalculateRanForce, propagate.cu:667-676) --------------------
// The reference draws curand_normal from per-node cuRAND states (closed library,
// not reproduced).  Here: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11) with counter {cell * 192 + node,
// step (low, high), attempt} and key {noise_seed, 0x43533344}, turned into
// standard normals by the Marsaglia polar method.  Every operation is integer or
// correctly rounded float (the logarithm is a fixed polynomial), so the host and
// the device produce the same bits.
__device__ __forceinline__ void philox4x32_10(uint32_t* c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// natural logarithm of s in (0, 1): s = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh((m - 1) / (m + 1)) by its series to z^9, e ln 2 in two parts
__device__ __forceinline__ float ln01(float s) {
    const uint32_t b = __float_as_uint(s);
    int e = (int)((b >> 23) & 0xffu) - 126;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f000000u);
    if (m < 0.70710678f) {
        m = m + m;
        e -= 1;
    }
    const float z = (m - 1.0f) / (m + 1.0f);
    const float z2 = z * z;
    float q = 0.11111111f;
    q = q * z2 + 0.14285715f;
    q = q * z2 + 0.2f;
    q = q * z2 + 0.33333334f;
    q = q * z2 + 1.0f;
    const float fe = (float)e;
    return fe * 0.693145751953125f + (2.0f * z * q + fe * 1.4286068e-06f);
}

// three standard normals of node `node` of a tissue at step `step`
__device__ __forceinline__ void noise3(uint32_t seed, uint32_t node, int64_t step, float* z) {
    int got = 0;
    z[0] = z[1] = z[2] = 0.f;
    for (uint32_t attempt = 0; got < 3 && attempt < 32; ++attempt) {
        uint32_t c[4] = {node, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), attempt};
        philox4x32_10(c, seed, 0x43533344u);
#pragma unroll
        for (int pr = 0; pr < 2; ++pr) {
            const float v1 = 2.0f * (((float)(c[2 * pr] >> 9) + 0.5f) * 1.1920929e-07f) - 1.0f;
            const float v2 = 2.0f * (((float)(c[2 * pr + 1] >> 9) + 0.5f) * 1.1920929e-07f) - 1.0f;
            const float s = v1 * v1 + v2 * v2;
            if (got < 3 && s > 0.0f && s < 1.0f) {
                const float f = sqrt_cr(-2.0f * ln01(s) / s);
                z[got++] = v1 * f;
                if (got < 3) z[got++] = v2 * f;
            }
        }
    }
}

// makeNNlist's voxel range of a cell (propagatebound.cu:224-245): the voxels of
// its bbox widened by attraction_range, clamped to the grid.
struct VoxRange {
    int sx, ex, sy, ey, sz, ez;
};
template <class MP, class PP> __device__ __forceinline__ VoxRange vox_range(const float* b, MP M, PP P) {
    const float DL = P->grid_dl, at = P->attraction_range;
    const float mx = M->gmin[0], my = M->gmin[1], mz = M->gmin[2];
    VoxRange v;
    v.sx = (int)((b[0] - at - mx) / DL); if (v.sx < 0) v.sx = 0;
    v.ex = (int)((b[1] + at - mx) / DL); if (v.ex >= M->dims[0]) v.ex = M->dims[0] - 1;
    v.sy = (int)((b[2] - at - my) / DL); if (v.sy < 0) v.sy = 0;
    v.ey = (int)((b[3] + at - my) / DL); if (v.ey >= M->dims[1]) v.ey = M->dims[1] - 1;
    v.sz = (int)((b[4] - at - mz) / DL); if (v.sz < 0) v.sz = 0;
    v.ez = (int)((b[5] + at - mz) / DL); if (v.ez >= M->dims[2]) v.ez = M->dims[2] - 1;
    return v;
}

// voxel range of a cell's bbox into VRANGE (lanes 0..5 of the caller store)
__device__ __forceinline__ void store_vrange(int32_t* vrange, int64_t slot, const VoxRange& v, int lane) {
    if (lane < 6) {
        const int w[6] = {v.sx, v.ex, v.sy, v.ey, v.sz, v.ez};
        vrange[slot * 6 + lane] = w[lane];
    }
}

// Grid geometry from the global extent e = {minx, maxx, miny, maxy, minz, maxz}:
// host Xdiv/Ydiv/Zdiv (GPUbounce.cu:1331-1334) with the voxel origin at the
// extent's minimum, and the search bins (edge a little above attraction_range +
// pair_skin so that every pair within that range lies in adjacent bins, two bins
// of margin, coarser when the extent outgrows the bin budget).  Returns the
// number of voxel counters to clear.  Called by one thread.
__device__ int64_t set_grid_geometry(SysMeta* M, const cs3d_params* P, const float* e) {
    if (!(e[1] >= e[0] && e[3] >= e[2] && e[5] >= e[4])) {
        // no cells (the extent kept its sentinels): a one-voxel grid, nothing to flag
        for (int a = 0; a < 3; ++a) {
            M->gmin[a] = 0.0f; M->dims[a] = 1;
            M->bmin[a] = 0.0f; M->bdims[a] = 3;
        }
        return 1;
    }
    const float DL = P->grid_dl;
    const int xd = (int)((e[1] - e[0]) / DL + 1.0f);
    const int yd = (int)((e[3] - e[2]) / DL + 1.0f);
    const int zd = (int)((e[5] - e[4]) / DL + 1.0f);
    const int64_t nv = (int64_t)xd * yd * zd;
    const bool ok = xd > 0 && yd > 0 && zd > 0 && nv <= M->vbudget;
    M->gmin[0] = e[0]; M->gmin[1] = e[2]; M->gmin[2] = e[4];
    M->dims[0] = ok ? xd : 1; M->dims[1] = ok ? yd : 1; M->dims[2] = ok ? zd : 1;
    if (!ok) atomicOr(&M->status, (uint32_t)CS3D_SYS_GRID_TOO_LARGE);
    // search bins of edge attraction_range + pair_skin covering the extent, or
    // fewer along the longest axes when the budget is short (periodic bins)
    float h = (P->attraction_range + P->pair_skin) * 1.001f;
    if (!(h > 0.f)) h = 1.0f;
    const float sz[3] = {e[1] - e[0], e[3] - e[2], e[5] - e[4]};
    int bd[3];
    for (int a = 0; a < 3; ++a) {
        const float q = sz[a] / h;
        bd[a] = (q >= 0.f && q < 1.0e6f) ? (int)q + 5 : 1000000;
    }
    for (int it = 0; it < 200; ++it) {
        if ((int64_t)bd[0] * bd[1] * bd[2] <= M->bbudget) break;
        const int a = bd[0] >= bd[1] && bd[0] >= bd[2] ? 0 : (bd[1] >= bd[2] ? 1 : 2);
        bd[a] = max(3, (int)(bd[a] * 0.9f));
    }
    for (int a = 0; a < 3; ++a) {
        M->bmin[a] = e[2 * a] - 2.0f * h;
        M->bdims[a] = bd[a];
    }
    M->bh = h;
    return ok ? nv : (int64_t)M->vbudget;
}

// {minx, maxx, miny, maxy, minz, maxz} of the node positions of one cell
// (workgroup of 192 lanes, nodes on lanes < 180) into out[6] and, through the
// shared part[3][6], to every lane's return in box[6].  Called by all lanes.
__device__ __forceinline__ void box_of_cell(float x, float y, float z, bool node, float (*part)[6], float* out,
                                            float* box = nullptr) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    float v[6] = {node ? x : INFINITY, node ? x : -INFINITY, node ? y : INFINITY,
                  node ? y : -INFINITY, node ? z : INFINITY, node ? z : -INFINITY};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 6; a += 2) {
            v[a] = fminf(v[a], __shfl_xor(v[a], o));
            v[a + 1] = fmaxf(v[a + 1], __shfl_xor(v[a + 1], o));
        }
    }
    if (lane == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) part[wv][a] = v[a];
    __syncthreads();
    float b[6];
#pragma unroll
    for (int a = 0; a < 6; a += 2) {
        b[a] = fminf(fminf(part[0][a], part[1][a]), part[2][a]);
        b[a + 1] = fmaxf(fmaxf(part[0][a + 1], part[1][a + 1]), part[2][a + 1]);
    }
    __syncthreads();  // part is reused
    if (out && tid < 6) {
        float val = b[0];
#pragma unroll
        for (int a = 1; a < 6; ++a)
            if (tid == a) val = b[a];
        out[tid] = val;
    }
    if (box)
#pragma unroll
        for (int a = 0; a < 6; ++a) box[a] = b[a];
}

// ------------------------------------------------------------------------------
// Integrate + ForwardTime (propagate.cu:681-712, 738-763), PressureUpdate
// (PressureKernels.cu:4-41) and insertion of the new positions into the
// search bins.
// ------------------------------------------------------------------------------
// 6 waves per SIMD (at most 80 VGPRs): 8 workgroups per CU, one round for 2048 cells
// kLive: walk Pool::LIVE (a multi-system step) instead of the range's slots
template <int kLive>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_integrate(Pool p, Range r,
                                                                                              int32_t flags,
                                                                                              int64_t step_override) {
    __shared__ float ibox[3][6];
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    // a multi-system step walks the compact list of live slots (k_live)
    const int nit = kLive ? *p.n_live : r.slot_hi - r.slot_lo;
    for (int it = xcd_block(); it < nit; it += gridDim.x) {
        const int slot = kLive ? p.LIVE[it] : r.slot_lo + it;
        // the cell's node data are loaded before the system checks (every slot of the
        // range is allocated), so their latency overlaps the metadata reads
        // (plain 64-bit node indices here: measured faster for this short kernel than
        // the uniform-base form of the other cell kernels)
        const int64_t gi = (int64_t)slot * kSlots + tid;
        float x = p.X[gi], y = p.Y[gi], z = p.Z[gi];
        const bool kick = flags & I_KICK;
        float vx = 0.f, vy = 0.f, vz = 0.f, fcx = 0.f, fcy = 0.f, fcz = 0.f, fdx = 0.f, fdy = 0.f, fdz = 0.f;
        if (kick) {
            vx = p.VX[gi]; vy = p.VY[gi]; vz = p.VZ[gi];
            fcx = p.FCX[gi]; fcy = p.FCY[gi]; fcz = p.FCZ[gi];
            fdx = p.FDX[gi]; fdy = p.FDY[gi]; fdz = p.FDZ[gi];
        }
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        // metadata and parameters through scalar loads: this kernel writes no field it
        // reads (only status bits, by atomics)
        SysMeta* M = p.meta + s;
        const CMeta MC = cmeta(M);
        const CParams P = cparams(p.prm + s);
        if (!sys_live(MC)) continue;
        const int c = slot - MC->base;
        if (c >= MC->n) continue;
        const CPool q = pool_late();  // pointers read again where used (see pool_late)
        if ((flags & I_GRIDCOUNT) && MC->rebuild && !(MC->status & CS3D_SYS_GRID_TOO_LARGE)) {
            // the grid was rebuilt at the end of the last step: this cell's new bbox
            // becomes the one the contact rule uses, and its voxels are counted
            // (makeNNlist's atomicAdd, propagatebound.cu:246-250) to catch overflow
            const float* bn = q->BBOXN + (int64_t)slot * 6;
            float b[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) b[a] = bn[a];
            if (tid < 6) q->BBOX[(int64_t)slot * 6 + tid] = b[tid];
            const VoxRange v = vox_range(b, MC, P);
            store_vrange(q->VRANGE, slot, v, tid);
            const int nx = v.ex - v.sx + 1, ny = v.ey - v.sy + 1, nz = v.ez - v.sz + 1;
            const int total = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;
            int32_t* counts = q->vox + MC->vbase;
            const int xd = MC->dims[0], yd = MC->dims[1], cap = MC->vcap;
            bool over = false;
            for (int t = tid; t < total; t += blockDim.x) {
                const int j3 = v.sz + t % nz, j2 = v.sy + (t / nz) % ny, j1 = v.sx + t / (nz * ny);
                if (atomicAdd(&counts[j3 * xd * yd + j2 * xd + j1], 1) >= cap) over = true;
            }
            if (__any(over) && (tid & 63) == __ffsll((unsigned long long)__ballot(over)) - 1)
                atomicOr(&M->status, (uint32_t)CS3D_SYS_VOXEL_OVERFLOW);
        }
        if (node && kick) {
            const float dt = P->dt, m = P->mass;
            const float rdt = sqrt_cr(dt);
            const bool rnd = P->add_rands != 0;  // the previous step's random force (d_fRanList)
            vx = half_kick(vx, fcx, fdx, rnd ? q->FRX[gi] : 0.f, dt, rdt, m, MC);
            vy = half_kick(vy, fcy, fdy, rnd ? q->FRY[gi] : 0.f, dt, rdt, m, MC);
            vz = half_kick(vz, fcz, fdz, rnd ? q->FRZ[gi] : 0.f, dt, rdt, m, MC);
            x = x + vx * dt;
            y = y + vy * dt;
            z = z + vz * dt;
            q->VAX[gi] = vx; q->VAY[gi] = vy; q->VAZ[gi] = vz;
            if (flags & I_WRITEV) { q->VX[gi] = vx; q->VY[gi] = vy; q->VZ[gi] = vz; }
            q->X[gi] = x; q->Y[gi] = y; q->Z[gi] = z;
        }
        if (node && (flags & I_CLEAR)) q->bins[MC->bbase + bin_of(MC, x, y, z)] = 0;
        if ((flags & I_BINS) && (!(flags & I_IFSEARCH) || MC->need_search)) {
            // insertion into the search bins, one atomic per node (the order of a bin's
            // entries does not matter: every candidate list is sorted).  Grouping the
            // lanes of a wave that share a bin behind one atomic was slower: the loop
            // over a wave's distinct bins cost more than the atomics it saved (the colony
            // at 20,000 cells: 0.215 against 0.188 ms for this kernel).  A partial search
            // (after divisions) works on search positions: the nodes of cells that did
            // not divide are binned at their last search position
            float qx = x, qy = y, qz = z;
            if ((flags & I_IFSEARCH) && MC->need_search == 2 && node && !q->RESET[slot]) {
                qx = q->XBX[gi]; qy = q->XBY[gi]; qz = q->XBZ[gi];
            }
            const int b = node ? bin_of(MC, qx, qy, qz) : -1;
            int32_t* counts = q->bins + MC->bbase;
            if (node) {
                const int k = atomicAdd(&counts[b], 1);
                if (k < kBCap) {
                    float4* ent = reinterpret_cast<float4*>(q->bins + MC->ebase);
                    ent[(int64_t)b * kBCap + k] = make_float4(qx, qy, qz, __int_as_float(c * kSlots + tid));
                } else {
                    flag(M, CS3D_SYS_BIN_OVERFLOW);
                }
            }
        }
        if (tid == 0 && (flags & I_PRESSURE)) {
            const int64_t stepno = step_override >= 0 ? step_override : MC->step + 1;
            const float inc = rgrowth(P, MC->n) * q->GS[slot];
            const float mx = P->max_pressure;
            float pr = q->P[slot];
            if (P->use_different_stiffnesses) {
                if (stepno < P->phase_count) {
                    if (pr < mx) pr += inc;
                } else if (q->E[slot] == P->stiffness1) {
                    if (pr < mx) pr += inc;
                } else if (pr != 0.f) {
                    pr = 0.f;
                }
            } else if (pr < mx) {
                pr += inc;
            }
            q->P[slot] = pr;
        }
        if (kick && (flags & I_IFSEARCH) && MC->need_search == 2 && q->RESET[slot]) {
            // a cell that divided or left the skin in the last step publishes the box of
            // its new positions: the partial search re-searches the cells within reach
            box_of_cell(x, y, z, node, ibox, q->SBOX + (int64_t)slot * 6);
        }
    }
}

// ------------------------------------------------------------------------------
// Contact candidates of one node (called per lane by a cell workgroup): every
// node j of another cell within attraction_range + pair_skin, found by scanning
// the 27 search bins around the node.  Keys b*192 + j (local cell ids) are left
// ascending in lst[0 .. n)[tid]; returns n.
// ------------------------------------------------------------------------------
template <class MP, class PP>
__device__ int find_candidates(const Pool& p, MP M, PP P, int c, float X, float Y, float Z, int32_t (*lst)[kSlots],
                               int tid, uint32_t& st) {
    int nc = 0;
    const float rc = P->attraction_range + P->pair_skin;
    const float rc2 = rc * rc;
    const int bx = bin1(X, M->bmin[0], M->bh, M->bdims[0]);
    const int by = bin1(Y, M->bmin[1], M->bh, M->bdims[1]);
    const int bz = bin1(Z, M->bmin[2], M->bh, M->bdims[2]);
    const int bdx = M->bdims[0], bdy = M->bdims[1], bdz = M->bdims[2];
    const int32_t* counts = p.bins + M->bbase;
    const float4* ents = reinterpret_cast<const float4*>(p.bins + M->ebase);
    auto consider = [&](const float4 en) {
        const int key = __float_as_int(en.w);
        if ((unsigned)(key - c * kSlots) < (unsigned)kSlots) return;  // a node of this cell
        const float dX = X - en.x, dY = Y - en.y, dZ = Z - en.z;
        if (dX * dX + dY * dY + dZ * dZ > rc2) return;
        if (nc < kPMax) lst[nc][tid] = key;
        ++nc;
    };
    // 9 rows of 3 x-adjacent bins; the first 4 entries of each bin of a row are
    // loaded together (independent loads), longer bins finish in a tail loop; the
    // next row's counts load while a row is scanned
    constexpr int kHead = 4;
    int bin_n[3], cnt_n[3];
    {
        const int y3 = wrap1(by - 1, bdy), z3 = wrap1(bz - 1, bdz);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            bin_n[q] = (z3 * bdy + y3) * bdx + wrap1(bx + q - 1, bdx);
            cnt_n[q] = counts[bin_n[q]];
        }
    }
    for (int row = 0; row < 9; ++row) {
        int bi[3], cn[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            bi[q] = bin_n[q];
            cn[q] = min(cnt_n[q], kBCap);
        }
        if (row < 8) {
            const int y3 = wrap1(by + (row + 1) % 3 - 1, bdy), z3 = wrap1(bz + (row + 1) / 3 - 1, bdz);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                bin_n[q] = (z3 * bdy + y3) * bdx + wrap1(bx + q - 1, bdx);
                cnt_n[q] = counts[bin_n[q]];
            }
        }
        float4 head[3][kHead];  // loads of empty slots are masked off
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int e = 0; e < kHead; ++e)
                head[q][e] = e < cn[q] ? ents[(int64_t)bi[q] * kBCap + e] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int e = 0; e < kHead; ++e)
                if (e < cn[q]) consider(head[q][e]);
            for (int e = kHead; e < cn[q]; ++e) consider(ents[(int64_t)bi[q] * kBCap + e]);
        }
    }
    if (nc > kPMax) {
        st |= CS3D_SYS_CAND_OVERFLOW;
        nc = kPMax;
    }
    for (int i = 1; i < nc; ++i) {  // insertion sort of a short list
        const int key = lst[i][tid];
        int j = i - 1;
        while (j >= 0 && lst[j][tid] > key) {
            lst[j + 1][tid] = lst[j][tid];
            --j;
        }
        lst[j + 1][tid] = key;
    }
    return nc;
}

// The reference's candidate rule for node (X, Y, Z) in voxel (vx, vy, vz) and
// cell b: b is in the node's voxel list -- makeNNlist puts b into exactly the
// voxels of its bbox widened by attraction_range (propagatebound.cu:224-259),
// whose range is kept per cell in VRANGE -- and the node lies within
// attraction_range of b's bbox (propagate.cu:382-392).
struct RuleCtx {
    float ar2, DL, gx, gy, gz;
    int xd, yd, zd;
};

// sVR, sBB: voxel ranges and boxes from the system's first cell (uniform); b: the
// partner cell's local index
__device__ __forceinline__ bool rule_ok(const int32_t* __restrict__ sVR, const float* __restrict__ sBB, uint32_t b,
                                        int vx, int vy, int vz, float X, float Y, float Z, float ar2) {
    // all twelve words are loaded before any test (8-byte aligned pairs)
    const int2* vr = reinterpret_cast<const int2*>(sVR);
    const float2* bb = reinterpret_cast<const float2*>(sBB);
    const int2 r01 = ldu(vr, 3 * b), r23 = ldu(vr, 3 * b + 1), r45 = ldu(vr, 3 * b + 2);
    const float2 c01 = ldu(bb, 3 * b), c23 = ldu(bb, 3 * b + 1), c45 = ldu(bb, 3 * b + 2);
    const bool inlist = (r01.x <= vx) & (vx <= r01.y) & (r23.x <= vy) & (vy <= r23.y) & (r45.x <= vz) & (vz <= r45.y);
    const float b0 = c01.x, b1 = c01.y, b2 = c23.x, b3 = c23.y, b4 = c45.x, b5 = c45.y;
    float qx = (float)(X - b1 > 0.0f) * (X - b1);
    qx += (float)(b0 - X > 0.0f) * (b0 - X);
    float qy = (float)(Y - b3 > 0.0f) * (Y - b3);
    qy += (float)(b2 - Y > 0.0f) * (b2 - Y);
    float qz = (float)(Z - b5 > 0.0f) * (Z - b5);
    qz += (float)(b4 - Z > 0.0f) * (b4 - Z);
    return inlist && !(qx * qx + qy * qy + qz * qz > ar2);
}

template <class MP, class PP>
__device__ __forceinline__ RuleCtx rule_ctx(MP M, PP P) {
    RuleCtx R;
    const float ar = P->attraction_range;
    R.ar2 = ar * ar;
    R.DL = P->grid_dl;
    R.gx = M->gmin[0]; R.gy = M->gmin[1]; R.gz = M->gmin[2];
    R.xd = M->dims[0]; R.yd = M->dims[1]; R.zd = M->dims[2];
    return R;
}

// The angle force's quotients idk/ra2 and idk/rb2 from the reciprocals ya2, yb2
// of ra2, rb2 (cs3d_div.h) where the operands are in range (quot_ok, quot_num_ok),
// by division otherwise: the same bits either way.  (-1/sqrt(inner) is -inv_rn.)
__device__ __forceinline__ void angle_quotients(float idk, float ra2, float ya2, float rb2, float yb2, float& ca,
                                                float& cb) {
    ca = quot_rn(idk, ra2, ya2);
    cb = quot_rn(idk, rb2, yb2);
    bool ok = quot_num_ok(idk);
    ok &= quot_ok(ra2);
    ok &= quot_ok(rb2);
    if (!ok) {
        slow_path();
        ca = idk / ra2;
        cb = idk / rb2;
    }
}

// Shared GetAngleForce pair (propagate.cu:26-60): Fa = F(a-v, b-v), Fb = F(b-v, a-v).
// dot, |.|^2, c1 and theta are symmetric in the two arguments, so both come from
// one evaluation with bit-identical values.
__device__ __forceinline__ uint32_t angle_pair(float ax, float ay, float az, float bx, float by, float bz,
                                               float th0, float k, float* out) {
    const float idk = ax * bx + ay * by + az * bz;
    const float ra2 = ax * ax + ay * ay + az * az;
    const float rb2 = bx * bx + by * by + bz * bz;
    const float inner = (float)((double)(ra2 * rb2 - idk * idk) + 1e-3);
    const float c1 = -inv_rn(sqrt_cr(inner));
    const float den = sqrt_cr(ra2) * sqrt_cr(rb2);
    const float theta = (float)cs3d_acos((double)idk / ((double)den + 1e-3));
    const float s = ((-k) * c1) * (theta - th0);
    float ca, cb;
    angle_quotients(idk, ra2, rcp_rn(ra2), rb2, rcp_rn(rb2), ca, cb);
    out[0] = s * (bx - ca * ax); out[1] = s * (by - ca * ay); out[2] = s * (bz - ca * az);
    out[3] = s * (ax - cb * bx); out[4] = s * (ay - cb * by); out[5] = s * (az - cb * bz);
    bool ok = isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2]) && isfinite(out[3]) && isfinite(out[4]) &&
              isfinite(out[5]);
    return ok ? 0u : (uint32_t)CS3D_SYS_ANGLE_NONFINITE;
}

// GetAngleForce pair (as angle_pair) with the squared lengths ra2, rb2 of the two
// bond vectors and their square roots la, lb computed by the caller: the same
// operations on the same vectors as in angle_pair (the springs of the node use
// them too), so every bit is the same.  Its three quotients stay divisions here:
// from reciprocals, with the division path kept for operands out of range, they
// would keep more values live across the acos than the six-wave register budget
// of k_con_dis holds.
__device__ __forceinline__ uint32_t angle_pair_l(float ax, float ay, float az, float ra2, float la, float bx, float by,
                                                 float bz, float rb2, float lb, float th0, float k, float* out) {
    const float idk = ax * bx + ay * by + az * bz;
    const float inner = (float)((double)(ra2 * rb2 - idk * idk) + 1e-3);
    const float c1 = -1.0f / sqrt_cr(inner);
    const float den = la * lb;
    const float theta = (float)cs3d_acos((double)idk / ((double)den + 1e-3));
    const float s = ((-k) * c1) * (theta - th0);
    const float ca = idk / ra2, cb = idk / rb2;
    out[0] = s * (bx - ca * ax); out[1] = s * (by - ca * ay); out[2] = s * (bz - ca * az);
    out[3] = s * (ax - cb * bx); out[4] = s * (ay - cb * by); out[5] = s * (az - cb * bz);
    bool ok = isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2]) && isfinite(out[3]) && isfinite(out[4]) &&
              isfinite(out[5]);
    return ok ? 0u : (uint32_t)CS3D_SYS_ANGLE_NONFINITE;
}

// friction normal of node tid (propagate.cu:558-572), positions in LDS
__device__ __forceinline__ void node_normal(const float* sx, const float* sy, const float* sz, const MeshDev* mesh,
                                            int tid, int N1, int N2, int N3, float& nx, float& ny, float& nz) {
    const float Ax = sx[N2] - sx[N1], Ay = sy[N2] - sy[N1], Az = sz[N2] - sz[N1];
    const float Bx = sx[N3] - sx[N1], By = sy[N3] - sy[N1], Bz = sz[N3] - sz[N1];
    const float cx = Ay * Bz - Az * By, cy = Az * Bx - Ax * Bz, cz = Ax * By - Ay * Bx;
    const float inv = inv_rn(sqrt_cr(cx * cx + cy * cy + cz * cz));
    const float sg = (float)ldu(mesh->sign, lane_index());
    nx = sg * (inv * cx);
    ny = sg * (inv * cy);
    nz = sg * (inv * cz);
}

struct PairConsts {
    float ar, ar2, rr, ka, kr, gext;
};

// Partner keys are local cell * 192 + node: the index of the node in a system's
// arrays, read through the system's first slot (a uniform pointer).

constexpr int kChunk = 4;  // partners whose loads are issued together

// Friction (propagate.cu:633-653) over the node's stored partners, in stored
// (reference) order.  The keys, then the partner velocities, of kChunk partners
// are loaded before any of them is used.
// lst: the cell's list (uniform), entries at i * 192 + tid; sVX/sVY/sVZ: the
// velocity arrays from the system's first slot (uniform).
__device__ __forceinline__ void pair_friction(const int32_t* lst, int tid, int np, float vix, float viy, float viz,
                                              float nx, float ny, float nz, float gext,
                                              const float* __restrict__ sVX, const float* __restrict__ sVY,
                                              const float* __restrict__ sVZ, float& fdx, float& fdy, float& fdz) {
    for (int i0 = 0; i0 < np; i0 += kChunk) {
        uint32_t g[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) g[j] = i0 + j < np ? (uint32_t)ldu(lst, (uint32_t)((i0 + j) * kSlots + tid)) : 0u;
        float wx[kChunk], wy[kChunk], wz[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const bool in = i0 + j < np;
            wx[j] = in ? ldu(sVX, g[j]) : 0.f;
            wy[j] = in ? ldu(sVY, g[j]) : 0.f;
            wz[j] = in ? ldu(sVZ, g[j]) : 0.f;
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            if (i0 + j < np) {
                const float vx = vix - wx[j], vy = viy - wy[j], vz = viz - wz[j];
                const float d = vx * nx + vy * ny + vz * nz;
                const float tx = vx - d * nx, ty = vy - d * ny, tz = vz - d * nz;
                fdx = fdx - gext * tx;
                fdy = fdy - gext * ty;
                fdz = fdz - gext * tz;
            }
        }
    }
}

// Reference tests over the node's candidates in stored (= reference) order:
// distance (ConForce skips R^2 >= ar^2, DisForce R^2 > ar^2, propagate.cu:421,641)
// and the candidate rule per partner cell.  Accepted keys go to rec (for the
// second friction pass); contact (CON) and friction (DIS) are summed as in
// propagate.cu:411-440 and 633-653.  The loop is software-pipelined: while one
// candidate is tested, the partner data of the next and the key of the one after
// are already loading, so the key -> position dependence of each candidate does
// not serialise the loop.
// Returns the number accepted.
// cand, rec: the cell's lists (uniform), entries at i * 192 + tid; sX .. sVZ,
// sVR, sBB: node arrays, voxel ranges and boxes from the system's first slot
// (uniform).
template <bool CON, bool DIS>
__device__ __forceinline__ int accept_forces(const int32_t* cand, int tid, int nc, int32_t* rec, float X, float Y,
                                             float Z, const RuleCtx& R, const int32_t* __restrict__ sVR,
                                             const float* __restrict__ sBB, float vix,
                                             float viy, float viz, float nx, float ny, float nz, const PairConsts& k,
                                             const float* __restrict__ sX, const float* __restrict__ sY,
                                             const float* __restrict__ sZ, const float* __restrict__ sVX,
                                             const float* __restrict__ sVY, const float* __restrict__ sVZ,
                                             float& cfx, float& cfy, float& cfz, float& fdx, float& fdy,
                                             float& fdz) {
    int np = 0;
    int lastb = -1;
    bool lastok = false;
    int vx = -1, vy = -1, vz = -1;  // the node's voxel (propagate.cu:360-372), found at first use
    int key_n = nc > 0 ? ldu(cand, (uint32_t)tid) : 0;
    int key_nn = nc > 1 ? ldu(cand, (uint32_t)(kSlots + tid)) : 0;
    float pxn = 0.f, pyn = 0.f, pzn = 0.f, wxn = 0.f, wyn = 0.f, wzn = 0.f;
    if (nc > 0) {
        const uint32_t g = (uint32_t)key_n;
        pxn = ldu(sX, g); pyn = ldu(sY, g); pzn = ldu(sZ, g);
        if (DIS) { wxn = ldu(sVX, g); wyn = ldu(sVY, g); wzn = ldu(sVZ, g); }
    }
    for (int i = 0; i < nc; ++i) {
        const int key = key_n;
        const float px = pxn, py = pyn, pz = pzn;
        const float wx = wxn, wy = wyn, wz = wzn;
        key_n = key_nn;
        if (i + 1 < nc) {
            const uint32_t g = (uint32_t)key_n;
            pxn = ldu(sX, g); pyn = ldu(sY, g); pzn = ldu(sZ, g);
            if (DIS) { wxn = ldu(sVX, g); wyn = ldu(sVY, g); wzn = ldu(sVZ, g); }
        }
        if (i + 2 < nc) key_nn = ldu(cand, (uint32_t)((i + 2) * kSlots + tid));
        const int b = key / kSlots;
        const float dX = X - px, dY = Y - py, dZ = Z - pz;
        const float R2 = dX * dX + dY * dY + dZ * dZ;
        if (R2 > k.ar2) continue;
        if (b != lastb) {
            if (vx < 0) {
                vx = vox1(X, R.gx, R.DL, R.xd);
                vy = vox1(Y, R.gy, R.DL, R.yd);
                vz = vox1(Z, R.gz, R.DL, R.zd);
            }
            lastb = b;
            lastok = rule_ok(sVR, sBB, (uint32_t)b, vx, vy, vz, X, Y, Z, R.ar2);
        }
        if (!lastok) continue;
        stu(rec, (uint32_t)(np * kSlots + tid), key);
        ++np;
        if (CON && R2 < k.ar2) {
            const float Rl = sqrt_cr(R2);
            if (Rl < k.ar) {
                cfx += k.ka * (k.ar - Rl) / Rl * dX;
                cfy += k.ka * (k.ar - Rl) / Rl * dY;
                cfz += k.ka * (k.ar - Rl) / Rl * dZ;
            }
            if (Rl <= k.rr) {
                cfx += k.kr * (k.rr - Rl) / Rl * dX;
                cfy += k.kr * (k.rr - Rl) / Rl * dY;
                cfz += k.kr * (k.rr - Rl) / Rl * dZ;
            }
        }
        if (DIS) {
            const float ux = vix - wx, uy = viy - wy, uz = viz - wz;
            const float d = ux * nx + uy * ny + uz * nz;
            const float tx = ux - d * nx, ty = uy - d * ny, tz = uz - d * nz;
            fdx = fdx - k.gext * tx;
            fdy = fdy - k.gext * ty;
            fdz = fdz - k.gext * tz;
        }
    }
    return np;
}

// ------------------------------------------------------------------------------
// Contact-candidate search (one cell per workgroup): on systems whose nodes
// moved pair_skin/2 since the last search (or always, outside the fused step),
// rebuild each node's sorted candidate list and remember the positions.
// ------------------------------------------------------------------------------
// kLive: walk Pool::LIVE (a multi-system step) instead of the range's slots
template <int kLive>
__global__ __launch_bounds__(192) void k_search(Pool p, Range r, int32_t always) {
    __shared__ int32_t lst[kPMax][kSlots];
    __shared__ float part[3][6];
    __shared__ int32_t near;
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    if (r.one_sys && !always) {  // one system: its flag decides for every workgroup
        const CMeta M0 = cmeta(p.meta + r.sys_lo);
        if (!M0->need_search || !sys_live(M0)) return;
    }
    // a multi-system step walks the compact list of live slots (k_live)
    const int nit = kLive ? *p.n_live : r.slot_hi - r.slot_lo;
    for (int it = xcd_block(); it < nit; it += gridDim.x) {
        const int slot = cell_slot<kLive>(p, r, it);
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        // metadata and parameters through scalar loads: this kernel writes none of
        // the fields it reads (only status bits, by atomics)
        SysMeta* M = p.meta + s;
        const CMeta MC = cmeta(M);
        const CParams P = cparams(p.prm + s);
        if (!sys_live(MC) || !(always || MC->need_search)) continue;
        const int c = slot - MC->base;
        if (c >= MC->n) continue;
        const size_t cb = (size_t)slot * kSlots;  // the cell's first node (uniform)
        const uint32_t ln = lane_index();
        float X = ldu(p.X + cb, ln), Y = ldu(p.Y + cb, ln), Z = ldu(p.Z + cb, ln);
        bool reset = true;
        if (!always && MC->need_search == 2) {
            // Partial search, in search positions.  The cells listed in RSL (RESET: a
            // node of the cell would leave the skin, or the cell divided) take their
            // current positions as search positions; every other node is within
            // pair_skin/2 of its search position.  Invariant: every list holds every
            // node of another cell within attraction_range + pair_skin in search
            // positions; with the pair_skin/2 bound it holds every node that is, or
            // comes, within attraction_range before the next search.  The nodes were
            // binned at their search positions (k_integrate), and SBOX holds the box of
            // every cell's search positions (the listed cells' new boxes from
            // k_integrate).  A cell whose box is farther than that radius from every
            // listed cell's box keeps its list: the invariant holds for it with the new
            // search positions.  The listed cells and the cells near them search again
            // around their search positions, so their lists are as long as after a full
            // search.
            const bool own = p.RESET[slot] != 0;
            if (!own) {
                const float* ob = p.SBOX + (int64_t)slot * 6;
                const float b[6] = {ob[0], ob[1], ob[2], ob[3], ob[4], ob[5]};
                const int nr = min(MC->n_reset, kMaxPartial);
                const float mg = P->attraction_range + P->pair_skin;
                bool hit = false;
                if (tid < nr) {
                    const int q = p.RSL[MC->base + tid];
                    const float* o = p.SBOX + (int64_t)(MC->base + q) * 6;
                    hit = b[0] - mg <= o[1] && o[0] <= b[1] + mg && b[2] - mg <= o[3] && o[2] <= b[3] + mg &&
                          b[4] - mg <= o[5] && o[4] <= b[5] + mg;
                }
                if (tid == 0) near = 0;
                __syncthreads();
                if (hit) near = 1;
                __syncthreads();
                const bool go = near != 0;
                __syncthreads();  // near is reset by the next cell
                if (!go) continue;
                X = ldu(p.XBX + cb, ln); Y = ldu(p.XBY + cb, ln); Z = ldu(p.XBZ + cb, ln);
            }
            reset = own;
        } else {
            // full search: the current positions become the search positions; their box
            // is kept for later partial searches
            box_of_cell(X, Y, Z, node, part, p.SBOX + (int64_t)slot * 6);
        }
        if (node) {
            uint32_t st = 0;
            const int nc = find_candidates(p, MC, P, c, X, Y, Z, lst, tid, st);
            int32_t* cstore = p.CAND + (size_t)slot * kPMax * kSlots;
            for (int i = 0; i < nc; ++i) stu(cstore, (uint32_t)(i * kSlots) + ln, lst[i][tid]);
            stu(p.NCAND + cb, ln, nc);
            if (reset) { stu(p.XBX + cb, ln, X); stu(p.XBY + cb, ln, Y); stu(p.XBZ + cb, ln, Z); }
            if (st) flag(M, st);
        }
    }
}

// ------------------------------------------------------------------------------
// Conservative force (CalculateConForce, propagate.cu:183-520) fused with the
// first dissipative force (CalculateDisForce with the post-Integrate
// velocities, propagate.cu:523-665) and VelocityUpdateA/B (propagate.cu:714-736),
// over the stored contact candidates.  Modes: M_STEP (FC, V, accepted
// partners), M_CON / M_INIT (FC, accepted partners), M_BONDED (spring +
// pressure + angle -> FC), M_PAIRS (accepted partners only).
// ------------------------------------------------------------------------------
// 6 waves per SIMD (80 VGPRs): 2048 workgroups fit on the chip at once, so 2000
// cells run in one round instead of 1.5.  Measured against 4 waves (100 VGPRs):
// this kernel 28.0 instead of 29.5 us, and k_dis_post, which reads its outputs
// next, 11.5 instead of 12.8 us, even with 18 values spilled to scratch then.  The
// spills are gone: the angles come before the springs, so the spring sums are not
// live across them; the node's position is read again from LDS after them; the
// neighbour, rest-length, rest-angle and gather tables are read at a uniform base
// plus the opaque lane offset where they are used, not through per-lane 64-bit
// pointers kept across the cell loop.
// kLive: walk Pool::LIVE (a multi-system step) instead of the range's slots
template <int kLive>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_con_dis(Pool p, Range r, int32_t mode, const float* __restrict__ rX,
                                                 const float* __restrict__ rY, const float* __restrict__ rZ,
                                                 const float* __restrict__ rVX, const float* __restrict__ rVY,
                                                 const float* __restrict__ rVZ, const float* __restrict__ rBB) {
    __shared__ float sx[kSlots], sy[kSlots], sz[kSlots];
    __shared__ float svx[kSlots], svy[kSlots], svz[kSlots];
    __shared__ struct {
        float angF[kNodes * 18];
    } u;
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    const MeshDev* mesh = p.mesh;
    const bool fused = mode == M_STEP;
    const bool forces = mode != M_PAIRS;
    const bool search = mode != M_BONDED;
    // a multi-system step walks the compact list of live slots (k_live)
    const int nit = kLive ? *p.n_live : r.slot_hi - r.slot_lo;
    for (int it = xcd_block(); it < nit; it += gridDim.x) {
        const int slot = cell_slot<kLive>(p, r, it);
        // node data first (every slot of the range is allocated): their latency
        // overlaps the metadata reads
        const size_t cb = (size_t)slot * kSlots;  // the cell's first node (uniform)
        const uint32_t ln = lane_index();
        const float X = ldu(rX + cb, ln), Y = ldu(rY + cb, ln), Z = ldu(rZ + cb, ln);
        float vX = 0.f, vY = 0.f, vZ = 0.f;
        if (fused) { vX = ldu(rVX + cb, ln); vY = ldu(rVY + cb, ln); vZ = ldu(rVZ + cb, ln); }
        const int nc_pre = search ? ldu(p.NCAND + cb, ln) : 0;
        const int s = __builtin_amdgcn_readfirstlane(sys_of(p, r, slot));
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!sys_live(M)) continue;
        // parameters and the fields of the metadata that this kernel does not write,
        // through scalar loads
        const CMeta MC = cmeta(M);
        const int sbase = MC->base;
        const int c = slot - sbase;
        if (c >= MC->n) continue;
        const CParams P = cparams(p.prm + s);
        sx[tid] = X; sy[tid] = Y; sz[tid] = Z;
        if (fused) { svx[tid] = vX; svy[tid] = vY; svz[tid] = vZ; }
        if (fused && tid == 0) {  // the search of this step has used the reset list
            p.RESET[slot] = 0;
            if (c == 0) M->n_reset = 0;
        }
        __syncthreads();
        uint32_t st = 0;
        const bool angles = forces && P->constrain_angles;
        float FX = 0.f, FY = 0.f, FZ = 0.f;
        if (forces) {
            if (node && (isnan(X) || isnan(Y) || isnan(Z))) st |= CS3D_SYS_NAN;  // propagate.cu:241-248
            if (node) {
                // the node's three bonds: vectors, squared lengths and lengths, used by the
                // node's own three angles (propagate.cu:26-60) and by the springs
                // (propagate.cu:298-314), which compute the same quantities from the same
                // vectors.  The angles write to LDS only, so evaluating them before the
                // springs leaves every sum in its reference order.
                float bx[3], by[3], bz[3], b2[3], bl[3];
                const uint32_t lb = lane_index();
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int N1 = ldu(mesh->nn, (uint32_t)(i * kSlots) + lb);
                    bx[i] = sx[N1] - X; by[i] = sy[N1] - Y; bz[i] = sz[N1] - Z;
                    b2[i] = bx[i] * bx[i] + by[i] * by[i] + bz[i] * bz[i];
                    bl[i] = sqrt_cr(b2[i]);
                }
                if (angles) {  // each angle once (propagate.cu:26-60): (n0,n1), (n1,n2), (n0,n2)
                    const float k = P->angle_k;
                    const uint32_t lt = lane_index();
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const int a = q == 1 ? 1 : 0;
                        const int b = q == 0 ? 1 : 2;
                        st |= angle_pair_l(bx[a], by[a], bz[a], b2[a], bl[a], bx[b], by[b], bz[b], b2[b], bl[b],
                                           ldu(mesh->theta0, 3 * lt + q), k, u.angF + (tid * 3 + q) * 6);
                    }
                }
                // ---- springs, propagate.cu:298-314
                const float stiffness = cld(p.E + slot);
                bool ok = true;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const uint32_t li = (uint32_t)(i * kSlots) + lane_index();
                    float tx, ty, tz;
                    ok &= spring_rcp(stiffness, bl[i], ldu(mesh->R0, li), ldu(mesh->R0inv, li), bx[i], by[i], bz[i],
                                     tx, ty, tz);
                    FX += tx;
                    FY += ty;
                    FZ += tz;
                }
                if (!ok) {  // an operand out of range: the three springs by division
                    slow_path();
                    FX = 0.f; FY = 0.f; FZ = 0.f;
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        float tx, ty, tz;
                        spring_div(stiffness, bl[i], ldu(mesh->R0, (uint32_t)(i * kSlots) + lane_index()), bx[i], by[i],
                                   bz[i], tx, ty, tz);
                        FX += tx;
                        FY += ty;
                        FZ += tz;
                    }
                }
            }
            if (angles) __syncthreads();  // the angle forces of every node are in LDS
            if (node) {
                const float X = lds_reread(sx, tid), Y = lds_reread(sy, tid), Z = lds_reread(sz, tid);
                // ---- pressure, propagate.cu:318-330 (centre of mass from the last CenterOfMass)
                const float rx = X - cld(p.CMX + slot), ry = Y - cld(p.CMY + slot), rz = Z - cld(p.CMZ + slot);
                const float inv = inv_rn(sqrt_cr(rx * rx + ry * ry + rz * rz));
                const float g = 3.0f * cld(p.P + slot);
                FX += g * (inv * rx);
                FY += g * (inv * ry);
                FZ += g * (inv * rz);
            }
            // ---- angle forces gathered in the summation order of CalculateAngleForce (propagate.cu:81-181)
            if (angles && node) {
                const float* a = u.angF + tid * 18;
                float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    ax = ax - (a[q * 6 + 0] + a[q * 6 + 3]);
                    ay = ay - (a[q * 6 + 1] + a[q * 6 + 4]);
                    az = az - (a[q * 6 + 2] + a[q * 6 + 5]);
                }
                const uint32_t lg = lane_index() * 6;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const float* f = u.angF + ldu(mesh->gather, lg + q);
                    ax = ax + f[0];
                    ay = ay + f[1];
                    az = az + f[2];
                }
                FX += ax; FY += ay; FZ += az;
            }
        }
        if (search && node) {
            const float X = lds_reread(sx, tid), Y = lds_reread(sy, tid), Z = lds_reread(sz, tid);
            const int nc = nc_pre;
            const int32_t* cand = p.CAND + (size_t)slot * kPMax * kSlots;  // the cell's lists (uniform)
            const size_t sb = (size_t)sbase * kSlots;                      // the system's first node
            const RuleCtx R = rule_ctx(MC, P);
            PairConsts k;
            k.ar = P->attraction_range;
            k.ar2 = k.ar * k.ar;
            k.rr = P->repulsion_range;
            k.ka = -P->attraction_strength * P->youngs_mod;
            k.kr = +P->repulsion_strength * P->youngs_mod;
            k.gext = P->viscotic_damping;
            float cfx = 0.f, cfy = 0.f, cfz = 0.f;
            float fdx = 0.f, fdy = 0.f, fdz = 0.f;
            int32_t* rec = p.PAIRS + (size_t)slot * kPMax * kSlots;
            const int32_t* sVR = p.VRANGE + (size_t)sbase * 6;
            const float* sBB = rBB + (size_t)sbase * 6;
            int np;
            if (fused) {
                const float vix = svx[tid], viy = svy[tid], viz = svz[tid];
                const float gint = P->internal_damping;
                int nn3[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {  // bond damping, propagate.cu:547-554
                    nn3[i] = ldu(mesh->nn, (uint32_t)(i * kSlots) + lane_index());
                    const int N = nn3[i];
                    fdx = fdx - gint * (vix - svx[N]);
                    fdy = fdy - gint * (viy - svy[N]);
                    fdz = fdz - gint * (viz - svz[N]);
                }
                float nx, ny, nz;
                node_normal(sx, sy, sz, mesh, tid, nn3[0], nn3[1], nn3[2], nx, ny, nz);
                np = accept_forces<true, true>(cand, (int)ln, nc, rec, X, Y, Z, R, sVR, sBB, vix, viy, viz, nx, ny, nz, k,
                                               rX + sb, rY + sb, rZ + sb, rVX + sb, rVY + sb, rVZ + sb, cfx, cfy,
                                               cfz, fdx, fdy, fdz);
                fdx = fdx - P->gamma_visc * vix;  // drag, propagate.cu:658
                fdy = fdy - P->gamma_visc * viy;
                fdz = fdz - P->gamma_visc * viz;
            } else if (forces) {
                np = accept_forces<true, false>(cand, (int)ln, nc, rec, X, Y, Z, R, sVR, sBB, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, k,
                                                rX + sb, rY + sb, rZ + sb, rVX + sb, rVY + sb, rVZ + sb, cfx, cfy,
                                                cfz, fdx, fdy, fdz);
            } else {
                np = accept_forces<false, false>(cand, (int)ln, nc, rec, X, Y, Z, R, sVR, sBB, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                                                 k, rX + sb, rY + sb, rZ + sb, rVX + sb, rVY + sb, rVZ + sb, cfx, cfy,
                                                 cfz, fdx, fdy, fdz);
            }
            stu(pool_late()->NPAIR + cb, ln, np);
            if (forces) {
                FX += cfx; FY += cfy; FZ += cfz;
                if (P->use_rigid_box) {  // propagate.cu:473-510 (lower wall at 0)
                    const float E = P->youngs_mod, thr = P->thresh_dist;
                    float g1 = X, g2 = P->box_max[0] - X;
                    if (g1 < thr) FX += -100.0f * E * (g1 - thr);
                    if (g2 < thr) FX += 100.0f * E * (g2 - thr);
                    g1 = Y; g2 = P->box_max[1] - Y;
                    if (g1 < thr) FY += -100.0f * E * (g1 - thr);
                    if (g2 < thr) FY += 100.0f * E * (g2 - thr);
                    g1 = Z; g2 = P->box_max[2] - Z;
                    if (g1 < thr) FZ += -100.0f * E * (g1 - thr);
                    if (g2 < thr) FZ += 100.0f * E * (g2 - thr);
                }
            }
            if (fused) {
                const float dt = P->dt, m = P->mass;
                const float rdt = sqrt_cr(dt);
                float frx = 0.f, fry = 0.f, frz = 0.f;
                if (P->add_rands) {  // CalculateRanForce for this step (propagate.cu:667-676)
                    float zr[3];
                    noise3(P->noise_seed, (uint32_t)(c * kSlots + tid), MC->step + 1, zr);
                    const float sc = P->rand_scale_factor;
                    frx = sc * zr[0]; fry = sc * zr[1]; frz = sc * zr[2];
                    const CPool q = pool_late();
                    stu(q->FRX + cb, ln, frx); stu(q->FRY + cb, ln, fry); stu(q->FRZ + cb, ln, frz);
                }
                const float vix = svx[tid], viy = svy[tid], viz = svz[tid];
                // VelocityUpdateA (propagate.cu:721-723) then VelocityUpdateB (:732-734)
                float vx = (float)((double)vix + kick_a(dt * FX + rdt * frx, m, MC));
                float vy = (float)((double)viy + kick_a(dt * FY + rdt * fry, m, MC));
                float vz = (float)((double)viz + kick_a(dt * FZ + rdt * frz, m, MC));
                vx = (float)((double)vx + kick_b(fdx, dt, m, MC));
                vy = (float)((double)vy + kick_b(fdy, dt, m, MC));
                vz = (float)((double)vz + kick_b(fdz, dt, m, MC));
                const CPool q = pool_late();
                stu(q->VX + cb, ln, vx); stu(q->VY + cb, ln, vy); stu(q->VZ + cb, ln, vz);
            }
        }
        if (node && forces) {
            const CPool q = pool_late();
            stu(q->FCX + cb, ln, FX); stu(q->FCY + cb, ln, FY); stu(q->FCZ + cb, ln, FZ);
        }
        if (st) flag(M, st);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------
// Bonded part of CalculateConForce alone (springs, pressure, and angles when
// constrainAngles): propagate.cu:253-337, the same arithmetic and summation
// order as k_con_dis, in a lean streaming kernel for the standalone bonded
// entry point.  Without angles it moves x in and F_con out (24 B per node): the
// per-lane mesh data stay in registers across the cells a workgroup walks, and
// the next cell's positions are loaded while the current one is computed.
// ------------------------------------------------------------------------------
template <bool ANGLES>
__global__ __launch_bounds__(192) void k_bonded(Pool p, Range r) {
    __shared__ float sx[kSlots], sy[kSlots], sz[kSlots];
    __shared__ float angF[ANGLES ? kNodes * 18 : 1];
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    const MeshDev* mesh = p.mesh;
    int nb[3] = {0, 0, 0};
    float r0[3] = {1.f, 1.f, 1.f}, ir0[3] = {1.f, 1.f, 1.f}, th[3] = {0.f, 0.f, 0.f};
    if (node) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            nb[i] = mesh->nn[i * kSlots + tid];
            r0[i] = mesh->R0[i * kSlots + tid];
            ir0[i] = mesh->R0inv[i * kSlots + tid];
            if (ANGLES) th[i] = mesh->theta0[3 * tid + i];
        }
    }
    int slot = r.slot_lo + blockIdx.x;
    float X = 0.f, Y = 0.f, Z = 0.f;
    if (slot < r.slot_hi) {
        const int64_t gi = (int64_t)slot * kSlots + tid;
        X = p.X[gi]; Y = p.Y[gi]; Z = p.Z[gi];
    }
    for (; slot < r.slot_hi; slot += gridDim.x) {
        // prefetch the next cell of this workgroup (in bounds: slot < slot_hi <= capacity)
        const int nslot = slot + gridDim.x;
        float Xn = 0.f, Yn = 0.f, Zn = 0.f;
        if (nslot < r.slot_hi) {
            const int64_t gn = (int64_t)nslot * kSlots + tid;
            Xn = p.X[gn]; Yn = p.Y[gn]; Zn = p.Z[gn];
        }
        const int s = sys_of(p, r, slot);
        bool use = s >= 0;
        SysMeta* M = use ? p.meta + s : nullptr;
        const cs3d_params* P = use ? p.prm + s : nullptr;
        use = use && sys_live(M) && slot - M->base < M->n && (P->constrain_angles != 0) == ANGLES;
        if (use) {
            const int64_t gi = (int64_t)slot * kSlots + tid;
            sx[tid] = X; sy[tid] = Y; sz[tid] = Z;
            __syncthreads();
            uint32_t st = 0;
            if (ANGLES) {
                if (node) {
                    const float k = P->angle_k;
#pragma unroll 1
                    for (int q = 0; q < 3; ++q) {
                        const int a = q == 1 ? nb[1] : nb[0];
                        const int b = q == 0 ? nb[1] : nb[2];
                        const float t0 = q == 0 ? th[0] : (q == 1 ? th[1] : th[2]);
                        st |= angle_pair(sx[a] - X, sy[a] - Y, sz[a] - Z, sx[b] - X, sy[b] - Y, sz[b] - Z, t0, k,
                                         angF + (tid * 3 + q) * 6);
                    }
                }
                __syncthreads();
            }
            if (node) {
                if (isnan(X) || isnan(Y) || isnan(Z)) st |= CS3D_SYS_NAN;
                float FX = 0.f, FY = 0.f, FZ = 0.f;
                const float stiffness = p.E[slot];
                float dX[3], dY[3], dZ[3], R[3];
                bool ok = true;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int N1 = nb[i];
                    dX[i] = sx[N1] - X; dY[i] = sy[N1] - Y; dZ[i] = sz[N1] - Z;
                    R[i] = sqrt_cr(dX[i] * dX[i] + dY[i] * dY[i] + dZ[i] * dZ[i]);
                    float tx, ty, tz;
                    ok &= spring_rcp(stiffness, R[i], r0[i], ir0[i], dX[i], dY[i], dZ[i], tx, ty, tz);
                    FX += tx;
                    FY += ty;
                    FZ += tz;
                }
                if (!ok) {  // an operand out of range: the three springs by division
                    slow_path();
                    FX = 0.f; FY = 0.f; FZ = 0.f;
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        float tx, ty, tz;
                        spring_div(stiffness, R[i], r0[i], dX[i], dY[i], dZ[i], tx, ty, tz);
                        FX += tx;
                        FY += ty;
                        FZ += tz;
                    }
                }
                const float rx = X - p.CMX[slot], ry = Y - p.CMY[slot], rz = Z - p.CMZ[slot];
                const float inv = inv_rn(sqrt_cr(rx * rx + ry * ry + rz * rz));
                const float g = 3.0f * p.P[slot];
                FX += g * (inv * rx);
                FY += g * (inv * ry);
                FZ += g * (inv * rz);
                if (ANGLES) {
                    const float* a = angF + tid * 18;
                    float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        ax = ax - (a[q * 6 + 0] + a[q * 6 + 3]);
                        ay = ay - (a[q * 6 + 1] + a[q * 6 + 4]);
                        az = az - (a[q * 6 + 2] + a[q * 6 + 5]);
                    }
                    const int32_t* gt = mesh->gather + tid * 6;
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        const float* f = angF + gt[q];
                        ax = ax + f[0];
                        ay = ay + f[1];
                        az = az + f[2];
                    }
                    FX += ax; FY += ay; FZ += az;
                }
                p.FCX[gi] = FX; p.FCY[gi] = FY; p.FCZ[gi] = FZ;
            }
            if (st) flag(M, st);
            __syncthreads();
        }
        X = Xn; Y = Yn; Z = Zn;
    }
}

// ------------------------------------------------------------------------------
// Dissipative force (CalculateDisForce, propagate.cu:523-665) with the
// end-of-step velocities, then CenterOfMass (centermass.cu:8-50) and, on growth
// steps, volumes (volume.cu:5-163).  Modes: M_STEP, M_DIS_V, M_DIS_VA, M_COM,
// M_VOLUMES.
// ------------------------------------------------------------------------------
// kLive: walk Pool::LIVE (a multi-system step) instead of the range's slots
template <int kLive>
__global__ __launch_bounds__(192) void k_dis_post(Pool p, Range r, int32_t mode, int64_t step_override,
                                                  int32_t clear_bins, const float* __restrict__ rX,
                                                  const float* __restrict__ rY, const float* __restrict__ rZ,
                                                  const float* __restrict__ rVX, const float* __restrict__ rVY,
                                                  const float* __restrict__ rVZ) {
    __shared__ float sx[kSlots], sy[kSlots], sz[kSlots];
    __shared__ float svx[kSlots], svy[kSlots], svz[kSlots];
    __shared__ float red[3][128];
    __shared__ float cm[3];
    __shared__ float bpart[3][6];
    __shared__ int32_t moved;
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    const MeshDev* mesh = p.mesh;
    const bool do_dis = mode == M_STEP || mode == M_DIS_V || mode == M_DIS_VA;
    const bool fused = mode == M_STEP;
    // a multi-system step walks the compact list of live slots (k_live)
    const int nit = kLive ? *p.n_live : r.slot_hi - r.slot_lo;
    for (int it = xcd_block(); it < nit; it += gridDim.x) {
        const int slot = cell_slot<kLive>(p, r, it);
        // node data first (every slot of the range is allocated): their latency
        // overlaps the metadata reads
        const size_t cb = (size_t)slot * kSlots;  // the cell's first node (uniform)
        const uint32_t ln = lane_index();
        const float X = ldu(rX + cb, ln), Y = ldu(rY + cb, ln), Z = ldu(rZ + cb, ln);
        float vX = 0.f, vY = 0.f, vZ = 0.f;
        if (do_dis) { vX = ldu(rVX + cb, ln); vY = ldu(rVY + cb, ln); vZ = ldu(rVZ + cb, ln); }
        const int np = (do_dis && node) ? ldu(p.NPAIR + cb, ln) : 0;
        float fcX = 0.f, fcY = 0.f, fcZ = 0.f, xbX = 0.f, xbY = 0.f, xbZ = 0.f;
        if (fused && do_dis && node) {
            fcX = ldu(p.FCX + cb, ln); fcY = ldu(p.FCY + cb, ln); fcZ = ldu(p.FCZ + cb, ln);
            xbX = ldu(p.XBX + cb, ln); xbY = ldu(p.XBY + cb, ln); xbZ = ldu(p.XBZ + cb, ln);
        }
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        // metadata and parameters through scalar loads: the fields read here are not
        // written by this kernel (search_next and n_reset, which it updates by atomics,
        // are read by atomic loads)
        SysMeta* M = p.meta + s;
        const CMeta MC = cmeta(M);
        const CParams P = cparams(p.prm + s);
        if (!sys_live(MC)) continue;
        const int c = slot - MC->base;
        if (c >= MC->n) continue;
        if (tid == 0) moved = 0;
        sx[tid] = X; sy[tid] = Y; sz[tid] = Z;
        if (do_dis) { svx[tid] = vX; svy[tid] = vY; svz[tid] = vZ; }
        __syncthreads();
        uint32_t st = 0;
        bool far = false;
        if (do_dis && node) {
            float fx = 0.f, fy = 0.f, fz = 0.f;
            const float vix = svx[tid], viy = svy[tid], viz = svz[tid];
            const float gint = P->internal_damping;
            int nn3[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {  // bond damping, propagate.cu:547-554
                const int N = mesh->nn[i * kSlots + tid];
                nn3[i] = N;
                fx = fx - gint * (vix - svx[N]);
                fy = fy - gint * (viy - svy[N]);
                fz = fz - gint * (viz - svz[N]);
            }
            float nx, ny, nz;
            node_normal(sx, sy, sz, mesh, tid, nn3[0], nn3[1], nn3[2], nx, ny, nz);
            PairConsts k;
            k.ar = P->attraction_range;
            k.ar2 = k.ar * k.ar;
            k.rr = P->repulsion_range;
            k.ka = 0.f;
            k.kr = 0.f;
            k.gext = P->viscotic_damping;
            const int32_t* lst = p.PAIRS + (size_t)slot * kPMax * kSlots;
            const size_t sb = (size_t)MC->base * kSlots;  // the system's first node
            pair_friction(lst, (int)ln, np, vix, viy, viz, nx, ny, nz, k.gext, rVX + sb, rVY + sb, rVZ + sb, fx, fy, fz);
            fx = fx - P->gamma_visc * vix;  // drag, propagate.cu:658
            fy = fy - P->gamma_visc * viy;
            fz = fz - P->gamma_visc * viz;
            const CPool q = pool_late();  // pointers read again where used (see pool_late)
            stu(q->FDX + cb, ln, fx); stu(q->FDY + cb, ln, fy); stu(q->FDZ + cb, ln, fz);
            if (fused) {
                // the next step's Integrate moves this node to X + v' dt with v' from the
                // values now in V, F_con and F_dis (bit-identical to k_integrate): search
                // again next step when that is pair_skin/2 or more from the last search
                const float dt = P->dt, m = P->mass, rdt = sqrt_cr(dt);
                const bool rnd = P->add_rands != 0;
                const float qx = X + half_kick(vix, fcX, fx, rnd ? ldu(q->FRX + cb, ln) : 0.f, dt, rdt, m, MC) * dt - xbX;
                const float qy = Y + half_kick(viy, fcY, fy, rnd ? ldu(q->FRY + cb, ln) : 0.f, dt, rdt, m, MC) * dt - xbY;
                const float qz = Z + half_kick(viz, fcZ, fz, rnd ? ldu(q->FRZ + cb, ln) : 0.f, dt, rdt, m, MC) * dt - xbZ;
                const float h = 0.5f * P->pair_skin;
                far = !(qx * qx + qy * qy + qz * qz <= h * h);
            }
        }
        // a node of this cell leaves the skin next step: the cell is listed for the next
        // (partial) search, up to the list size (one atomic per such cell); the flag
        // atomic is skipped once it is visibly set
        if (far) moved = 1;
        __syncthreads();
        if (fused && tid == 0 && moved) {
            pool_late()->RESET[slot] = 1;
            if (__hip_atomic_load(&M->search_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                atomicOr(&M->search_next, 1);
            if (__hip_atomic_load(&M->n_reset, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= kMaxPartial) {
                const int k = atomicAdd(&M->n_reset, 1);
                if (k < kMaxPartial) pool_late()->RSL[MC->base + k] = c;
            }
        }
        if (clear_bins && (mode != M_STEP || MC->need_search)) {
            // the search of this step is done: this cell's share of the system's
            // search-bin counters (those of the current bin grid, which every
            // insertion since the last clear used) is cleared in one coalesced run
            const int64_t nb = min((int64_t)MC->bbudget, (int64_t)MC->bdims[0] * MC->bdims[1] * MC->bdims[2]);
            const int64_t per = (nb + MC->n - 1) / MC->n;
            const int64_t lo = (int64_t)c * per, hi = min(nb, lo + per);
            int32_t* counts = pool_late()->bins + MC->bbase;
            for (int64_t b = lo + tid; b < hi; b += blockDim.x) counts[b] = 0;
        }
        if (fused) {  // this cell's bbox at the end of the step (bounding_boxes, propagatebound.cu:10-141)
            float b[6] = {node ? X : INFINITY, node ? X : -INFINITY, node ? Y : INFINITY,
                          node ? Y : -INFINITY, node ? Z : INFINITY, node ? Z : -INFINITY};
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
                for (int a = 0; a < 6; a += 2) {
                    b[a] = fminf(b[a], __shfl_xor(b[a], o));
                    b[a + 1] = fmaxf(b[a + 1], __shfl_xor(b[a + 1], o));
                }
            }
            if ((tid & 63) == 0)
#pragma unroll
                for (int a = 0; a < 6; ++a) bpart[tid >> 6][a] = b[a];
        }  // combined after the next barrier

        const bool do_com = mode == M_STEP || mode == M_COM;
        const int64_t stepno = step_override >= 0 ? step_override : MC->step + 1;
        const bool do_vol =
            mode == M_VOLUMES || (mode == M_STEP && stepno <= P->div_time_steps && rgrowth(P, MC->n) > 0.0f);
        if (do_com) {
            // 256-slot tree of centermass.cu:32-41: levels 128 and 64 through LDS, 32..1 by shuffles
            if (tid < 128) {
                const bool hi = tid + 128 < kNodes;
                red[0][tid] = sx[tid] + (hi ? sx[tid + 128] : 0.0f);
                red[1][tid] = sy[tid] + (hi ? sy[tid + 128] : 0.0f);
                red[2][tid] = sz[tid] + (hi ? sz[tid + 128] : 0.0f);
            }
            __syncthreads();
            if (tid < 64) {
                float a = red[0][tid] + red[0][tid + 64];
                float b = red[1][tid] + red[1][tid + 64];
                float d = red[2][tid] + red[2][tid + 64];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    a = a + __shfl_down(a, o);
                    b = b + __shfl_down(b, o);
                    d = d + __shfl_down(d, o);
                }
                if (tid == 0) {
                    cm[0] = a / 180.0f; cm[1] = b / 180.0f; cm[2] = d / 180.0f;
                    const CPool q = pool_late();
                    q->CMX[slot] = cm[0]; q->CMY[slot] = cm[1]; q->CMZ[slot] = cm[2];
                }
            }
            if (fused && tid < 6) {
                float b = bpart[0][tid];
                b = (tid & 1) ? fmaxf(b, bpart[1][tid]) : fminf(b, bpart[1][tid]);
                b = (tid & 1) ? fmaxf(b, bpart[2][tid]) : fminf(b, bpart[2][tid]);
                pool_late()->BBOXN[(int64_t)slot * 6 + tid] = b;
            }
            __syncthreads();
        } else if (do_vol) {
            if (tid == 0) {
                const CPool q = pool_late();
                cm[0] = q->CMX[slot]; cm[1] = q->CMY[slot]; cm[2] = q->CMZ[slot];
            }
            __syncthreads();
        }
        if (do_vol) {
            // volume.cu:22-26: local coordinates about the (fresh) centre of mass
            const float qx = X - cm[0], qy = Y - cm[1], qz = Z - cm[2];
            __syncthreads();
            sx[tid] = qx; sy[tid] = qy; sz[tid] = qz;
            __syncthreads();
            const bool sph = P->check_sphericity != 0;
            float totvol = 0.f, faceArea = 0.f;
            if (tid < CS3D_FACES) {  // volume.cu:37-119
                const int32_t* rr = mesh->faces + 7 * tid;
                float avX = 0.0f, avY = 0.0f, avZ = 0.0f;
                for (int i = 0; i < 5; ++i) { avX += sx[rr[i]]; avY += sy[rr[i]]; avZ += sz[rr[i]]; }
                float fac = 0.166666667f;
                if (tid < 12) fac = 0.2f;
                else { avX += sx[rr[5]]; avY += sy[rr[5]]; avZ += sz[rr[5]]; }
                avX *= fac; avY *= fac; avZ *= fac;
                for (int i = 0; i < 6; ++i) {
                    const int a = rr[i], b = rr[i + 1];
                    const float n1 = (sy[b] * avZ - avY * sz[b]) * sx[a];
                    const float n2 = (sz[b] * avX - avZ * sx[b]) * sy[a];
                    const float n3 = (sx[b] * avY - avX * sy[b]) * sz[a];
                    totvol += fabsf(n1 + n2 + n3);
                    if (sph) {
                        const float x1 = sx[a] - avX, y1 = sy[a] - avY, z1 = sz[a] - avZ;
                        const float x2 = sx[b] - avX, y2 = sy[b] - avY, z2 = sz[b] - avZ;
                        const float xx = y1 * z2 - z1 * y2, yy = z1 * x2 - x1 * z2, zz = x1 * y2 - y1 * x2;
                        faceArea = (float)((double)faceArea + 0.5 * (double)sqrt_cr(xx * xx + yy * yy + zz * zz));
                    }
                }
            }
            // canonical 128-slot tree over the faces (faces >= 92 contribute 0)
            if (tid < 128) {
                red[0][tid] = tid < CS3D_FACES ? totvol : 0.0f;
                red[1][tid] = tid < CS3D_FACES ? faceArea : 0.0f;
            }
            __syncthreads();
            if (tid < 64) {
                float a = red[0][tid] + red[0][tid + 64];
                float b = red[1][tid] + red[1][tid + 64];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    a = a + __shfl_down(a, o);
                    b = b + __shfl_down(b, o);
                }
                if (tid == 0) {  // volume.cu:129-161
                    const float volume = (float)((double)a / 6.0);
                    pool_late()->VOL[slot] = volume;
                    if (!isfinite(volume)) st |= CS3D_SYS_VOLUME_NONFINITE;
                    int8_t fl = volume > P->division_vol ? 1 : 0;
                    if (sph) {
                        pool_late()->AREA[slot] = b;
                        const float cbr = cbrtf(volume);
                        const float psi = (float)(4.835975862049408 * (double)cbr * (double)cbr / (double)b);
                        if ((double)fabsf(1.0f - psi) > 0.05) fl = 0;
                    }
                    if (P->use_different_stiffnesses && P->recalc_r0 && pool_late()->E[slot] != P->stiffness1) fl = 0;
                    pool_late()->FLAG[slot] = fl;
                }
            }
        }
        if (st) flag(M, st);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------
// Device RANMAR (marsaglia.h:115-144) and plane normals (RandomVector.h:8-78).
// The generator values are multiples of 2^-24 and every operation is exact in
// float, so the device stream equals the host one.
// ------------------------------------------------------------------------------
__device__ float ranmar1(cs3d_ranmar_state* st) {
    float uni = st->u[st->i97] - st->u[st->j97];
    if (uni < 0.0f) uni += 1.0f;
    st->u[st->i97] = uni;
    st->i97--;
    if (st->i97 == 0) st->i97 = 97;
    st->j97--;
    if (st->j97 == 0) st->j97 = 97;
    st->c -= st->cd;
    if (st->c < 0.0f) st->c += st->cm;
    uni -= st->c;
    if (uni < 0.0f) uni += 1.0f;
    return uni;
}

__device__ void random_vector(cs3d_ranmar_state* st, float* n) {
    float x1, x2, s;
    do {
        x1 = ranmar1(st);
        x2 = ranmar1(st);
        x1 = 2.0f * x1 - 1.0f;
        x2 = 2.0f * x2 - 1.0f;
        s = x1 * x1 + x2 * x2;
    } while (s >= 1.0f);
    n[0] = 2.0f * x1 * sqrt_cr(1.0f - x1 * x1 - x2 * x2);
    n[1] = 2.0f * x2 * sqrt_cr(1.0f - x1 * x1 - x2 * x2);
    n[2] = 1.0f - 2.0f * (x1 * x1 + x2 * x2);
}

__device__ void random_vector_basis(cs3d_ranmar_state* st, float* n, const float* basis) {
    const float r = ranmar1(st);
    const float Nv = (float)cos(2 * 3.14159 * (double)r);
    const float Nw = (float)sin(2 * 3.14159 * (double)r);
    float v[3], w[3];
    if (basis[1] != 0) {
        v[0] = 0; v[1] = basis[2]; v[2] = -1 * basis[1];
        w[0] = basis[1]; w[1] = -1 * basis[0]; w[2] = 0;
    } else {
        v[0] = 0; v[1] = 1; v[2] = 0;
        w[0] = basis[2]; w[1] = 0; w[2] = -1 * basis[0];
    }
    float f = (w[0] * v[0] + w[1] * v[1] + w[2] * w[2]) / (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    w[0] = w[0] - f * v[0]; w[1] = w[1] - f * v[1]; w[2] = w[2] - f * v[2];
    f = sqrt_cr(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] / f; v[1] = v[1] / f; v[2] = v[2] / f;
    f = sqrt_cr(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    w[0] = w[0] / f; w[1] = w[1] / f; w[2] = w[2] / f;
    n[0] = Nv * v[0] + Nw * w[0];
    n[1] = Nv * v[1] + Nw * w[1];
    n[2] = Nv * v[2] + Nw * w[2];
}

// ------------------------------------------------------------------------------
// Division phase of one system per workgroup (256 lanes): count_and_get_div
// (GPUbounce.cu:2489-2500) as an ascending compaction of the flags, one RANMAR
// normal per event in event order, cell_division (postscriptinit.cu:185-256)
// with the daughter at index n, PressureReset (PressureKernels.cu:45-65) and
// the daughter stiffness rule (GPUbounce.cu:1254-1260).  In M_STEP it also ends
// the step: step counter, and whether the grid is rebuilt this step.
// ------------------------------------------------------------------------------
// min (even a) / max (odd a) of 12 values over a 256-lane workgroup; the result
// lands in out[0..12) (LDS), visible after the trailing barrier
__device__ __forceinline__ void block_minmax12(float* v, float (*part)[12], float* out) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 12; a += 2) {
            v[a] = fminf(v[a], __shfl_xor(v[a], o));
            v[a + 1] = fmaxf(v[a + 1], __shfl_xor(v[a + 1], o));
        }
    }
    if ((tid & 63) == 0)
#pragma unroll
        for (int a = 0; a < 12; ++a) part[tid >> 6][a] = v[a];
    __syncthreads();
    if (tid < 12) {
        float b = part[0][tid];
        const int nw = (int)(blockDim.x >> 6);
        for (int w = 1; w < nw; ++w) b = (tid & 1) ? fmaxf(b, part[w][tid]) : fminf(b, part[w][tid]);
        out[tid] = b;
    }
    __syncthreads();
}

// Division phase of system s by one 256-lane workgroup (every lane calls it; the
// early returns depend on the system only, so the workgroup stays together).
__device__ void divide_system(const Pool& p, int s, int32_t mode) {
    __shared__ int32_t wsum[4];
    __shared__ int32_t nev, ncur, stop;
    __shared__ float nrm[3];
    __shared__ int32_t bad;
    __shared__ float redc[3][128];
    __shared__ float mpart[4][12];
    __shared__ float mres[12];
    __shared__ int64_t nclear;
    const int tid = threadIdx.x;
    SysMeta* M = p.meta + s;
    const bool step_mode = mode == M_STEP;
    // the metadata and parameters that a step without divisions needs, read up front
    // in one batch of scalar loads (this kernel writes these fields only after
    // reading them): one memory round trip instead of one per branch
    const CMeta MC = cmeta(M);
    const CParams PC = cparams(p.prm + s);
    const int32_t m_active = MC->active, n0 = MC->n, base = MC->base, cap = MC->cap;
    const int64_t m_step = MC->step, m_horizon = MC->horizon, m_cell_steps = MC->cell_steps;
    const uint32_t m_status = MC->status;
    const int32_t m_nreset = MC->n_reset, m_search_next = MC->search_next;
    const int64_t div_time = PC->div_time_steps;
    const int32_t m_gint = PC->grid_interval, m_ccom = PC->correct_com;
    const float m_rgrowth = rgrowth(PC, n0);
    const bool live = m_active && (m_horizon < 0 || m_step < m_horizon) && !(m_status & CS3D_SYS_FATAL_MASK);
    // a stopped system keeps a pending grid adoption: its bboxes are adopted when it
    // resumes or when a caller reads the grid
    if (!m_active || (step_mode && !live)) {
        if (tid == 0 && step_mode) M->shift_pending = 0;
        return;
    }
    const cs3d_params* P = p.prm + s;
    const int64_t stepno = m_step + 1;
    bool growth = true;
    if (step_mode) growth = stepno <= div_time && m_rgrowth > 0.0f;
    if (tid == 0) { nev = 0; ncur = n0; stop = 0; bad = 0; }
    __syncthreads();
    if (growth) {
        // ascending event list in EVL[base .. base + nev), 4096 cells per pass: each
        // thread loads the flags of 16 consecutive cells at once and counts them, a scan
        // of the counts (shuffles within a wave, then the four wave totals) gives the
        // thread's first event, and it lists its events in cell order
        constexpr int kRun = 16;
        int listed = 0;  // events listed by the passes before (the same in every thread)
        for (int c0 = 0; c0 < n0; c0 += kRun * 256) {
            const int cr = c0 + kRun * tid;
            int8_t f[kRun];
#pragma unroll
            for (int j = 0; j < kRun; ++j) f[j] = cr + j < n0 ? p.FLAG[base + cr + j] : (int8_t)0;
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < kRun; ++j) cnt += f[j] == 1 ? 1 : 0;
            const int lane = tid & 63, wave = tid >> 6;
            int inc = cnt;  // inclusive scan within the wave
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            if (lane == 63) wsum[wave] = inc;
            __syncthreads();
            int e = listed + inc - cnt;
            for (int w = 0; w < wave; ++w) e += wsum[w];
            if (cnt) {
#pragma unroll
                for (int j = 0; j < kRun; ++j)
                    if (f[j] == 1) {
                        p.EVL[base + e++] = cr + j;
                        p.FLAG[base + cr + j] = 0;
                    }
            }
            listed += wsum[0] + wsum[1] + wsum[2] + wsum[3];
            __syncthreads();  // wsum is rewritten by the next pass
        }
        if (tid == 0) nev = listed;
        __syncthreads();
        const int ne = nev;
        const float half = (float)(0.5 * (double)P->repulsion_range);
        for (int e = 0; e < ne; ++e) {
            if (tid == 0) {
                if (P->use_div_plane_basis) random_vector_basis(&M->rng, nrm, P->div_plane_basis);
                else random_vector(&M->rng, nrm);
                const float nn = sqrt_cr(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                if ((double)fabsf(nn - 1) > 1e-3) bad |= CS3D_SYS_BAD_NORMAL;
                if (ncur >= cap) { bad |= CS3D_SYS_CAPACITY; stop = 1; }
            }
            __syncthreads();
            if (stop) break;
            const int rank = p.EVL[base + e];
            const int d = ncur;
            const int64_t pi = (int64_t)(base + rank) * kSlots + tid;
            const int64_t di = (int64_t)(base + d) * kSlots + tid;
            const float cmx = p.CMX[base + rank], cmy = p.CMY[base + rank], cmz = p.CMZ[base + rank];
            // bboxes of parent and daughter after the split (min/max pairs: parent x, y,
            // z, then daughter x, y, z)
            float bx[12];
#pragma unroll
            for (int a = 0; a < 12; a += 2) { bx[a] = INFINITY; bx[a + 1] = -INFINITY; }
            if (tid < kNodes) {
                const float x0 = p.X[pi], y0 = p.Y[pi], z0 = p.Z[pi];
#pragma unroll
                for (int side = 0; side < 2; ++side) {  // CalcAndUpdateDaughtPos, postscriptinit.cu:148-182
                    const float Nx = side ? -1 * nrm[0] : nrm[0];
                    const float Ny = side ? -1 * nrm[1] : nrm[1];
                    const float Nz = side ? -1 * nrm[2] : nrm[2];
                    float x = x0 - cmx - half * Nx, y = y0 - cmy - half * Ny, z = z0 - cmz - half * Nz;
                    const float pd = x * Nx + y * Ny + z * Nz;
                    if (pd < 0.0f || pd == 0) { x = x - pd * Nx; y = y - pd * Ny; z = z - pd * Nz; }
                    x = x + (cmx + half * Nx); y = y + (cmy + half * Ny); z = z + (cmz + half * Nz);
                    const int64_t o = side ? di : pi;
                    p.X[o] = x; p.Y[o] = y; p.Z[o] = z;
                    bx[6 * side + 0] = x; bx[6 * side + 1] = x;
                    bx[6 * side + 2] = y; bx[6 * side + 3] = y;
                    bx[6 * side + 4] = z; bx[6 * side + 5] = z;
                }
                p.VX[di] = p.VX[pi]; p.VY[di] = p.VY[pi]; p.VZ[di] = p.VZ[pi];
            } else if (tid < kSlots) {
                p.X[di] = 0.f; p.Y[di] = 0.f; p.Z[di] = 0.f;
                p.VX[di] = 0.f; p.VY[di] = 0.f; p.VZ[di] = 0.f;
            }
            if (tid < kSlots) {  // a new cell's forces start at zero (never-written slots in the reference)
                p.FCX[di] = 0.f; p.FCY[di] = 0.f; p.FCZ[di] = 0.f;
                p.FDX[di] = 0.f; p.FDY[di] = 0.f; p.FDZ[di] = 0.f;
            }
            if (step_mode) {
                block_minmax12(bx, mpart, mres);
                if (tid < 12) p.BBOXN[(int64_t)(base + (tid < 6 ? rank : d)) * 6 + tid % 6] = mres[tid];
            }
            if (tid == 0) {
                const int gd = base + d, gp = base + rank;
                p.E[gd] = P->daught_same_stiffness ? p.E[gp] : P->stiffness1;
                p.GS[gd] = p.GS[gp];
                p.CMX[gd] = 0.f; p.CMY[gd] = 0.f; p.CMZ[gd] = 0.f;
                p.P[gp] = P->min_pressure;
                p.P[gd] = P->min_pressure;
                ncur = d + 1;
            }
            __syncthreads();
        }
        if (P->daughter_com_fix) {  // opt-in: centre of mass of parent and daughter from their new nodes
            const int done = ncur - n0;
            for (int e = 0; e < done; ++e) {
                for (int q = 0; q < 2; ++q) {
                    const int cell = q ? n0 + e : p.EVL[base + e];
                    const int64_t o = (int64_t)(base + cell) * kSlots;
                    if (tid < 128) {  // same 256-slot tree as CenterOfMass
                        const bool hi = tid + 128 < kNodes;
                        redc[0][tid] = p.X[o + tid] + (hi ? p.X[o + tid + 128] : 0.0f);
                        redc[1][tid] = p.Y[o + tid] + (hi ? p.Y[o + tid + 128] : 0.0f);
                        redc[2][tid] = p.Z[o + tid] + (hi ? p.Z[o + tid + 128] : 0.0f);
                    }
                    __syncthreads();
                    if (tid < 64) {
                        float a = redc[0][tid] + redc[0][tid + 64];
                        float b = redc[1][tid] + redc[1][tid + 64];
                        float d = redc[2][tid] + redc[2][tid + 64];
#pragma unroll
                        for (int k = 32; k > 0; k >>= 1) {
                            a = a + __shfl_down(a, k);
                            b = b + __shfl_down(b, k);
                            d = d + __shfl_down(d, k);
                        }
                        if (tid == 0) {
                            p.CMX[base + cell] = a / 180.0f;
                            p.CMY[base + cell] = b / 180.0f;
                            p.CMZ[base + cell] = d / 180.0f;
                        }
                    }
                    __syncthreads();
                }
            }
        }
    }
    __syncthreads();
    const int events = ncur - n0;
    const int gint = m_gint > 0 ? m_gint : 1;
    const bool rebuild = step_mode && (events > 0 || stepno % gint == 0);
    if (rebuild) {
        // grid rebuild from the end-of-step bboxes: minmaxpre/minmaxpost
        // (propagatebound.cu:145-204, 266-323, sentinels +-1e8) and the grid
        // geometry; the per-cell part (bbox adoption, voxel counts) is done by the
        // next step's Integrate
        float e[12];
#pragma unroll
        for (int a = 0; a < 12; a += 2) { e[a] = 1.0E8f; e[a + 1] = -1.0E8f; }
        // the boxes of 8 cells per thread are loaded before any is reduced (min and max
        // are exact in any order)
        constexpr int kUnroll = 8;
        for (int c0 = 0; c0 < ncur; c0 += kUnroll * 256) {
            float bb[kUnroll][6];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const int c = c0 + u * 256 + tid;
                const float* b = p.BBOXN + (int64_t)(base + min(c, ncur - 1)) * 6;
#pragma unroll
                for (int a = 0; a < 6; ++a) bb[u][a] = b[a];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (c0 + u * 256 + tid < ncur) {
#pragma unroll
                    for (int a = 0; a < 6; a += 2) {
                        e[a] = fminf(e[a], bb[u][a]);
                        e[a + 1] = fmaxf(e[a + 1], bb[u][a + 1]);
                    }
                }
            }
        }
        block_minmax12(e, mpart, mres);
        if (tid == 0) nclear = set_grid_geometry(M, P, mres);
        __syncthreads();
        int32_t* counts = p.vox + M->vbase;
        for (int64_t v = tid; v < nclear; v += blockDim.x) counts[v] = 0;
    }
    const bool shift = step_mode && m_ccom;
    if (shift) {
        // calc_sys_CM (GPUbounce.cu:2504-2520): float sums of the cell centres in
        // cell order (daughters of this step still hold 0), divided by the count;
        // the nodes are shifted by k_com_shift (CorrectCoMMotion, propagate.cu:766-776)
        float acc[3] = {0.f, 0.f, 0.f};
        for (int c0 = 0; c0 < ncur; c0 += 128) {
            if (tid < 128 && c0 + tid < ncur) {
                redc[0][tid] = p.CMX[base + c0 + tid];
                redc[1][tid] = p.CMY[base + c0 + tid];
                redc[2][tid] = p.CMZ[base + c0 + tid];
            }
            __syncthreads();
            if (tid < 3) {
                const int m = min(128, ncur - c0);
                for (int k = 0; k < m; ++k) acc[tid] = acc[tid] + redc[tid][k];
            }
            __syncthreads();
        }
        if (tid < 3) M->sys_cm[tid] = acc[tid] / (float)ncur;
    }
    if (tid == 0) {
        M->shift_pending = shift ? 1 : 0;
        if (bad) atomicOr(&M->status, (uint32_t)bad);
        M->n_events = events;
        M->n = ncur;
        if (step_mode) {
            M->rebuild = rebuild ? 1 : 0;
            M->step = stepno;
            M->cell_steps = m_cell_steps + n0;  // the n0 cells that were stepped
            // After divisions, the divided cells and the cells listed by k_dis_post (a
            // node leaves the skin next step) are searched again next step, with the
            // cells near them (partial search).  Without divisions a node leaving the
            // skin means a full search: it resets every search position, so the next one
            // comes many steps later, where partial searches of the cells that cross the
            // skin one after another would search on almost every step.  A full search
            // also when the list is too long, or when CorrectCoMMotion shifts the nodes
            // (their search positions are not shifted).
            // RSL holds one entry per cell slot of the system (M->cap entries at
            // RSL[base]): a list that would not fit there means a full search
            int nr = m_nreset;
            const int lim = min(kMaxPartial, cap);
            bool full = (events == 0 && m_search_next) || nr > lim || nr + 2 * events > lim ||
                        (shift && (events > 0 || m_search_next));
            if (!full)
                for (int e = 0; e < events; ++e) {
                    const int pc = p.EVL[base + e], dc = n0 + e;
                    p.RSL[base + nr++] = pc;
                    p.RSL[base + nr++] = dc;
                    p.RESET[base + pc] = 1;
                    p.RESET[base + dc] = 1;
                }
            M->need_search = full ? 1 : (nr > 0 ? 2 : 0);
            M->n_reset = full ? 0 : nr;
            M->search_next = 0;
        } else if (events > 0) {
            M->need_search = 1;
        }
    }
    __syncthreads();  // the shared state is reused by the workgroup's next system
}

// One workgroup per system, striding over the systems of the range in use.
__global__ __launch_bounds__(256) void k_sys_divide(Pool p, Range r, int32_t mode) {
    const int hi = r.one_sys ? r.sys_hi : min(r.sys_hi, *p.nsys);
    for (int s = r.sys_lo + blockIdx.x; s < hi; s += gridDim.x) divide_system(p, s, mode);
}

// Compact list of the live cell slots of a multi-system range (cells c < n of the
// systems that step), rebuilt at the start of every step: the step's cell kernels
// walk it instead of every slot of each system's room.  Every workgroup computes
// the offsets of all systems in use (an exclusive scan of their live counts, 1024
// at a time) and writes the slots of the systems s with s mod gridDim.x =
// blockIdx.x, all lanes on one system at a time; workgroup 0 writes the count.
__device__ __forceinline__ int live_cells(const SysMeta* m) { return sys_live(m) ? m->n : 0; }

constexpr int kLiveThreads = 1024;

__global__ __launch_bounds__(kLiveThreads) void k_live(Pool p, Range r) {
    __shared__ int32_t wsum[kLiveThreads / 64];
    __shared__ int32_t offs[kLiveThreads];
    __shared__ int32_t carry;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lo = r.sys_lo, hi = min(r.sys_hi, *p.nsys);
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int c0 = lo; c0 < hi; c0 += kLiveThreads) {
        const int s = c0 + tid;
        const int n = s < hi ? live_cells(p.meta + s) : 0;
        // inclusive scan within the wave, then over the wave totals
        int v = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o);
            if (lane >= o) v += t;
        }
        if (lane == 63) wsum[wv] = v;
        __syncthreads();
        if (tid == 0) {
            int acc = carry;
            for (int w = 0; w < kLiveThreads / 64; ++w) {
                const int t = wsum[w];
                wsum[w] = acc;
                acc += t;
            }
            carry = acc;
        }
        __syncthreads();
        offs[tid] = wsum[wv] + v - n;  // exclusive offset of system s
        __syncthreads();
        const int m = min(kLiveThreads, hi - c0);
        for (int k = 0; k < m; ++k) {
            const int sk = c0 + k;
            if (sk % (int)gridDim.x != (int)blockIdx.x) continue;
            const SysMeta* M = p.meta + sk;
            const int nk = live_cells(M);
            const int off = offs[k], base = M->base;
            for (int c = tid; c < nk; c += kLiveThreads) p.LIVE[off + c] = base + c;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0 && tid == 0) *p.n_live = carry;
}

// Sets the count of system ids in use (stream-ordered with the steps).
__global__ void k_set_nsys(Pool p, int32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p.nsys = n;
}

// CorrectCoMMotion (propagate.cu:766-776) for the systems whose division phase
// just computed the system centre: every node of every cell minus sys_cm.  The
// reference's kernel also shifts the 12 padding slots of each cell; they are no
// mesh nodes and stay 0 here.  The bboxes and cell centres of the step are left
// as they are, as in the reference (they were computed before the shift).
__global__ __launch_bounds__(192) void k_com_shift(Pool p, Range r) {
    const int tid = threadIdx.x;
    if (tid >= kNodes) return;
    for (int slot = r.slot_lo + xcd_block(); slot < r.slot_hi; slot += gridDim.x) {
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        const SysMeta* M = p.meta + s;
        if (!M->shift_pending || slot - M->base >= M->n) continue;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        p.X[gi] = p.X[gi] - M->sys_cm[0];
        p.Y[gi] = p.Y[gi] - M->sys_cm[1];
        p.Z[gi] = p.Z[gi] - M->sys_cm[2];
    }
}

// ------------------------------------------------------------------------------
// Grid rebuild: bounding_boxes (propagatebound.cu:10-141), minmaxpre/minmaxpost
// (:145-204, 266-323) as order-free min/max (one atomic per workgroup and
// system), host Xdiv/Ydiv/Zdiv (GPUbounce.cu:1331-1334) and the search-bin
// geometry on the device, makeNNlist (:209-263) with atomic slots and a
// per-voxel sort into ascending order.
// ------------------------------------------------------------------------------
constexpr int kBoxCells = 8;  // cells per workgroup of k_cellbox

__global__ __launch_bounds__(192) void k_cellbox(Pool p, Range r) {
    __shared__ float part[3][6];
    __shared__ int32_t acc[6];
    __shared__ int32_t accsys;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    if (tid == 0) accsys = -1;
    __syncthreads();
    const int first = r.slot_lo + blockIdx.x * kBoxCells;
    for (int q = 0; q < kBoxCells; ++q) {
        const int slot = first + q;
        if (slot >= r.slot_hi) break;
        const int s = sys_of(p, r, slot);
        const SysMeta* M = s >= 0 ? p.meta + s : nullptr;
        const bool use = M && M->rebuild && (slot - M->base) < M->n;
        if (!use) continue;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        const bool node = tid < kNodes;
        float v[6];
        const float x = p.X[gi], y = p.Y[gi], z = p.Z[gi];
        v[0] = node ? x : INFINITY; v[1] = node ? x : -INFINITY;
        v[2] = node ? y : INFINITY; v[3] = node ? y : -INFINITY;
        v[4] = node ? z : INFINITY; v[5] = node ? z : -INFINITY;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int a = 0; a < 6; a += 2) {
                v[a] = fminf(v[a], __shfl_xor(v[a], o));
                v[a + 1] = fmaxf(v[a + 1], __shfl_xor(v[a + 1], o));
            }
        }
        if (lane == 0)
            for (int a = 0; a < 6; ++a) part[wv][a] = v[a];
        __syncthreads();
        if (tid < 6) {
            float b = part[0][tid];
            b = (tid & 1) ? fmaxf(b, part[1][tid]) : fminf(b, part[1][tid]);
            b = (tid & 1) ? fmaxf(b, part[2][tid]) : fminf(b, part[2][tid]);
            p.BBOX[(int64_t)slot * 6 + tid] = b;
            if (accsys != s) {  // flush the previous system's partial extent
                if (accsys >= 0) {
                    SysMeta* A = p.meta + accsys;
                    if (tid & 1) atomicMax(&A->ext[tid], acc[tid]);
                    else atomicMin(&A->ext[tid], acc[tid]);
                }
                acc[tid] = f2o(b);
            } else {
                acc[tid] = (tid & 1) ? max(acc[tid], f2o(b)) : min(acc[tid], f2o(b));
            }
        }
        __syncthreads();
        if (tid == 0) accsys = s;
        __syncthreads();
    }
    if (tid < 6 && accsys >= 0) {
        SysMeta* A = p.meta + accsys;
        if (tid & 1) atomicMax(&A->ext[tid], acc[tid]);
        else atomicMin(&A->ext[tid], acc[tid]);
    }
}

// blockIdx.y = system, blockIdx.x strides over its voxels (clearing counts)
__global__ __launch_bounds__(256) void k_grid_dims(Pool p, Range r) {
    const int s = r.sys_lo + blockIdx.y;
    if (s >= r.sys_hi) return;
    SysMeta* M = p.meta + s;
    if (!M->rebuild) return;
    const cs3d_params* P = p.prm + s;
    float e[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) e[a] = o2f(M->ext[a]);
    const float DL = P->grid_dl;
    const int xd = (int)((e[1] - e[0]) / DL + 1.0f);
    const int yd = (int)((e[3] - e[2]) / DL + 1.0f);
    const int zd = (int)((e[5] - e[4]) / DL + 1.0f);
    const int64_t nv = (int64_t)xd * yd * zd;
    const bool ok = xd > 0 && yd > 0 && zd > 0 && nv <= M->vbudget;
    const int64_t nclear = ok ? nv : (int64_t)M->vbudget;
    int32_t* counts = p.vox + M->vbase;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nclear; v += (int64_t)gridDim.x * blockDim.x)
        counts[v] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) (void)set_grid_geometry(M, P, e);
}

// force = 1: fill the lists of the current grid whatever the rebuild flag says
__global__ void k_insert(Pool p, Range r, int32_t force) {
    for (int slot = r.slot_lo + blockIdx.x * blockDim.x + threadIdx.x; slot < r.slot_hi;
         slot += gridDim.x * blockDim.x) {
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!force && !M->rebuild) continue;
        const int c = slot - M->base;
        if (c == 0) {  // extent accumulators back to the sentinels of minmaxpre (+-1e8)
#pragma unroll
            for (int a = 0; a < 3; ++a) { M->ext[2 * a] = f2o(1.0E8f); M->ext[2 * a + 1] = f2o(-1.0E8f); }
        }
        if (c >= M->n || (M->status & CS3D_SYS_GRID_TOO_LARGE)) continue;
        const cs3d_params* P = p.prm + s;
        const VoxRange vr = vox_range(p.BBOX + (int64_t)slot * 6, M, P);
        {
            const int w[6] = {vr.sx, vr.ex, vr.sy, vr.ey, vr.sz, vr.ez};
            for (int a = 0; a < 6; ++a) p.VRANGE[(int64_t)slot * 6 + a] = w[a];
        }
        const int xd = M->dims[0], yd = M->dims[1];
        int32_t* counts = p.vox + M->vbase;
        int32_t* lists = p.vox + M->lbase;
        const int cap = M->vcap;
        bool over = false;
        for (int j1 = vr.sx; j1 <= vr.ex; ++j1)
            for (int j2 = vr.sy; j2 <= vr.ey; ++j2)
                for (int j3 = vr.sz; j3 <= vr.ez; ++j3) {
                    const int v = j3 * xd * yd + j2 * xd + j1;
                    const int k = atomicAdd(&counts[v], 1);
                    if (k < cap) lists[(int64_t)v * cap + k] = c;
                    else over = true;
                }
        if (over) atomicOr(&M->status, (uint32_t)CS3D_SYS_VOXEL_OVERFLOW);
    }
}

__global__ __launch_bounds__(256) void k_sort(Pool p, Range r, int32_t force) {
    const int s = r.sys_lo + blockIdx.y;
    if (s >= r.sys_hi) return;
    SysMeta* M = p.meta + s;
    if (!force && !M->rebuild) return;
    if (M->status & CS3D_SYS_GRID_TOO_LARGE) return;
    const int64_t nv = (int64_t)M->dims[0] * M->dims[1] * M->dims[2];
    const int cap = M->vcap;
    const int32_t* counts = p.vox + M->vbase;
    int32_t* lists = p.vox + M->lbase;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const int n = min(counts[v], cap);
        int32_t* L = lists + v * cap;
        for (int i = 1; i < n; ++i) {  // insertion sort of one short list
            const int key = L[i];
            int j = i - 1;
            while (j >= 0 && L[j] > key) { L[j + 1] = L[j]; --j; }
            L[j + 1] = key;
        }
    }
}

// Completes a grid whose per-cell part is pending (the fused step leaves it to
// the next step's Integrate): the new bboxes become current.  Then the voxel
// counters of the current grid are cleared so that k_insert can fill the lists.
__global__ __launch_bounds__(192) void k_grid_adopt(Pool p, Range r) {
    for (int slot = r.slot_lo + xcd_block(); slot < r.slot_hi; slot += gridDim.x) {
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        const SysMeta* M = p.meta + s;
        if (!M->rebuild || slot - M->base >= M->n) continue;
        if (threadIdx.x < 6) p.BBOX[(int64_t)slot * 6 + threadIdx.x] = p.BBOXN[(int64_t)slot * 6 + threadIdx.x];
        if (!(M->status & CS3D_SYS_GRID_TOO_LARGE))
            store_vrange(p.VRANGE, slot, vox_range(p.BBOXN + (int64_t)slot * 6, M, p.prm + s), threadIdx.x);
    }
}

// voxel ranges of the current bboxes (after a caller sets BBOX or the grid)
__global__ void k_vrange(Pool p, Range r) {
    for (int slot = r.slot_lo + blockIdx.x * blockDim.x + threadIdx.x; slot < r.slot_hi;
         slot += gridDim.x * blockDim.x) {
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        const SysMeta* M = p.meta + s;
        if (slot - M->base >= M->n || (M->status & CS3D_SYS_GRID_TOO_LARGE)) continue;
        const VoxRange v = vox_range(p.BBOX + (int64_t)slot * 6, M, p.prm + s);
        const int w[6] = {v.sx, v.ex, v.sy, v.ey, v.sz, v.ez};
        for (int a = 0; a < 6; ++a) p.VRANGE[(int64_t)slot * 6 + a] = w[a];
    }
}

__global__ __launch_bounds__(256) void k_vox_clear(Pool p, Range r) {
    const int s = r.sys_lo + blockIdx.y;
    if (s >= r.sys_hi) return;
    const SysMeta* M = p.meta + s;
    if (M->status & CS3D_SYS_GRID_TOO_LARGE) return;
    const int64_t nv = (int64_t)M->dims[0] * M->dims[1] * M->dims[2];
    int32_t* counts = p.vox + M->vbase;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
        counts[v] = 0;
}

// VelocityUpdateA then VelocityUpdateB (propagate.cu:714-736) on V (per-kernel entry point)
// which: 1 = VelocityUpdateA, 2 = VelocityUpdateB, 3 = A then B
__global__ __launch_bounds__(192) void k_velocity_ab(Pool p, Range r, int32_t which) {
    const int tid = threadIdx.x;
    for (int slot = r.slot_lo + xcd_block(); slot < r.slot_hi; slot += gridDim.x) {
        const int s = sys_of(p, r, slot);
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (slot - M->base >= M->n || tid >= kNodes) continue;
        const cs3d_params* P = p.prm + s;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        const float dt = P->dt, m = P->mass, rdt = sqrt_cr(dt);
        float v[3] = {p.VX[gi], p.VY[gi], p.VZ[gi]};
        const float fc[3] = {p.FCX[gi], p.FCY[gi], p.FCZ[gi]};
        const float fd[3] = {p.FDX[gi], p.FDY[gi], p.FDZ[gi]};
        const bool rnd = P->add_rands != 0;  // the stored random force (d_fRanList)
        const float fr[3] = {rnd ? p.FRX[gi] : 0.f, rnd ? p.FRY[gi] : 0.f, rnd ? p.FRZ[gi] : 0.f};
        if (which & 1) {  // VelocityUpdateA (propagate.cu:714-725)
#pragma unroll
            for (int a = 0; a < 3; ++a)
                v[a] = (float)((double)v[a] + kick_a(dt * fc[a] + rdt * fr[a], m, cmeta(M)));
        }
        if (which & 2) {  // VelocityUpdateB (propagate.cu:727-736)
#pragma unroll
            for (int a = 0; a < 3; ++a) v[a] = (float)((double)v[a] + kick_b(fd[a], dt, m, cmeta(M)));
        }
        p.VX[gi] = v[0]; p.VY[gi] = v[1]; p.VZ[gi] = v[2];
    }
}

// minmaxpre + minmaxpost (propagatebound.cu:145-204, 266-323) over the cell boxes
// in BBOX of system r.sys_lo: {minx, maxx, miny, maxy, minz, maxz} from the
// sentinels +-1e8 with fminf / fmaxf, into out[6].  One workgroup.
__global__ __launch_bounds__(256) void k_extent(Pool p, Range r, float* __restrict__ out) {
    __shared__ float red[6][256];
    const int tid = threadIdx.x;
    const SysMeta* M = p.meta + r.sys_lo;
    float e[6] = {1.0E8f, -1.0E8f, 1.0E8f, -1.0E8f, 1.0E8f, -1.0E8f};
    for (int c = tid; c < M->n; c += blockDim.x) {
        const float* b = p.BBOX + (int64_t)(M->base + c) * 6;
#pragma unroll
        for (int a = 0; a < 6; a += 2) {
            e[a] = fminf(e[a], b[a]);
            e[a + 1] = fmaxf(e[a + 1], b[a + 1]);
        }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) red[a][tid] = e[a];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int a = 0; a < 6; a += 2) {
                red[a][tid] = fminf(red[a][tid], red[a][tid + h]);
                red[a + 1][tid] = fmaxf(red[a + 1][tid], red[a + 1][tid + h]);
            }
        }
        __syncthreads();
    }
    if (tid < 6) out[tid] = red[tid][0];
}

// PressureReset (PressureKernels.cu:45-65): pressure of the listed cells (local
// indices of system r.sys_lo, checked on the host) set to value
__global__ __launch_bounds__(256) void k_pressure_reset(Pool p, Range r, const int32_t* __restrict__ cells,
                                                        int32_t count, float value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p.P[p.meta[r.sys_lo].base + cells[i]] = value;
}

// plain streaming copy used to calibrate HBM byte counters
// Arithmetic self-checks of the kernels' own functions (for tests): sqrt_cr against
// sqrtf over all 2^32 inputs (count of differing results, NaNs equal), and
// cs3d_acos of n arguments.
__global__ __launch_bounds__(256) void k_check_sqrt(unsigned long long* bad) {
    unsigned long long nb = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < (1ull << 32);
         u += (uint64_t)gridDim.x * blockDim.x) {
        const float x = __uint_as_float((uint32_t)u);
        const float a = sqrt_cr(x), b = sqrtf(x);
        nb += !(__float_as_uint(a) == __float_as_uint(b) || (isnan(a) && isnan(b)));
    }
    if (nb) atomicAdd(bad, nb);
}
// Each check also counts what shows that it ran and can see a difference: the
// reciprocal check the floats it compared, the quotient checks the pairs where
// the uncorrected product RN(a * y) differs from a / b (about a quarter).
// rcp_rn against 1.0f / b for every float b with quot_ok(b): out[0] differing,
// out[1] compared
__global__ __launch_bounds__(256) void k_check_rcp(unsigned long long* out) {
    unsigned long long nb = 0, nc = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < (1ull << 32);
         u += (uint64_t)gridDim.x * blockDim.x) {
        const float b = __uint_as_float((uint32_t)u);
        if (quot_ok(b)) {
            nb += __float_as_uint(rcp_rn(b)) != __float_as_uint(1.0f / b);
            ++nc;
        }
    }
    if (nb) atomicAdd(out, nb);
    atomicAdd(out + 1, nc);
}
// quot_rn(a, b, rcp_rn(b)) against a / b: b in [1, 2), one significand per thread
// (the grid is exactly 2^23 threads), a in [1, 2) with the a_count significands
// from a_first (a_first + a_count <= 2^23, checked by the caller), and a = +-0
// with the first chunk; out[0] differing, out[1] the control
__global__ __launch_bounds__(256) void k_check_quot(unsigned long long* out, uint32_t a_first, uint32_t a_count) {
    const uint32_t bs = blockIdx.x * blockDim.x + threadIdx.x;
    const float b = __uint_as_float(0x3f800000u | bs);
    const float y = rcp_rn(b);
    uint32_t nb = 0, nc = 0;
    if (a_first == 0) {  // a = +0 and -0 keep their sign
        const float zp = __uint_as_float(0u), zm = __uint_as_float(0x80000000u);
        nb += __float_as_uint(quot_rn(zp, b, y)) != __float_as_uint(zp / b);
        nb += __float_as_uint(quot_rn(zm, b, y)) != __float_as_uint(zm / b);
    }
    for (uint32_t i = 0; i < a_count; ++i) {
        const float a = __uint_as_float(0x3f800000u | (a_first + i));
        const uint32_t t = __float_as_uint(a / b);
        nb += __float_as_uint(quot_rn(a, b, y)) != t;
        nc += __float_as_uint(a * y) != t;
    }
    if (nb) atomicAdd(out, (unsigned long long)nb);
    atomicAdd(out + 1, (unsigned long long)nc);
}
// the same on n pseudo-random pairs over the whole range quot_ok admits: a of
// either sign, both with exponents in [-40, 39] and random significands
__global__ __launch_bounds__(256) void k_check_quot_scaled(unsigned long long* out, uint64_t seed, uint64_t n) {
    unsigned long long nb = 0, nc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t h = (i + seed) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 29; h *= 0x94D049BB133111EBull; h ^= h >> 32;
        const uint32_t ea = (uint32_t)(h % 80), eb = (uint32_t)((h >> 8) % 80);
        const float a = __uint_as_float((uint32_t)((h >> 63) << 31) | ((87u + ea) << 23) | (uint32_t)((h >> 16) & 0x7fffff));
        const float b = __uint_as_float(((87u + eb) << 23) | (uint32_t)((h >> 39) & 0x7fffff));
        const float y = rcp_rn(b);
        const uint32_t t = __float_as_uint(a / b);
        nb += __float_as_uint(quot_rn(a, b, y)) != t;
        nc += __float_as_uint(a * y) != t;
    }
    if (nb) atomicAdd(out, nb);
    atomicAdd(out + 1, nc);
}
__global__ __launch_bounds__(256) void k_acos(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = cs3d_acos(x[i]);
}

__global__ __launch_bounds__(256) void k_copy_calibrate(const float4* __restrict__ src, float4* __restrict__ dst,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// sets the grid-rebuild and candidate-search flags of systems (-1 = unchanged)
__global__ void k_set_flags(Pool p, Range r, int32_t rebuild, int32_t search) {
    const int s = r.sys_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < r.sys_hi && p.meta[s].active) {
        if (rebuild >= 0) p.meta[s].rebuild = rebuild;
        if (search >= 0) {
            p.meta[s].need_search = search;
            p.meta[s].n_reset = 0;
        }
    }
}

// ------------------------------------------------------------------------------
// launchers (called by cs3d_api.hip)
// ------------------------------------------------------------------------------
// Kernel timing for cs3d_step_timed: while kernel_timer is set, every launch
// carries its own start/stop events (hipExtLaunchKernelGGL), so a phase's time is
// the sum of its kernels' durations, without the launch gaps.
thread_local std::vector<std::pair<hipEvent_t, hipEvent_t>>* kernel_timer = nullptr;

template <class F, class... A>
static void launch(F kernel, dim3 grid, dim3 block, hipStream_t st, A... args) {
    if (kernel_timer) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
            hipExtLaunchKernelGGL(kernel, grid, block, 0, st, e0, e1, 0, args...);
            kernel_timer->push_back({e0, e1});
            return;
        }
        if (e0) (void)hipEventDestroy(e0);
    }
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

// Grid of a cell kernel: one workgroup per cell slot up to 256 CUs x wg_per_cu
// workgroups; beyond that each workgroup strides over several cells.
static int grid_cells(const Range& r, int per_cu) {
    const int n = r.slot_hi - r.slot_lo;
    const int cap = 256 * (per_cu > 0 ? per_cu : 16);
    return n < 1 ? 1 : (n < cap ? n : cap);
}

// workgroups of k_live and of the division phase over several systems
constexpr int kSysBlocks = 256;

void launch_live(const Pool& p, const Range& r, hipStream_t st) {
    launch(k_live, dim3(kSysBlocks / 4), dim3(kLiveThreads), st, p, r);
}
void launch_set_nsys(const Pool& p, int n, hipStream_t st) { launch(k_set_nsys, dim3(1), dim3(64), st, p, n); }

void launch_integrate(const Pool& p, const Range& r, int flags, int64_t step_override, hipStream_t st) {
    if (r.live)
        launch(k_integrate<1>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, flags, step_override);
    else
        launch(k_integrate<0>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, flags, step_override);
}
void launch_search(const Pool& p, const Range& r, int always, hipStream_t st) {
    if (r.live)
        launch(k_search<1>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, always);
    else
        launch(k_search<0>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, always);
}
void launch_con(const Pool& p, const Range& r, int mode, bool use_va, hipStream_t st) {
    const float* vx = use_va ? p.VAX : p.VX;
    const float* vy = use_va ? p.VAY : p.VY;
    const float* vz = use_va ? p.VAZ : p.VZ;
    if (r.live)
        launch(k_con_dis<1>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, mode, p.X, p.Y, p.Z, vx, vy, vz,
                               p.BBOX);
    else
        launch(k_con_dis<0>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, mode, p.X, p.Y, p.Z, vx, vy, vz,
                               p.BBOX);
}
// angles: 0 = systems without bending, 1 = with, -1 = both launches
void launch_bonded(const Pool& p, const Range& r, int angles, hipStream_t st) {
    if (angles <= 0) launch(k_bonded<false>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r);
    if (angles != 0) launch(k_bonded<true>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r);
}
void launch_dis(const Pool& p, const Range& r, int mode, int64_t step_override, int clear_bins, bool use_va,
                hipStream_t st) {
    const float* vx = use_va ? p.VAX : p.VX;
    const float* vy = use_va ? p.VAY : p.VY;
    const float* vz = use_va ? p.VAZ : p.VZ;
    if (r.live)
        launch(k_dis_post<1>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, mode, step_override, clear_bins,
                               p.X, p.Y, p.Z, vx, vy, vz);
    else
        launch(k_dis_post<0>, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, mode, step_override, clear_bins,
                               p.X, p.Y, p.Z, vx, vy, vz);
}
void launch_divide(const Pool& p, const Range& r, int mode, hipStream_t st) {
    const int ns = r.sys_hi - r.sys_lo;
    launch(k_sys_divide, dim3(ns < 1 ? 1 : (ns < kSysBlocks ? ns : kSysBlocks)), dim3(256), st, p, r, mode);
}
void launch_com_shift(const Pool& p, const Range& r, hipStream_t st) {
    launch(k_com_shift, dim3(grid_cells(r, p.wg_per_cu)), dim3(kSlots), st, p, r);
}
void launch_grid(const Pool& p, const Range& r, int vox_blocks, hipStream_t st) {
    const int n = r.slot_hi - r.slot_lo;
    const int cb = (n + 255) / 256;
    const int bb = (n + kBoxCells - 1) / kBoxCells;
    const int ns = r.sys_hi - r.sys_lo;
    launch(k_cellbox, dim3(bb > 0 ? bb : 1), dim3(192), st, p, r);
    launch(k_grid_dims, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), st, p, r);
    launch(k_insert, dim3(cb > 0 ? cb : 1), dim3(256), st, p, r, 0);
    launch(k_sort, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), st, p, r, 0);
}
void launch_vrange(const Pool& p, const Range& r, hipStream_t st) {
    const int n = r.slot_hi - r.slot_lo;
    launch(k_vrange, dim3(n > 0 ? (n + 255) / 256 : 1), dim3(256), st, p, r);
}
// voxel lists of the current grid (the fused step keeps only the counts)
void launch_vox_lists(const Pool& p, const Range& r, int vox_blocks, hipStream_t st) {
    const int n = r.slot_hi - r.slot_lo;
    const int cb = (n + 255) / 256;
    const int ns = r.sys_hi - r.sys_lo;
    launch(k_grid_adopt, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r);
    launch(k_vox_clear, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), st, p, r);
    launch(k_insert, dim3(cb > 0 ? cb : 1), dim3(256), st, p, r, 1);
    launch(k_sort, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), st, p, r, 1);
}
void launch_velocity_ab(const Pool& p, const Range& r, int which, hipStream_t st) {
    launch(k_velocity_ab, dim3(grid_cells(r, p.wg_per_cu)), dim3(192), st, p, r, which);
}
void launch_cellbox(const Pool& p, const Range& r, hipStream_t st) {
    const int n = r.slot_hi - r.slot_lo;
    const int bb = (n + kBoxCells - 1) / kBoxCells;
    launch(k_cellbox, dim3(bb > 0 ? bb : 1), dim3(192), st, p, r);
}
void launch_extent(const Pool& p, const Range& r, float* out, hipStream_t st) {
    launch(k_extent, dim3(1), dim3(256), st, p, r, out);
}
void launch_pressure_reset(const Pool& p, const Range& r, const int32_t* cells, int count, float value,
                           hipStream_t st) {
    launch(k_pressure_reset, dim3((count + 255) / 256), dim3(256), st, p, r, cells, count, value);
}
void launch_set_flags(const Pool& p, const Range& r, int rebuild, int search, hipStream_t st) {
    const int ns = r.sys_hi - r.sys_lo;
    launch(k_set_flags, dim3((ns + 255) / 256 > 0 ? (ns + 255) / 256 : 1), dim3(256), st, p, r, rebuild,
                       search);
}
void launch_copy_calibrate(const void* src, void* dst, int64_t n4, hipStream_t st) {
    launch(k_copy_calibrate, dim3(2048), dim3(256), st, (const float4*)src, (float4*)dst, n4);
}
void launch_check_sqrt(unsigned long long* bad, hipStream_t st) {
    launch(k_check_sqrt, dim3(16384), dim3(256), st, bad);
}
void launch_check_div(unsigned long long* bad, uint32_t a_count, uint64_t scaled_pairs, hipStream_t st) {
    launch(k_check_rcp, dim3(16384), dim3(256), st, bad);
    constexpr uint32_t kChunkA = 1024;  // about 5 ms per launch
    for (uint32_t a0 = 0; a0 < a_count; a0 += kChunkA)
        launch(k_check_quot, dim3((1u << 23) / 256), dim3(256), st, bad + 2, a0,
               a_count - a0 < kChunkA ? a_count - a0 : kChunkA);
    if (scaled_pairs > 0) launch(k_check_quot_scaled, dim3(16384), dim3(256), st, bad + 4, (uint64_t)12345, scaled_pairs);
}
void launch_acos(const double* x, double* y, int64_t n, hipStream_t st) {
    launch(k_acos, dim3(n < 1 ? 1 : (n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384), dim3(256), st, x, y, n);
}

}  // namespace cs3d
