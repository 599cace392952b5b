// cs3d_kernels.hip - CDNA4 (gfx950) kernels of the CellSim3D hot path.
//
// One cell per 192-lane workgroup (3 x wave64; lanes 180..191 are the padding
// slots), grid-stride over cell slots so a fixed launch covers a dynamic cell
// count.  Every kernel reproduces the floating-point operation order of the
// reference kernel it replaces (cited per function) so that results match the
// reference's arithmetic; compiled with -ffp-contract=off.
//
// Contact / friction: the reference scans, per node, the cells listed in its
// voxel and tests all 180 nodes of each kept cell (propagate.cu:375-442,
// 603-655).  Here each workgroup first builds the ascending list of cells that
// any of its nodes can keep (union of the voxels under the cell's current box,
// filtered by a box-to-box distance bound that is never larger than a node's),
// then walks it wave-uniformly: neighbour node data are read at uniform
// addresses (scalar loads into SGPRs), each lane applies the reference's own
// per-node test (voxel-list membership via a merge cursor + bbox distance), and
// 16-node patches whose box is farther than attraction_range from the node are
// skipped.  A skipped patch contains no pair the reference would accept (its
// float distance bound is <= every pair distance), and accepted pairs are summed
// in the reference order (ascending cell, ascending node), so the result equals
// the brute-force loop.
#include <hip/hip_runtime.h>
#include <math.h>

#include "cs3d_dev.h"

namespace cs3d {

__device__ __forceinline__ bool sys_live(const SysMeta* m) {
    return m->active && (m->horizon < 0 || m->step < m->horizon) && !(m->status & CS3D_SYS_FATAL_MASK);
}

__device__ __forceinline__ void flag(SysMeta* m, uint32_t bits) {
    if (bits) atomicOr(&m->status, bits);
}

// rGrowth of the population model, GPUbounce.cu:1077-1088
__device__ __forceinline__ float rgrowth(const cs3d_params* P, int n) {
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

// order-preserving int encoding of floats for atomic min/max
__device__ __forceinline__ int f2o(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

// ------------------------------------------------------------------------------
// Integrate + ForwardTime (propagate.cu:681-712, 738-763), PressureUpdate
// (PressureKernels.cu:4-41) and the 16-node patch boxes of the new positions.
// ------------------------------------------------------------------------------
enum : int32_t { I_KICK = 1, I_PRESSURE = 2, I_PBOX = 4 };

__global__ __launch_bounds__(192) void k_integrate(Pool p, Range r, int32_t flags, int64_t step_override) {
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    for (int slot = r.slot_lo + blockIdx.x; slot < r.slot_hi; slot += gridDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!sys_live(M)) continue;
        const int c = slot - M->base;
        if (c >= M->n) continue;
        const cs3d_params* P = p.prm + s;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        float x = p.X[gi], y = p.Y[gi], z = p.Z[gi];
        if (node && (flags & I_KICK)) {
            const float dt = P->dt, m = P->mass;
            const float rdt = sqrtf(dt);
            const float fr = 0.0f;  // random force disabled (add_rands = 0)
            float vx = p.VX[gi], vy = p.VY[gi], vz = p.VZ[gi];
            vx = (float)((double)vx + 0.5 * (double)(dt * p.FCX[gi] + dt * p.FDX[gi] + rdt * fr) / (double)m);
            vy = (float)((double)vy + 0.5 * (double)(dt * p.FCY[gi] + dt * p.FDY[gi] + rdt * fr) / (double)m);
            vz = (float)((double)vz + 0.5 * (double)(dt * p.FCZ[gi] + dt * p.FDZ[gi] + rdt * fr) / (double)m);
            x = x + vx * dt;
            y = y + vy * dt;
            z = z + vz * dt;
            p.VAX[gi] = vx; p.VAY[gi] = vy; p.VAZ[gi] = vz;
            p.X[gi] = x; p.Y[gi] = y; p.Z[gi] = z;
        }
        if (flags & I_PBOX) {
            float mnx = node ? x : INFINITY, mxx = node ? x : -INFINITY;
            float mny = node ? y : INFINITY, mxy = node ? y : -INFINITY;
            float mnz = node ? z : INFINITY, mxz = node ? z : -INFINITY;
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                mnx = fminf(mnx, __shfl_xor(mnx, o, 16)); mxx = fmaxf(mxx, __shfl_xor(mxx, o, 16));
                mny = fminf(mny, __shfl_xor(mny, o, 16)); mxy = fmaxf(mxy, __shfl_xor(mxy, o, 16));
                mnz = fminf(mnz, __shfl_xor(mnz, o, 16)); mxz = fmaxf(mxz, __shfl_xor(mxz, o, 16));
            }
            if ((tid & 15) == 0) {
                float* pb = p.PBOX + (int64_t)slot * kPBox + (tid >> 4) * 6;
                pb[0] = mnx; pb[1] = mxx; pb[2] = mny; pb[3] = mxy; pb[4] = mnz; pb[5] = mxz;
            }
        }
        if (tid == 0 && (flags & I_PRESSURE)) {
            const int64_t stepno = step_override >= 0 ? step_override : M->step + 1;
            const float inc = rgrowth(P, M->n) * p.GS[slot];
            const float mx = P->max_pressure;
            float pr = p.P[slot];
            if (P->use_different_stiffnesses) {
                if (stepno < P->phase_count) {
                    if (pr < mx) pr += inc;
                } else if (p.E[slot] == P->stiffness1) {
                    if (pr < mx) pr += inc;
                } else if (pr != 0.f) {
                    pr = 0.f;
                }
            } else if (pr < mx) {
                pr += inc;
            }
            p.P[slot] = pr;
        }
    }
}

// ------------------------------------------------------------------------------
// Candidate neighbour cells of one cell (workgroup-cooperative).
// ------------------------------------------------------------------------------
struct NCShared {
    int32_t ent[1024];
    int8_t keep[1024];
    int32_t vid[kNVMax];
    int32_t pre[kNVMax];
    int32_t list[kNCMax];
    int32_t ne;
    int32_t nlist;
    int32_t bad;
};

// Builds nc.list (ascending local ids) of cells b != c listed in any voxel under
// [cb0..cb5] whose bbox lies within attraction_range of that box.
__device__ void build_nc(const Pool& p, SysMeta* M, const cs3d_params* P, int c, const float* cb,
                         const float* __restrict__ BB, NCShared& nc) {
    const int tid = threadIdx.x;
    const float DL = P->grid_dl, ar = P->attraction_range;
    const float ar2 = ar * ar;
    const int xd = M->dims[0], yd = M->dims[1], zd = M->dims[2];
    const int vx0 = vox1(cb[0], M->gmin[0], DL, xd), vx1 = vox1(cb[1], M->gmin[0], DL, xd);
    const int vy0 = vox1(cb[2], M->gmin[1], DL, yd), vy1 = vox1(cb[3], M->gmin[1], DL, yd);
    const int vz0 = vox1(cb[4], M->gmin[2], DL, zd), vz1 = vox1(cb[5], M->gmin[2], DL, zd);
    const int nx = vx1 - vx0 + 1, ny = vy1 - vy0 + 1, nz = vz1 - vz0 + 1;
    const int nv = nx * ny * nz;
    const int32_t* counts = p.vox + M->vbase;
    const int32_t* lists = p.vox + M->lbase;
    const int cap = M->vcap;
    if (tid < 64) {
        int cnt = 0, v = 0;
        if (tid < nv && nv <= kNVMax) {
            const int ix = tid % nx, iy = (tid / nx) % ny, iz = tid / (nx * ny);
            v = (vz0 + iz) * xd * yd + (vy0 + iy) * xd + (vx0 + ix);
            cnt = min(counts[v], cap);
        }
        int inc = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(inc, o);
            if (tid >= o) inc += t;
        }
        if (tid < kNVMax) { nc.vid[tid] = v; nc.pre[tid] = inc - cnt; }
        if (tid == 63) {
            nc.ne = inc;
            nc.nlist = 0;
            nc.bad = (nv > kNVMax || inc > 1024) ? 1 : 0;
        }
    }
    __syncthreads();
    if (nc.bad) {
        if (tid == 0) { flag(M, CS3D_SYS_CAND_OVERFLOW); nc.nlist = 0; }
        __syncthreads();
        return;
    }
    const int ne = nc.ne;
    for (int e = tid; e < ne; e += blockDim.x) {
        int t = 0;
        while (t + 1 < nv && nc.pre[t + 1] <= e) ++t;
        nc.ent[e] = lists[(int64_t)nc.vid[t] * cap + (e - nc.pre[t])];
    }
    __syncthreads();
    for (int e = tid; e < ne; e += blockDim.x) {
        const int b = nc.ent[e];
        bool keep = b != c;
        for (int e2 = 0; keep && e2 < e; ++e2) keep = nc.ent[e2] != b;
        if (keep) {  // box-to-box bound of the node test of propagate.cu:382-392
            const float* bb = BB + (int64_t)(M->base + b) * 6;
            float dx = (float)(cb[0] - bb[1] > 0.0f) * (cb[0] - bb[1]);
            dx += (float)(bb[0] - cb[1] > 0.0f) * (bb[0] - cb[1]);
            float dy = (float)(cb[2] - bb[3] > 0.0f) * (cb[2] - bb[3]);
            dy += (float)(bb[2] - cb[3] > 0.0f) * (bb[2] - cb[3]);
            float dz = (float)(cb[4] - bb[5] > 0.0f) * (cb[4] - bb[5]);
            dz += (float)(bb[4] - cb[5] > 0.0f) * (bb[4] - cb[5]);
            keep = !(dx * dx + dy * dy + dz * dz > ar2);
        }
        nc.keep[e] = keep ? 1 : 0;
    }
    __syncthreads();
    for (int e = tid; e < ne; e += blockDim.x) {
        if (!nc.keep[e]) continue;
        const int b = nc.ent[e];
        int rank = 0;
        for (int e2 = 0; e2 < ne; ++e2) rank += (nc.keep[e2] && nc.ent[e2] < b) ? 1 : 0;
        if (rank < kNCMax) nc.list[rank] = b;
        atomicAdd(&nc.nlist, 1);
    }
    __syncthreads();
    if (nc.nlist > kNCMax) {
        __syncthreads();
        if (tid == 0) { flag(M, CS3D_SYS_CAND_OVERFLOW); nc.nlist = kNCMax; }
        __syncthreads();
    }
}

// Per-lane state of the reference's candidate test for one node.
struct LaneCand {
    const int32_t* list;
    int cnt;
    int cur;
    int ncand;
};

__device__ __forceinline__ LaneCand lane_cand(const Pool& p, const SysMeta* M, const cs3d_params* P, float X,
                                              float Y, float Z, bool node) {
    LaneCand L;
    L.cur = 0;
    L.ncand = 0;
    L.cnt = 0;
    L.list = p.vox;
    if (node) {
        const float DL = P->grid_dl;
        const int xd = M->dims[0], yd = M->dims[1], zd = M->dims[2];
        const int sx = vox1(X, M->gmin[0], DL, xd), sy = vox1(Y, M->gmin[1], DL, yd), sz = vox1(Z, M->gmin[2], DL, zd);
        const int index = sz * xd * yd + sy * xd + sx;
        L.cnt = min(p.vox[M->vbase + index], M->vcap);
        L.list = p.vox + M->lbase + (int64_t)index * M->vcap;
    }
    return L;
}

// Reference node test (propagate.cu:375-392): b in the node's voxel list and the
// node within attraction_range of b's bbox.  b values arrive ascending.
__device__ __forceinline__ bool lane_test(LaneCand& L, int b, float X, float Y, float Z, const float* bb,
                                          float ar2) {
    while (L.cur < L.cnt && L.list[L.cur] < b) ++L.cur;
    if (!(L.cur < L.cnt && L.list[L.cur] == b)) return false;
    ++L.cur;
    float dX = (float)(X - bb[1] > 0.0f) * (X - bb[1]);
    dX += (float)(bb[0] - X > 0.0f) * (bb[0] - X);
    float dY = (float)(Y - bb[3] > 0.0f) * (Y - bb[3]);
    dY += (float)(bb[2] - Y > 0.0f) * (bb[2] - Y);
    float dZ = (float)(Z - bb[5] > 0.0f) * (Z - bb[5]);
    dZ += (float)(bb[4] - Z > 0.0f) * (bb[4] - Z);
    if (dX * dX + dY * dY + dZ * dZ > ar2) return false;
    ++L.ncand;
    return true;
}

__device__ __forceinline__ bool patch_near(float X, float Y, float Z, const float* pb, float ar2) {
    float gx = (float)(X - pb[1] > 0.0f) * (X - pb[1]);
    gx += (float)(pb[0] - X > 0.0f) * (pb[0] - X);
    float gy = (float)(Y - pb[3] > 0.0f) * (Y - pb[3]);
    gy += (float)(pb[2] - Y > 0.0f) * (pb[2] - Y);
    float gz = (float)(Z - pb[5] > 0.0f) * (Z - pb[5]);
    gz += (float)(pb[4] - Z > 0.0f) * (pb[4] - Z);
    return !(gx * gx + gy * gy + gz * gz > ar2);
}

struct PairConsts {
    float ar, ar2, rr, ka, kr, gext;
};

// Contact (CON: propagate.cu:406-442) and friction (DIS: propagate.cu:629-655)
// over all kept neighbour nodes, in reference order.
template <bool CON, bool DIS>
__device__ __forceinline__ void pair_loop(const NCShared& nc, int base, const SysMeta* M, LaneCand& L, bool node,
                                          float X, float Y, float Z, float vix, float viy, float viz, float nx,
                                          float ny, float nz, const PairConsts& k,
                                          const float* __restrict__ rX, const float* __restrict__ rY,
                                          const float* __restrict__ rZ, const float* __restrict__ rVX,
                                          const float* __restrict__ rVY, const float* __restrict__ rVZ,
                                          const float* __restrict__ rPB, const float* __restrict__ rBB,
                                          float& cfx, float& cfy, float& cfz, float& fdx, float& fdy,
                                          float& fdz) {
    const int nl = nc.nlist;
    for (int kk = 0; kk < nl; ++kk) {
        const int b = __builtin_amdgcn_readfirstlane(nc.list[kk]);
        const int64_t gb = (int64_t)(base + b);
        const float* bb = rBB + gb * 6;
        const bool cand = node && lane_test(L, b, X, Y, Z, bb, k.ar2);
        if (!__any(cand)) continue;
        const float* pbs = rPB + gb * kPBox;
        const int64_t nb = gb * kSlots;
        for (int pt = 0; pt < kNPatch; ++pt) {
            const bool act = cand && patch_near(X, Y, Z, pbs + pt * 6, k.ar2);
            if (!__any(act)) continue;
            const int j0 = pt * kPatch;
            const int j1 = min(j0 + kPatch, kNodes);
            for (int j = j0; j < j1; ++j) {
                const float xj = rX[nb + j], yj = rY[nb + j], zj = rZ[nb + j];
                float vjx = 0.f, vjy = 0.f, vjz = 0.f;
                if (DIS) { vjx = rVX[nb + j]; vjy = rVY[nb + j]; vjz = rVZ[nb + j]; }
                if (act) {
                    const float dX = X - xj, dY = Y - yj, dZ = Z - zj;
                    const float R2 = dX * dX + dY * dY + dZ * dZ;
                    if (CON && R2 < k.ar2) {
                        const float R = sqrtf(R2);
                        if (R < k.ar) {
                            cfx += k.ka * (k.ar - R) / R * dX;
                            cfy += k.ka * (k.ar - R) / R * dY;
                            cfz += k.ka * (k.ar - R) / R * dZ;
                        }
                        if (R <= k.rr) {
                            cfx += k.kr * (k.rr - R) / R * dX;
                            cfy += k.kr * (k.rr - R) / R * dY;
                            cfz += k.kr * (k.rr - R) / R * dZ;
                        }
                    }
                    if (DIS && !(R2 > k.ar2)) {
                        const float vx = vix - vjx, vy = viy - vjy, vz = viz - vjz;
                        const float d = vx * nx + vy * ny + vz * nz;
                        const float tx = vx - d * nx, ty = vy - d * ny, tz = vz - d * nz;
                        fdx = fdx - k.gext * tx;
                        fdy = fdy - k.gext * ty;
                        fdz = fdz - k.gext * tz;
                    }
                }
            }
        }
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
    const float c1 = -1.0f / sqrtf(inner);
    const float den = sqrtf(ra2) * sqrtf(rb2);
    const float theta = (float)acos((double)idk / ((double)den + 1e-3));
    const float s = ((-k) * c1) * (theta - th0);
    const float ca = idk / ra2, cb = idk / rb2;
    out[0] = s * (bx - ca * ax); out[1] = s * (by - ca * ay); out[2] = s * (bz - ca * az);
    out[3] = s * (ax - cb * bx); out[4] = s * (ay - cb * by); out[5] = s * (az - cb * bz);
    bool ok = isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2]) && isfinite(out[3]) && isfinite(out[4]) &&
              isfinite(out[5]);
    return ok ? 0u : (uint32_t)CS3D_SYS_ANGLE_NONFINITE;
}

// own box of a cell from its 12 patch boxes
__device__ __forceinline__ void cell_box_from_patches(const float* __restrict__ pb, float* cb) {
    const int tid = threadIdx.x;
    if (tid < 6) {
        float v = pb[tid];
        for (int q = 1; q < kNPatch; ++q) {
            const float w = pb[q * 6 + tid];
            v = (tid & 1) ? fmaxf(v, w) : fminf(v, w);
        }
        cb[tid] = v;
    }
}

// ------------------------------------------------------------------------------
// Conservative force (CalculateConForce, propagate.cu:183-520) fused with the
// first dissipative force (CalculateDisForce with the post-Integrate
// velocities, propagate.cu:523-665) and VelocityUpdateA/B (propagate.cu:714-736).
// Modes: M_STEP (FC, V), M_CON / M_INIT (FC only), M_BONDED (spring + pressure
// + angle -> FC).
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(192) void k_con_dis(Pool p, Range r, int32_t mode, const float* __restrict__ rX,
                                                 const float* __restrict__ rY, const float* __restrict__ rZ,
                                                 const float* __restrict__ rVX, const float* __restrict__ rVY,
                                                 const float* __restrict__ rVZ, const float* __restrict__ rPB,
                                                 const float* __restrict__ rBB) {
    __shared__ float sx[kSlots], sy[kSlots], sz[kSlots];
    __shared__ float svx[kSlots], svy[kSlots], svz[kSlots];
    __shared__ float angF[kNodes * 18];
    __shared__ float cbox[6];
    __shared__ NCShared nc;
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    const MeshDev* mesh = p.mesh;
    for (int slot = r.slot_lo + blockIdx.x; slot < r.slot_hi; slot += gridDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!sys_live(M)) continue;
        const int c = slot - M->base;
        if (c >= M->n) continue;
        const cs3d_params* P = p.prm + s;
        const int64_t cb0 = (int64_t)slot * kSlots;
        const int64_t gi = cb0 + tid;
        const float X = rX[gi], Y = rY[gi], Z = rZ[gi];
        sx[tid] = X; sy[tid] = Y; sz[tid] = Z;
        const bool fused = mode == M_STEP;
        if (fused) { svx[tid] = rVX[gi]; svy[tid] = rVY[gi]; svz[tid] = rVZ[gi]; }
        if (mode != M_BONDED) cell_box_from_patches(rPB + (int64_t)slot * kPBox, cbox);
        __syncthreads();
        uint32_t st = 0;
        if (node && (isnan(X) || isnan(Y) || isnan(Z))) st |= CS3D_SYS_NAN;  // propagate.cu:241-248
        // ---- springs, propagate.cu:298-314
        const float stiffness = p.E[slot];
        float FX = 0.f, FY = 0.f, FZ = 0.f;
        int nn3[3] = {0, 0, 0};
        if (node) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int N1 = mesh->nn[i * kSlots + tid];
                nn3[i] = N1;
                const float dX = sx[N1] - X, dY = sy[N1] - Y, dZ = sz[N1] - Z;
                const float R = sqrtf(dX * dX + dY * dY + dZ * dZ);
                const float R0 = mesh->R0[i * kSlots + tid];
                FX += stiffness * (R - R0) / R0 * dX / R;
                FY += stiffness * (R - R0) / R0 * dY / R;
                FZ += stiffness * (R - R0) / R0 * dZ / R;
            }
            // ---- pressure, propagate.cu:318-330 (centre of mass from the last CenterOfMass)
            const float rx = X - p.CMX[slot], ry = Y - p.CMY[slot], rz = Z - p.CMZ[slot];
            const float inv = 1.0f / sqrtf(rx * rx + ry * ry + rz * rz);
            const float g = 3.0f * p.P[slot];
            FX += g * (inv * rx);
            FY += g * (inv * ry);
            FZ += g * (inv * rz);
        }
        // ---- angles, propagate.cu:81-181 (each angle once, gathered in reference order)
        if (P->constrain_angles) {
            const float k = P->angle_k;
            if (node) {
                const float* th = mesh->theta0 + 3 * tid;
                const int A[3] = {nn3[0], nn3[1], nn3[0]};
                const int B[3] = {nn3[1], nn3[2], nn3[2]};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    st |= angle_pair(sx[A[q]] - X, sy[A[q]] - Y, sz[A[q]] - Z, sx[B[q]] - X, sy[B[q]] - Y,
                                     sz[B[q]] - Z, th[q], k, angF + (tid * 3 + q) * 6);
                }
            }
            __syncthreads();
            if (node) {
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
        }
        float fdx = 0.f, fdy = 0.f, fdz = 0.f;
        float vix = 0.f, viy = 0.f, viz = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
        if (mode != M_BONDED) {
            build_nc(p, M, P, c, cbox, rBB, nc);
            if (tid < kNCMax) p.NCL[(int64_t)slot * kNCMax + tid] = tid < nc.nlist ? nc.list[tid] : -1;
            if (tid == 0) p.NCN[slot] = nc.nlist;
            PairConsts k;
            k.ar = P->attraction_range;
            k.ar2 = k.ar * k.ar;
            k.rr = P->repulsion_range;
            k.ka = -P->attraction_strength * P->youngs_mod;
            k.kr = +P->repulsion_strength * P->youngs_mod;
            k.gext = P->viscotic_damping;
            if (fused && node) {
                vix = svx[tid]; viy = svy[tid]; viz = svz[tid];
                const float gint = P->internal_damping;
#pragma unroll
                for (int i = 0; i < 3; ++i) {  // bond damping, propagate.cu:547-554
                    const int N = nn3[i];
                    fdx = fdx - gint * (vix - svx[N]);
                    fdy = fdy - gint * (viy - svy[N]);
                    fdz = fdz - gint * (viz - svz[N]);
                }
                const int N1 = nn3[0], N2 = nn3[1], N3 = nn3[2];  // normal, propagate.cu:558-572
                const float Ax = sx[N2] - sx[N1], Ay = sy[N2] - sy[N1], Az = sz[N2] - sz[N1];
                const float Bx = sx[N3] - sx[N1], By = sy[N3] - sy[N1], Bz = sz[N3] - sz[N1];
                const float cx = Ay * Bz - Az * By, cy = Az * Bx - Ax * Bz, cz = Ax * By - Ay * Bx;
                const float inv = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
                const float sg = (float)mesh->sign[tid];
                nx = sg * (inv * cx); ny = sg * (inv * cy); nz = sg * (inv * cz);
            }
            LaneCand L = lane_cand(p, M, P, X, Y, Z, node);
            float cfx = 0.f, cfy = 0.f, cfz = 0.f;
            if (fused)
                pair_loop<true, true>(nc, M->base, M, L, node, X, Y, Z, vix, viy, viz, nx, ny, nz, k, rX, rY, rZ,
                                      rVX, rVY, rVZ, rPB, rBB, cfx, cfy, cfz, fdx, fdy, fdz);
            else
                pair_loop<true, false>(nc, M->base, M, L, node, X, Y, Z, vix, viy, viz, nx, ny, nz, k, rX, rY, rZ,
                                       rVX, rVY, rVZ, rPB, rBB, cfx, cfy, cfz, fdx, fdy, fdz);
            if (L.ncand > 10) st |= CS3D_SYS_NN_OVER_10;
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
        if (node) {
            p.FCX[gi] = FX; p.FCY[gi] = FY; p.FCZ[gi] = FZ;
            if (fused) {
                fdx = fdx - P->gamma_visc * vix;  // drag, propagate.cu:658
                fdy = fdy - P->gamma_visc * viy;
                fdz = fdz - P->gamma_visc * viz;
                const float dt = P->dt, m = P->mass;
                const float rdt = sqrtf(dt);
                const float fr = 0.0f;
                // VelocityUpdateA (propagate.cu:721-723) then VelocityUpdateB (:732-734)
                float vx = (float)((double)vix + 0.5 * (double)(dt * FX + rdt * fr) / (double)m);
                float vy = (float)((double)viy + 0.5 * (double)(dt * FY + rdt * fr) / (double)m);
                float vz = (float)((double)viz + 0.5 * (double)(dt * FZ + rdt * fr) / (double)m);
                vx = (float)((double)vx + 0.5 * (double)dt * (double)fdx / (double)m);
                vy = (float)((double)vy + 0.5 * (double)dt * (double)fdy / (double)m);
                vz = (float)((double)vz + 0.5 * (double)dt * (double)fdz / (double)m);
                p.VX[gi] = vx; p.VY[gi] = vy; p.VZ[gi] = vz;
            }
        }
        if (st) flag(M, st);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------
// Dissipative force (CalculateDisForce, propagate.cu:523-665) with the
// end-of-step velocities, then CenterOfMass (centermass.cu:8-50) and, on growth
// steps, volumes (volume.cu:5-163).  Modes: M_STEP, M_DIS_V, M_DIS_VA, M_COM,
// M_VOLUMES.
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(192) void k_dis_post(Pool p, Range r, int32_t mode, int64_t step_override,
                                                  const float* __restrict__ rX, const float* __restrict__ rY,
                                                  const float* __restrict__ rZ, const float* __restrict__ rVX,
                                                  const float* __restrict__ rVY, const float* __restrict__ rVZ,
                                                  const float* __restrict__ rPB, const float* __restrict__ rBB) {
    __shared__ float sx[kSlots], sy[kSlots], sz[kSlots];
    __shared__ float svx[kSlots], svy[kSlots], svz[kSlots];
    __shared__ float red[3][128];
    __shared__ float cm[3];
    __shared__ float cbox[6];
    __shared__ NCShared nc;
    const int tid = threadIdx.x;
    const bool node = tid < kNodes;
    const MeshDev* mesh = p.mesh;
    const bool do_dis = mode == M_STEP || mode == M_DIS_V || mode == M_DIS_VA;
    for (int slot = r.slot_lo + blockIdx.x; slot < r.slot_hi; slot += gridDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!sys_live(M)) continue;
        const int c = slot - M->base;
        if (c >= M->n) continue;
        const cs3d_params* P = p.prm + s;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        const float X = rX[gi], Y = rY[gi], Z = rZ[gi];
        sx[tid] = X; sy[tid] = Y; sz[tid] = Z;
        if (do_dis) { svx[tid] = rVX[gi]; svy[tid] = rVY[gi]; svz[tid] = rVZ[gi]; }
        if (do_dis && mode != M_STEP) cell_box_from_patches(rPB + (int64_t)slot * kPBox, cbox);
        if (do_dis && mode == M_STEP) {
            if (tid == 0) nc.nlist = p.NCN[slot];
            if (tid < kNCMax) nc.list[tid] = p.NCL[(int64_t)slot * kNCMax + tid];
        }
        __syncthreads();
        uint32_t st = 0;
        if (do_dis) {
            if (mode != M_STEP) build_nc(p, M, P, c, cbox, rBB, nc);
            float fx = 0.f, fy = 0.f, fz = 0.f;
            float vix = 0.f, viy = 0.f, viz = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
            int nn3[3] = {0, 0, 0};
            if (node) {
                vix = svx[tid]; viy = svy[tid]; viz = svz[tid];
                const float gint = P->internal_damping;
#pragma unroll
                for (int i = 0; i < 3; ++i) {  // bond damping, propagate.cu:547-554
                    const int N = mesh->nn[i * kSlots + tid];
                    nn3[i] = N;
                    fx = fx - gint * (vix - svx[N]);
                    fy = fy - gint * (viy - svy[N]);
                    fz = fz - gint * (viz - svz[N]);
                }
                const int N1 = nn3[0], N2 = nn3[1], N3 = nn3[2];  // normal, propagate.cu:558-572
                const float Ax = sx[N2] - sx[N1], Ay = sy[N2] - sy[N1], Az = sz[N2] - sz[N1];
                const float Bx = sx[N3] - sx[N1], By = sy[N3] - sy[N1], Bz = sz[N3] - sz[N1];
                const float cx = Ay * Bz - Az * By, cy = Az * Bx - Ax * Bz, cz = Ax * By - Ay * Bx;
                const float inv = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
                const float sg = (float)mesh->sign[tid];
                nx = sg * (inv * cx); ny = sg * (inv * cy); nz = sg * (inv * cz);
            }
            PairConsts k;
            k.ar = P->attraction_range;
            k.ar2 = k.ar * k.ar;
            k.rr = P->repulsion_range;
            k.ka = 0.f;
            k.kr = 0.f;
            k.gext = P->viscotic_damping;
            LaneCand L = lane_cand(p, M, P, X, Y, Z, node);
            float cfx = 0.f, cfy = 0.f, cfz = 0.f;
            pair_loop<false, true>(nc, M->base, M, L, node, X, Y, Z, vix, viy, viz, nx, ny, nz, k, rX, rY, rZ, rVX,
                                   rVY, rVZ, rPB, rBB, cfx, cfy, cfz, fx, fy, fz);
            if (node) {
                fx = fx - P->gamma_visc * vix;  // drag, propagate.cu:658
                fy = fy - P->gamma_visc * viy;
                fz = fz - P->gamma_visc * viz;
                p.FDX[gi] = fx; p.FDY[gi] = fy; p.FDZ[gi] = fz;
            }
            if (L.ncand > 10) st |= CS3D_SYS_NN_OVER_10;
        }
        const bool do_com = mode == M_STEP || mode == M_COM;
        const int64_t stepno = step_override >= 0 ? step_override : M->step + 1;
        const bool do_vol =
            mode == M_VOLUMES || (mode == M_STEP && stepno <= P->div_time_steps && rgrowth(P, M->n) > 0.0f);
        if (do_com) {
            // 256-slot tree of centermass.cu:32-41: level 128 and 64 through LDS, 32..1 by shuffles
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
                    p.CMX[slot] = cm[0]; p.CMY[slot] = cm[1]; p.CMZ[slot] = cm[2];
                }
            }
            __syncthreads();
        } else if (do_vol) {
            if (tid == 0) { cm[0] = p.CMX[slot]; cm[1] = p.CMY[slot]; cm[2] = p.CMZ[slot]; }
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
                        faceArea = (float)((double)faceArea + 0.5 * (double)sqrtf(xx * xx + yy * yy + zz * zz));
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
                    p.VOL[slot] = volume;
                    if (!isfinite(volume)) st |= CS3D_SYS_VOLUME_NONFINITE;
                    int8_t fl = volume > P->division_vol ? 1 : 0;
                    if (sph) {
                        p.AREA[slot] = b;
                        const float cbr = cbrtf(volume);
                        const float psi = (float)(4.835975862049408 * (double)cbr * (double)cbr / (double)b);
                        if ((double)fabsf(1.0f - psi) > 0.05) fl = 0;
                    }
                    if (P->use_different_stiffnesses && P->recalc_r0 && p.E[slot] != P->stiffness1) fl = 0;
                    p.FLAG[slot] = fl;
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
    n[0] = 2.0f * x1 * sqrtf(1.0f - x1 * x1 - x2 * x2);
    n[1] = 2.0f * x2 * sqrtf(1.0f - x1 * x1 - x2 * x2);
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
    f = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] / f; v[1] = v[1] / f; v[2] = v[2] / f;
    f = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
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
__global__ __launch_bounds__(256) void k_sys_divide(Pool p, Range r, int32_t mode) {
    __shared__ int32_t scan[256];
    __shared__ int32_t nev, ncur, stop;
    __shared__ float nrm[3];
    __shared__ int32_t bad;
    __shared__ float redc[3][128];
    const int tid = threadIdx.x;
    const int s = r.sys_lo + blockIdx.x;
    if (s >= r.sys_hi) return;
    SysMeta* M = p.meta + s;
    const bool step_mode = mode == M_STEP;
    if (!M->active) {
        if (tid == 0 && step_mode) M->rebuild = 0;
        return;
    }
    if (step_mode && !sys_live(M)) {
        if (tid == 0) M->rebuild = 0;
        return;
    }
    const cs3d_params* P = p.prm + s;
    const int base = M->base;
    const int n0 = M->n;
    const int64_t stepno = M->step + 1;
    bool growth = true;
    if (step_mode) growth = stepno <= P->div_time_steps && rgrowth(P, n0) > 0.0f;
    if (tid == 0) { nev = 0; ncur = n0; stop = 0; bad = 0; }
    __syncthreads();
    if (growth) {
        // ascending event list in EVL[base .. base + nev)
        for (int c0 = 0; c0 < n0; c0 += 256) {
            const int c = c0 + tid;
            const int f = (c < n0 && p.FLAG[base + c] == 1) ? 1 : 0;
            scan[tid] = f;
            __syncthreads();
            for (int o = 1; o < 256; o <<= 1) {
                const int t = tid >= o ? scan[tid - o] : 0;
                __syncthreads();
                scan[tid] += t;
                __syncthreads();
            }
            if (f) {
                p.EVL[base + nev + scan[tid] - 1] = c;
                p.FLAG[base + c] = 0;
            }
            __syncthreads();
            if (tid == 0) nev += scan[255];
            __syncthreads();
        }
        const int ne = nev;
        const float half = (float)(0.5 * (double)P->repulsion_range);
        for (int e = 0; e < ne; ++e) {
            if (tid == 0) {
                if (P->use_div_plane_basis) random_vector_basis(&M->rng, nrm, P->div_plane_basis);
                else random_vector(&M->rng, nrm);
                const float nn = sqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                if ((double)fabsf(nn - 1) > 1e-3) bad |= CS3D_SYS_BAD_NORMAL;
                if (ncur >= M->cap) { bad |= CS3D_SYS_CAPACITY; stop = 1; }
            }
            __syncthreads();
            if (stop) break;
            const int rank = p.EVL[base + e];
            const int d = ncur;
            const int64_t pi = (int64_t)(base + rank) * kSlots + tid;
            const int64_t di = (int64_t)(base + d) * kSlots + tid;
            const float cmx = p.CMX[base + rank], cmy = p.CMY[base + rank], cmz = p.CMZ[base + rank];
            float px = 0, py = 0, pz = 0, dx = 0, dy = 0, dz = 0;
            const bool nd = tid < kNodes;
            if (nd) {
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
                    if (side) { dx = x; dy = y; dz = z; } else { px = x; py = y; pz = z; }
                }
                p.X[pi] = px; p.Y[pi] = py; p.Z[pi] = pz;
                p.X[di] = dx; p.Y[di] = dy; p.Z[di] = dz;
                p.VX[di] = p.VX[pi]; p.VY[di] = p.VY[pi]; p.VZ[di] = p.VZ[pi];
            } else if (tid < kSlots) {
                p.X[di] = 0.f; p.Y[di] = 0.f; p.Z[di] = 0.f;
                p.VX[di] = 0.f; p.VY[di] = 0.f; p.VZ[di] = 0.f;
            }
            if (tid < kSlots) {  // a new cell's forces start at zero (never-written slots, GPUbounce.cu:563-572)
                p.FCX[di] = 0.f; p.FCY[di] = 0.f; p.FCZ[di] = 0.f;
                p.FDX[di] = 0.f; p.FDY[di] = 0.f; p.FDZ[di] = 0.f;
            }
            // patch boxes of the parent and the daughter
            if (tid < kSlots) {
                for (int side = 0; side < 2; ++side) {
                    const float x = side ? dx : px, y = side ? dy : py, z = side ? dz : pz;
                    float mnx = nd ? x : INFINITY, mxx = nd ? x : -INFINITY;
                    float mny = nd ? y : INFINITY, mxy = nd ? y : -INFINITY;
                    float mnz = nd ? z : INFINITY, mxz = nd ? z : -INFINITY;
#pragma unroll
                    for (int o = 8; o > 0; o >>= 1) {
                        mnx = fminf(mnx, __shfl_xor(mnx, o, 16)); mxx = fmaxf(mxx, __shfl_xor(mxx, o, 16));
                        mny = fminf(mny, __shfl_xor(mny, o, 16)); mxy = fmaxf(mxy, __shfl_xor(mxy, o, 16));
                        mnz = fminf(mnz, __shfl_xor(mnz, o, 16)); mxz = fmaxf(mxz, __shfl_xor(mxz, o, 16));
                    }
                    if ((tid & 15) == 0) {
                        float* pb = p.PBOX + (int64_t)(base + (side ? d : rank)) * kPBox + (tid >> 4) * 6;
                        pb[0] = mnx; pb[1] = mxx; pb[2] = mny; pb[3] = mxy; pb[4] = mnz; pb[5] = mxz;
                    }
                }
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
    if (tid == 0) {
        if (bad) atomicOr(&M->status, (uint32_t)bad);
        const int events = ncur - n0;
        M->n_events = events;
        M->n = ncur;
        if (step_mode) {
            const int gi = P->grid_interval > 0 ? P->grid_interval : 1;
            M->rebuild = (events > 0 || stepno % gi == 0) ? 1 : 0;
            M->step = stepno;
        }
    }
}

// ------------------------------------------------------------------------------
// Grid rebuild: bounding_boxes (propagatebound.cu:10-141, from the patch
// boxes), minmaxpre/minmaxpost (:145-204, 266-323) as order-free atomic min/max,
// host Xdiv/Ydiv/Zdiv (GPUbounce.cu:1331-1334) on the device, makeNNlist
// (:209-263) with atomic slots and a per-voxel sort into ascending order.
// ------------------------------------------------------------------------------
__global__ void k_cellbox(Pool p, Range r) {
    for (int slot = r.slot_lo + blockIdx.x * blockDim.x + threadIdx.x; slot < r.slot_hi;
         slot += gridDim.x * blockDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!M->rebuild) continue;
        const int c = slot - M->base;
        if (c >= M->n) continue;
        const float* pb = p.PBOX + (int64_t)slot * kPBox;
        float b[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) b[a] = pb[a];
        for (int q = 1; q < kNPatch; ++q) {
#pragma unroll
            for (int a = 0; a < 6; a += 2) {
                b[a] = fminf(b[a], pb[q * 6 + a]);
                b[a + 1] = fmaxf(b[a + 1], pb[q * 6 + a + 1]);
            }
        }
        float* o = p.BBOX + (int64_t)slot * 6;
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = b[a];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&M->ext[2 * a], f2o(b[2 * a]));
            atomicMax(&M->ext[2 * a + 1], f2o(b[2 * a + 1]));
        }
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
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        M->gmin[0] = e[0]; M->gmin[1] = e[2]; M->gmin[2] = e[4];
        M->dims[0] = ok ? xd : 1; M->dims[1] = ok ? yd : 1; M->dims[2] = ok ? zd : 1;
        if (!ok) atomicOr(&M->status, (uint32_t)CS3D_SYS_GRID_TOO_LARGE);
    }
}

__global__ void k_insert(Pool p, Range r) {
    for (int slot = r.slot_lo + blockIdx.x * blockDim.x + threadIdx.x; slot < r.slot_hi;
         slot += gridDim.x * blockDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (!M->rebuild) continue;
        const int c = slot - M->base;
        if (c == 0) {  // extent accumulators back to the sentinels of minmaxpre (+-1e8)
#pragma unroll
            for (int a = 0; a < 3; ++a) { M->ext[2 * a] = f2o(1.0E8f); M->ext[2 * a + 1] = f2o(-1.0E8f); }
        }
        if (c >= M->n || (M->status & CS3D_SYS_GRID_TOO_LARGE)) continue;
        const cs3d_params* P = p.prm + s;
        const float* b = p.BBOX + (int64_t)slot * 6;
        const float DL = P->grid_dl, at = P->attraction_range;
        const float mx = M->gmin[0], my = M->gmin[1], mz = M->gmin[2];
        const int xd = M->dims[0], yd = M->dims[1], zd = M->dims[2];
        int sx = (int)((b[0] - at - mx) / DL); if (sx < 0) sx = 0;
        int ex = (int)((b[1] + at - mx) / DL); if (ex >= xd) ex = xd - 1;
        int sy = (int)((b[2] - at - my) / DL); if (sy < 0) sy = 0;
        int ey = (int)((b[3] + at - my) / DL); if (ey >= yd) ey = yd - 1;
        int sz = (int)((b[4] - at - mz) / DL); if (sz < 0) sz = 0;
        int ez = (int)((b[5] + at - mz) / DL); if (ez >= zd) ez = zd - 1;
        int32_t* counts = p.vox + M->vbase;
        int32_t* lists = p.vox + M->lbase;
        const int cap = M->vcap;
        bool over = false;
        for (int j1 = sx; j1 <= ex; ++j1)
            for (int j2 = sy; j2 <= ey; ++j2)
                for (int j3 = sz; j3 <= ez; ++j3) {
                    const int v = j3 * xd * yd + j2 * xd + j1;
                    const int k = atomicAdd(&counts[v], 1);
                    if (k < cap) lists[(int64_t)v * cap + k] = c;
                    else over = true;
                }
        if (over) atomicOr(&M->status, (uint32_t)CS3D_SYS_VOXEL_OVERFLOW);
    }
}

__global__ __launch_bounds__(256) void k_sort(Pool p, Range r) {
    const int s = r.sys_lo + blockIdx.y;
    if (s >= r.sys_hi) return;
    SysMeta* M = p.meta + s;
    if (!M->rebuild) return;
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

// VelocityUpdateA then VelocityUpdateB (propagate.cu:714-736) on V (per-kernel entry point)
__global__ __launch_bounds__(192) void k_velocity_ab(Pool p, Range r) {
    const int tid = threadIdx.x;
    for (int slot = r.slot_lo + blockIdx.x; slot < r.slot_hi; slot += gridDim.x) {
        const int s = p.slot_sys[slot];
        if (s < 0) continue;
        SysMeta* M = p.meta + s;
        if (slot - M->base >= M->n || tid >= kNodes) continue;
        const cs3d_params* P = p.prm + s;
        const int64_t gi = (int64_t)slot * kSlots + tid;
        const float dt = P->dt, m = P->mass, rdt = sqrtf(dt), fr = 0.0f;
        float v[3] = {p.VX[gi], p.VY[gi], p.VZ[gi]};
        const float fc[3] = {p.FCX[gi], p.FCY[gi], p.FCZ[gi]};
        const float fd[3] = {p.FDX[gi], p.FDY[gi], p.FDZ[gi]};
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = (float)((double)v[a] + 0.5 * (double)(dt * fc[a] + rdt * fr) / (double)m);
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = (float)((double)v[a] + 0.5 * (double)dt * (double)fd[a] / (double)m);
        p.VX[gi] = v[0]; p.VY[gi] = v[1]; p.VZ[gi] = v[2];
    }
}

// marks systems for a grid rebuild (per-kernel entry point and start-up)
__global__ void k_set_rebuild(Pool p, Range r, int32_t value) {
    const int s = r.sys_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < r.sys_hi && p.meta[s].active) p.meta[s].rebuild = value;
}

// ------------------------------------------------------------------------------
// launchers (called by cs3d_api.hip)
// ------------------------------------------------------------------------------
static int grid_cells(const Range& r, int per_cu) {
    const int n = r.slot_hi - r.slot_lo;
    const int cap = 256 * per_cu;
    return n < 1 ? 1 : (n < cap ? n : cap);
}

void launch_integrate(const Pool& p, const Range& r, int flags, int64_t step_override, hipStream_t st) {
    hipLaunchKernelGGL(k_integrate, dim3(grid_cells(r, 16)), dim3(192), 0, st, p, r, flags, step_override);
}
void launch_con(const Pool& p, const Range& r, int mode, bool use_va, hipStream_t st) {
    const float* vx = use_va ? p.VAX : p.VX;
    const float* vy = use_va ? p.VAY : p.VY;
    const float* vz = use_va ? p.VAZ : p.VZ;
    hipLaunchKernelGGL(k_con_dis, dim3(grid_cells(r, 8)), dim3(192), 0, st, p, r, mode, p.X, p.Y, p.Z, vx, vy, vz,
                       p.PBOX, p.BBOX);
}
void launch_dis(const Pool& p, const Range& r, int mode, int64_t step_override, bool use_va, hipStream_t st) {
    const float* vx = use_va ? p.VAX : p.VX;
    const float* vy = use_va ? p.VAY : p.VY;
    const float* vz = use_va ? p.VAZ : p.VZ;
    hipLaunchKernelGGL(k_dis_post, dim3(grid_cells(r, 8)), dim3(192), 0, st, p, r, mode, step_override, p.X, p.Y,
                       p.Z, vx, vy, vz, p.PBOX, p.BBOX);
}
void launch_divide(const Pool& p, const Range& r, int mode, hipStream_t st) {
    const int ns = r.sys_hi - r.sys_lo;
    hipLaunchKernelGGL(k_sys_divide, dim3(ns > 0 ? ns : 1), dim3(256), 0, st, p, r, mode);
}
void launch_grid(const Pool& p, const Range& r, int vox_blocks, hipStream_t st) {
    const int n = r.slot_hi - r.slot_lo;
    const int cb = (n + 255) / 256;
    const int ns = r.sys_hi - r.sys_lo;
    hipLaunchKernelGGL(k_cellbox, dim3(cb > 0 ? cb : 1), dim3(256), 0, st, p, r);
    hipLaunchKernelGGL(k_grid_dims, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), 0, st, p, r);
    hipLaunchKernelGGL(k_insert, dim3(cb > 0 ? cb : 1), dim3(256), 0, st, p, r);
    hipLaunchKernelGGL(k_sort, dim3(vox_blocks, ns > 0 ? ns : 1), dim3(256), 0, st, p, r);
}
void launch_velocity_ab(const Pool& p, const Range& r, hipStream_t st) {
    hipLaunchKernelGGL(k_velocity_ab, dim3(grid_cells(r, 8)), dim3(192), 0, st, p, r);
}
void launch_set_rebuild(const Pool& p, const Range& r, int value, hipStream_t st) {
    const int ns = r.sys_hi - r.sys_lo;
    hipLaunchKernelGGL(k_set_rebuild, dim3((ns + 255) / 256 > 0 ? (ns + 255) / 256 : 1), dim3(256), 0, st, p, r,
                       value);
}

}  // namespace cs3d
