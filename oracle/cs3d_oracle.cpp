// cs3d_oracle.cpp - CPU oracle (test infrastructure; see cs3d_oracle.h).
//
// A restatement, written for this project, of the reference kernel bodies of
// SoftSimu/CellSim3D in their operation order.  Every function cites the
// reference code it follows (paths relative to the reference root).  Built
// with -ffp-contract=off: every float operation rounds on its own.  That is this
// project's canonical order, shared with the device kernels, not a property of
// the reference's binary: nvcc contracts a*b+c into FMA by default
// (makefile:1-13 does not pass --fmad=false) wherever it chooses, so the
// reference's own last bits depend on those choices; the FP64 re-run bounds both.
#include "cs3d_oracle.h"

#include "../cellsim3d_amd/csrc/cs3d_acos.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int S = CS3D_SLOTS;   // 192-slot stride
constexpr int NN = CS3D_NODES;  // 180 real nodes

// ---- float3 semantics of inc/VectorFunctions.hpp:31-122 ----------------------
template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(T a, V3<T> b) { return mk(a * b.x, a * b.y, a * b.z); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T mag2(V3<T> a) { return dot(a, a); }
template <class T> inline T mag(V3<T> a) { return std::sqrt(mag2(a)); }
// operator/(b, a) is (1/a)*b (VectorFunctions.hpp:84-92); calcUnitVec is a/mag(a) (:107-109)
template <class T> inline V3<T> unit(V3<T> a) { return (T(1) / mag(a)) * a; }
template <class T> inline bool finite3(V3<T> a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
template <class T> constexpr bool is_f32() { return std::is_same<T, float>::value; }

template <class T> struct Grid {
    T minx, miny, minz;
    int xdiv, ydiv, zdiv;
    int budget, cap;
    const int32_t* counts;
    const int32_t* lists;
};

Grid<float> grid_from_meta(const int32_t* g, const int32_t* counts, const int32_t* lists) {
    Grid<float> r;
    std::memcpy(&r.minx, &g[0], 4);
    std::memcpy(&r.miny, &g[1], 4);
    std::memcpy(&r.minz, &g[2], 4);
    r.xdiv = g[3]; r.ydiv = g[4]; r.zdiv = g[5]; r.budget = g[6]; r.cap = g[7];
    r.counts = counts; r.lists = lists;
    return r;
}

// ---- GetAngleForce, propagate.cu:26-60 ------------------------------------
template <class T> V3<T> angle_force(V3<T> iPos, V3<T> kPos, T theta_o, T k, uint32_t& st) {
    T i_dot_k = dot(iPos, kPos);
    T ri_2 = mag2(iPos);
    T rk_2 = mag2(kPos);
    T c1, theta;
    if (is_f32<T>()) {
        // `+ 1e-3` is a double literal: the sum is double, then sqrtf takes it as float (:32)
        float inner = (float)((double)(ri_2 * rk_2 - i_dot_k * i_dot_k) + 1e-3);
        c1 = (T)(-1.0f / sqrtf(inner));
        // acos of a double argument, result stored into a float (:36); the acos is
        // the one the device kernels call (cs3d_acos.h), so that the float angle is
        // the same where acos lies near a float midpoint
        float den = sqrtf((float)ri_2) * sqrtf((float)rk_2);
        theta = (T)(float)cs3d_acos((double)i_dot_k / ((double)den + 1e-3));
    } else {
        c1 = T(-1) / std::sqrt(ri_2 * rk_2 - i_dot_k * i_dot_k + T(1e-3));
        theta = std::acos(i_dot_k / (std::sqrt(ri_2) * std::sqrt(rk_2) + T(1e-3)));
    }
    T c2 = i_dot_k / ri_2;
    T s = ((-k) * c1) * (theta - theta_o);
    V3<T> F = s * (kPos - c2 * iPos);
    if (!finite3(F)) st |= CS3D_SYS_ANGLE_NONFINITE;
    return F;
}

// ---- NeighNeighs, propagate.cu:63-78 --------------------------------------
template <class T>
void neigh_neighs(int nodeInd, int ni, int& nii, int& nij, int& nik, const T* th0, T& t1, T& t2) {
    if (nii == nodeInd) {
        t1 = th0[3 * ni + 0];  // aij
        t2 = th0[3 * ni + 2];  // aik
        nii = nij;
        nij = nik;
    } else if (nij == nodeInd) {
        t1 = th0[3 * ni + 0];  // aij
        t2 = th0[3 * ni + 1];  // ajk
        nij = nik;
    } else {
        t1 = th0[3 * ni + 2];  // aik
        t2 = th0[3 * ni + 1];  // ajk
    }
}

// ---- CalculateAngleForce, propagate.cu:81-181 ------------------------------
template <class T>
V3<T> angle_node(int p, const int32_t* nn, const T* x, const T* y, const T* z, const T* th0, T k, uint32_t& st) {
    auto P = [&](int i) { return mk(x[i], y[i], z[i]); };
    int ni = nn[0 * S + p], nj = nn[1 * S + p], nk = nn[2 * S + p];
    V3<T> nodeF = mk(T(0), T(0), T(0));
    V3<T> pos = P(p), niP = P(ni), njP = P(nj), nkP = P(nk);
    T aij = th0[3 * p + 0], ajk = th0[3 * p + 1], aik = th0[3 * p + 2];
    nodeF = nodeF - (angle_force(niP - pos, njP - pos, aij, k, st) + angle_force(njP - pos, niP - pos, aij, k, st));
    nodeF = nodeF - (angle_force(njP - pos, nkP - pos, ajk, k, st) + angle_force(nkP - pos, njP - pos, ajk, k, st));
    nodeF = nodeF - (angle_force(niP - pos, nkP - pos, aik, k, st) + angle_force(nkP - pos, niP - pos, aik, k, st));
    const int nbr[3] = {ni, nj, nk};
    for (int q = 0; q < 3; ++q) {
        int n = nbr[q];
        int a = nn[n + 0 * S], b = nn[n + 1 * S], c = nn[n + 2 * S];
        T t1, t2;
        neigh_neighs(p, n, a, b, c, th0, t1, t2);
        V3<T> nP = P(n);
        nodeF = nodeF + angle_force(pos - nP, P(a) - nP, t1, k, st);
        nodeF = nodeF + angle_force(pos - nP, P(b) - nP, t2, k, st);
    }
    return nodeF;
}

// voxel index of a node, propagate.cu:360-372
template <class T> inline int voxel_of(T X, T Y, T Z, const Grid<T>& g, T DL) {
    int sx = (int)((X - g.minx) / DL);
    if (sx < 0) sx = 0;
    if (sx >= g.xdiv) sx = g.xdiv - 1;
    int sy = (int)((Y - g.miny) / DL);
    if (sy < 0) sy = 0;
    if (sy >= g.ydiv) sy = g.ydiv - 1;
    int sz = (int)((Z - g.minz) / DL);
    if (sz < 0) sz = 0;
    if (sz >= g.zdiv) sz = g.zdiv - 1;
    return sz * g.xdiv * g.ydiv + sy * g.xdiv + sx;
}

// candidate-cell scan, propagate.cu:375-404 (and 603-627): list order is
// ascending (canonical).  Returns the candidates in `out`.
template <class T>
void candidates(int rank, T X, T Y, T Z, const Grid<T>& g, const T* bb, T ar, T DL, std::vector<int>& out,
                uint32_t& st) {
    out.clear();
    int index = voxel_of(X, Y, Z, g, DL);
    int cnt = g.counts[index];
    if (cnt > g.cap) cnt = g.cap;
    for (int r = 0; r < cnt; ++r) {
        int nb = g.lists[(int64_t)index * g.cap + r];
        if (nb == rank) continue;
        const T* b = bb + 6 * (int64_t)nb;
        T dX = (T)(X - b[1] > T(0)) * (X - b[1]);
        dX += (T)(b[0] - X > T(0)) * (b[0] - X);
        T dY = (T)(Y - b[3] > T(0)) * (Y - b[3]);
        dY += (T)(b[2] - Y > T(0)) * (b[2] - Y);
        T dZ = (T)(Z - b[5] > T(0)) * (Z - b[5]);
        dZ += (T)(b[4] - Z > T(0)) * (b[4] - Z);
        if (dX * dX + dY * dY + dZ * dZ > ar * ar) continue;
        out.push_back(nb);
    }
    if (out.size() > 10) st |= CS3D_SYS_NN_OVER_10;
}

// ---- CalculateConForce, propagate.cu:183-520 --------------------------------
template <class T>
uint32_t con_force(int n, const cs3d_params* p, const int32_t* nnT, const int32_t* /*sign*/, const T* R0,
                   const T* th0, const T* X, const T* Y, const T* Z, const T* CMx, const T* CMy, const T* CMz,
                   const T* Pr, const T* stiff, const T* bb, const Grid<T>& g, bool bonded_only, T* FXo,
                   T* FYo, T* FZo) {
    uint32_t status = 0;
    const T ar = (T)p->attraction_range, rr = (T)p->repulsion_range;
    const T as = (T)p->attraction_strength, rs = (T)p->repulsion_strength;
    const T E = (T)p->youngs_mod, DL = (T)p->grid_dl, thr = (T)p->thresh_dist, k = (T)p->angle_k;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : status)
    for (int rank = 0; rank < n; ++rank) {
        std::vector<int> local;
        local.reserve(64);
        const T* x = X + (int64_t)rank * S;
        const T* y = Y + (int64_t)rank * S;
        const T* z = Z + (int64_t)rank * S;
        for (int atom = 0; atom < NN; ++atom) {
            uint32_t st = 0;
            if (std::isnan(x[atom]) || std::isnan(y[atom]) || std::isnan(z[atom])) st |= CS3D_SYS_NAN;
            T stiffness = stiff[rank];
            T FX = 0, FY = 0, FZ = 0;
            for (int i = 0; i < 3; ++i) {  // springs, :298-314
                int N1 = nnT[i * S + atom];
                T dX = x[N1] - x[atom], dY = y[N1] - y[atom], dZ = z[N1] - z[atom];
                T R = std::sqrt(dX * dX + dY * dY + dZ * dZ);
                T r0 = R0[i * S + atom];
                FX += stiffness * (R - r0) / r0 * dX / R;
                FY += stiffness * (R - r0) / r0 * dY / R;
                FZ += stiffness * (R - r0) / r0 * dZ / R;
            }
            // pressure, :318-330 (CM is whatever CenterOfMass last wrote)
            V3<T> rCM = mk(x[atom] - CMx[rank], y[atom] - CMy[rank], z[atom] - CMz[rank]);
            V3<T> gF = (T(3) * Pr[rank]) * unit(rCM);
            FX += gF.x; FY += gF.y; FZ += gF.z;
            if (p->constrain_angles) {  // :332-337
                V3<T> t = angle_node(atom, nnT, x, y, z, th0, k, st);
                FX += t.x; FY += t.y; FZ += t.z;
            }
            if (!bonded_only) {
                T Xa = x[atom], Ya = y[atom], Za = z[atom];
                candidates(rank, Xa, Ya, Za, g, bb, ar, DL, local, st);
                V3<T> cf = mk(T(0), T(0), T(0));
                for (int nb : local) {  // :406-442
                    const T* xb = X + (int64_t)nb * S;
                    const T* yb = Y + (int64_t)nb * S;
                    const T* zb = Z + (int64_t)nb * S;
                    for (int j = 0; j < NN; ++j) {
                        T dX = Xa - xb[j], dY = Ya - yb[j], dZ = Za - zb[j];
                        T R = dX * dX + dY * dY + dZ * dZ;
                        if (R >= ar * ar) continue;
                        R = std::sqrt(R);
                        if (R < ar) {
                            cf.x += -as * E * (ar - R) / R * dX;
                            cf.y += -as * E * (ar - R) / R * dY;
                            cf.z += -as * E * (ar - R) / R * dZ;
                        }
                        if (R <= rr) {
                            cf.x += +rs * E * (rr - R) / R * dX;
                            cf.y += +rs * E * (rr - R) / R * dY;
                            cf.z += +rs * E * (rr - R) / R * dZ;
                        }
                    }
                }
                FX += cf.x; FY += cf.y; FZ += cf.z;
                if (p->use_rigid_box) {  // :473-510 (lower wall at 0)
                    T gap1 = Xa, gap2 = (T)p->box_max[0] - Xa;
                    if (gap1 < thr) FX += T(-100) * E * (gap1 - thr);
                    if (gap2 < thr) FX += T(100) * E * (gap2 - thr);
                    gap1 = Ya; gap2 = (T)p->box_max[1] - Ya;
                    if (gap1 < thr) FY += T(-100) * E * (gap1 - thr);
                    if (gap2 < thr) FY += T(100) * E * (gap2 - thr);
                    gap1 = Za; gap2 = (T)p->box_max[2] - Za;
                    if (gap1 < thr) FZ += T(-100) * E * (gap1 - thr);
                    if (gap2 < thr) FZ += T(100) * E * (gap2 - thr);
                }
            }
            int64_t gi = (int64_t)rank * S + atom;
            FXo[gi] = FX; FYo[gi] = FY; FZo[gi] = FZ;
            status |= st;
        }
    }
    return status;
}

// ---- CalculateDisForce, propagate.cu:523-665 --------------------------------
template <class T>
uint32_t dis_force(int n, const cs3d_params* p, const int32_t* nnT, const int32_t* sign, const T* X, const T* Y,
                   const T* Z, const T* VX, const T* VY, const T* VZ, const T* bb, const Grid<T>& g, T* FXo,
                   T* FYo, T* FZo) {
    uint32_t status = 0;
    const T ar = (T)p->attraction_range, DL = (T)p->grid_dl;
    const T gint = (T)p->internal_damping, gext = (T)p->viscotic_damping, go = (T)p->gamma_visc;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : status)
    for (int c = 0; c < n; ++c) {
        std::vector<int> local;
        local.reserve(64);
        const int64_t base = (int64_t)c * S;
        for (int i = 0; i < NN; ++i) {
            uint32_t st = 0;
            int64_t gi = base + i;
            V3<T> force = mk(T(0), T(0), T(0));
            V3<T> vi = mk(VX[gi], VY[gi], VZ[gi]);
            for (int k = 0; k < 3; ++k) {  // bond damping :547-554
                int N = nnT[k * S + i];
                V3<T> vn = mk(VX[base + N], VY[base + N], VZ[base + N]);
                force = force - gint * (vi - vn);
            }
            int N1 = nnT[i], N2 = nnT[S + i], N3 = nnT[2 * S + i];  // normal :558-572
            V3<T> A = mk(X[base + N2] - X[base + N1], Y[base + N2] - Y[base + N1], Z[base + N2] - Z[base + N1]);
            V3<T> B = mk(X[base + N3] - X[base + N1], Y[base + N3] - Y[base + N1], Z[base + N3] - Z[base + N1]);
            V3<T> nrm = unit(cross(A, B));
            nrm = (T)sign[i] * nrm;
            T Xa = X[gi], Ya = Y[gi], Za = Z[gi];
            candidates(c, Xa, Ya, Za, g, bb, ar, DL, local, st);
            for (int nb : local) {  // :629-655
                const int64_t ob = (int64_t)nb * S;
                for (int j = 0; j < NN; ++j) {
                    T dX = Xa - X[ob + j], dY = Ya - Y[ob + j], dZ = Za - Z[ob + j];
                    T R = dX * dX + dY * dY + dZ * dZ;
                    if (R > ar * ar) continue;
                    V3<T> vj = mk(VX[ob + j], VY[ob + j], VZ[ob + j]);
                    V3<T> vij = vi - vj;
                    V3<T> vTau = vij - dot(vij, nrm) * nrm;
                    force = force - gext * vTau;
                }
            }
            force = force - go * vi;  // drag :658
            FXo[gi] = force.x; FYo[gi] = force.y; FZo[gi] = force.z;
            status |= st;
        }
    }
    return status;
}

// ---- random force -----------------------------------------------------------
// The reference's CalculateRanForce (propagate.cu:667-676) draws curand_normal
// from cuRAND states, a closed library that is not reproduced.  The engine's
// documented replacement: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with
// counter {cell * 192 + node, step low, step high, attempt} and key {seed,
// 0x43533344}, standard normals by the Marsaglia polar method, the logarithm
// as ln m = 2 atanh((m-1)/(m+1)) to z^9 after splitting off the binary exponent.
inline uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
        const uint32_t h0 = mulhi32(m0, c[0]), l0 = m0 * c[0];
        const uint32_t h1 = mulhi32(m1, c[2]), l1 = m1 * c[2];
        const uint32_t x0 = h1 ^ c[1] ^ k0, x2 = h0 ^ c[3] ^ k1;
        c[0] = x0; c[1] = l1; c[2] = x2; c[3] = l0;
        k0 += 0x9E3779B9u;  // Weyl constants of Philox
        k1 += 0xBB67AE85u;
    }
}

float ln_unit(float s) {
    uint32_t b;
    std::memcpy(&b, &s, 4);
    int e = (int)((b >> 23) & 0xffu) - 126;
    const uint32_t mb = (b & 0x7fffffu) | 0x3f000000u;
    float m;
    std::memcpy(&m, &mb, 4);
    if (m < 0.70710678f) { m = m + m; e -= 1; }
    const float z = (m - 1.0f) / (m + 1.0f), z2 = z * z;
    float q = 0.11111111f;
    q = q * z2 + 0.14285715f;
    q = q * z2 + 0.2f;
    q = q * z2 + 0.33333334f;
    q = q * z2 + 1.0f;
    const float fe = (float)e;
    return fe * 0.693145751953125f + (2.0f * z * q + fe * 1.4286068e-06f);
}

void normals3(uint32_t seed, uint32_t node, int64_t step, float z[3]) {
    int got = 0;
    z[0] = z[1] = z[2] = 0.f;
    for (uint32_t attempt = 0; got < 3 && attempt < 32; ++attempt) {
        uint32_t c[4] = {node, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), attempt};
        philox10(c, seed, 0x43533344u);
        for (int pr = 0; pr < 2; ++pr) {
            const float v1 = 2.0f * (((float)(c[2 * pr] >> 9) + 0.5f) * 1.1920929e-07f) - 1.0f;
            const float v2 = 2.0f * (((float)(c[2 * pr + 1] >> 9) + 0.5f) * 1.1920929e-07f) - 1.0f;
            const float s = v1 * v1 + v2 * v2;
            if (got < 3 && s > 0.0f && s < 1.0f) {
                const float f = std::sqrt(-2.0f * ln_unit(s) / s);
                z[got++] = v1 * f;
                if (got < 3) z[got++] = v2 * f;
            }
        }
    }
}

// ---- Integrate + ForwardTime, propagate.cu:681-712, 738-763 -----------------
template <class T>
void integrate(int n, const cs3d_params* p, T* X, T* Y, T* Z, T* VX, T* VY, T* VZ, const T* FCX, const T* FCY,
               const T* FCZ, const T* FDX, const T* FDY, const T* FDZ, const T* FRX = nullptr,
               const T* FRY = nullptr, const T* FRZ = nullptr) {
    const T dt = (T)p->dt, m = (T)p->mass;
    const T root_dt = std::sqrt(dt);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < NN; ++i) {
            int64_t gi = (int64_t)c * S + i;
            T* V[3] = {VX + gi, VY + gi, VZ + gi};
            T* P[3] = {X + gi, Y + gi, Z + gi};
            const T FC[3] = {FCX[gi], FCY[gi], FCZ[gi]}, FD[3] = {FDX[gi], FDY[gi], FDZ[gi]};
            const T FR[3] = {FRX ? FRX[gi] : T(0), FRY ? FRY[gi] : T(0), FRZ ? FRZ[gi] : T(0)};
            for (int a = 0; a < 3; ++a) {
                const T fr = FR[a];
                if (is_f32<T>())
                    *V[a] = (T)((double)*V[a] + 0.5 * (double)(dt * FC[a] + dt * FD[a] + root_dt * fr) / (double)m);
                else
                    *V[a] = *V[a] + T(0.5) * (dt * FC[a] + dt * FD[a] + root_dt * fr) / m;
            }
            for (int a = 0; a < 3; ++a) *P[a] = *P[a] + *V[a] * dt;  // x_P = x + v dt; x <- x_P
        }
    }
}

// ---- VelocityUpdateA then VelocityUpdateB, propagate.cu:714-736 --------------
template <class T>
void velocity_ab(int n, const cs3d_params* p, T* VX, T* VY, T* VZ, const T* FCX, const T* FCY, const T* FCZ,
                 const T* FDX, const T* FDY, const T* FDZ, const T* FRX = nullptr, const T* FRY = nullptr,
                 const T* FRZ = nullptr) {
    const T dt = (T)p->dt, m = (T)p->mass;
    const T root_dt = std::sqrt(dt);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < NN; ++i) {
            int64_t gi = (int64_t)c * S + i;
            T* V[3] = {VX + gi, VY + gi, VZ + gi};
            const T FC[3] = {FCX[gi], FCY[gi], FCZ[gi]}, FD[3] = {FDX[gi], FDY[gi], FDZ[gi]};
            const T FR[3] = {FRX ? FRX[gi] : T(0), FRY ? FRY[gi] : T(0), FRZ ? FRZ[gi] : T(0)};
            for (int a = 0; a < 3; ++a) {
                const T fr = FR[a];
                if (is_f32<T>())
                    *V[a] = (T)((double)*V[a] + 0.5 * (double)(dt * FC[a] + root_dt * fr) / (double)m);
                else
                    *V[a] = *V[a] + T(0.5) * (dt * FC[a] + root_dt * fr) / m;
            }
            for (int a = 0; a < 3; ++a) {
                if (is_f32<T>())
                    *V[a] = (T)((double)*V[a] + 0.5 * (double)dt * (double)FD[a] / (double)m);
                else
                    *V[a] = *V[a] + T(0.5) * dt * FD[a] / m;
            }
        }
    }
}

// ---- CenterOfMass, centermass.cu:8-50 (256-slot tree) ------------------------
template <class T> void com(int n, const T* X, const T* Y, const T* Z, T* CMx, T* CMy, T* CMz) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        const T* src[3] = {X + (int64_t)c * S, Y + (int64_t)c * S, Z + (int64_t)c * S};
        T* dst[3] = {CMx + c, CMy + c, CMz + c};
        for (int a = 0; a < 3; ++a) {
            T s[256];
            for (int t = 0; t < 256; ++t) s[t] = t < NN ? src[a][t] : T(0);
            for (int st = 128; st > 0; st >>= 1)
                for (int t = 0; t < st; ++t) s[t] += s[t + st];
            *dst[a] = s[0] / T(180);
        }
    }
}

// pairwise tree over 128 slots (canonical order for the volume/area sums)
template <class T> T tree128(const T* v, int cnt) {
    T s[128];
    for (int t = 0; t < 128; ++t) s[t] = t < cnt ? v[t] : T(0);
    for (int st = 64; st > 0; st >>= 1)
        for (int t = 0; t < st; ++t) s[t] += s[t + st];
    return s[0];
}

// ---- volumes, volume.cu:5-163 --------------------------------------------------
template <class T>
uint32_t volumes(int n, const cs3d_params* p, const int32_t* F, const T* X, const T* Y, const T* Z, const T* CMx,
                 const T* CMy, const T* CMz, const T* stiff, T* vol, T* area, int8_t* flag) {
    uint32_t status = 0;
    const bool sph = p->check_sphericity != 0;
#pragma omp parallel for schedule(static) reduction(| : status)
    for (int c = 0; c < n; ++c) {
        T lx[S], ly[S], lz[S];
        for (int t = 0; t < NN; ++t) {
            lx[t] = X[(int64_t)c * S + t] - CMx[c];
            ly[t] = Y[(int64_t)c * S + t] - CMy[c];
            lz[t] = Z[(int64_t)c * S + t] - CMz[c];
        }
        T fv[CS3D_FACES], fa[CS3D_FACES];
        for (int f = 0; f < CS3D_FACES; ++f) {
            const int32_t* r = F + 7 * f;
            T avX = 0, avY = 0, avZ = 0;
            for (int i = 0; i < 5; ++i) { avX += lx[r[i]]; avY += ly[r[i]]; avZ += lz[r[i]]; }
            T fac = (T)0.166666667f;
            if (f < 12) fac = (T)0.2f;
            else { avX += lx[r[5]]; avY += ly[r[5]]; avZ += lz[r[5]]; }
            avX *= fac; avY *= fac; avZ *= fac;
            T totvol = 0, faceArea = 0;
            for (int i = 0; i < 6; ++i) {
                int a = r[i], b = r[i + 1];
                T n1 = (ly[b] * avZ - avY * lz[b]) * lx[a];
                T n2 = (lz[b] * avX - avZ * lx[b]) * ly[a];
                T n3 = (lx[b] * avY - avX * ly[b]) * lz[a];
                totvol += std::fabs(n1 + n2 + n3);
                if (sph) {
                    T x1 = lx[a] - avX, y1 = ly[a] - avY, z1 = lz[a] - avZ;
                    T x2 = lx[b] - avX, y2 = ly[b] - avY, z2 = lz[b] - avZ;
                    T xx = y1 * z2 - z1 * y2, yy = z1 * x2 - x1 * z2, zz = x1 * y2 - y1 * x2;
                    if (is_f32<T>())
                        faceArea = (T)((double)faceArea + 0.5 * (double)sqrtf((float)(xx * xx + yy * yy + zz * zz)));
                    else
                        faceArea += T(0.5) * std::sqrt(xx * xx + yy * yy + zz * zz);
                }
            }
            fv[f] = totvol;
            fa[f] = faceArea;
        }
        T volume = tree128(fv, CS3D_FACES);
        T ar = sph ? tree128(fa, CS3D_FACES) : T(0);
        if (is_f32<T>()) volume = (T)((double)volume / 6.0);
        else volume = volume / T(6);
        vol[c] = volume;
        if (!std::isfinite(volume)) status |= CS3D_SYS_VOLUME_NONFINITE;
        int8_t fl = 0;
        if (volume > (T)p->division_vol) fl = 1;
        if (sph) {
            area[c] = ar;
            T cb, psi;
            if (is_f32<T>()) {
                cb = cbrtf((float)volume);
                psi = (T)(float)(4.835975862049408 * (double)cb * (double)cb / (double)ar);
                if ((double)std::fabs(1.0f - (float)psi) > 0.05) fl = 0;
            } else {
                cb = std::cbrt(volume);
                psi = T(4.835975862049408) * cb * cb / ar;
                if (std::fabs(T(1) - psi) > T(0.05)) fl = 0;
            }
        }
        if (p->use_different_stiffnesses && p->recalc_r0 && stiff[c] != (T)p->stiffness1) fl = 0;
        flag[c] = fl;
    }
    return status;
}

// ---- bounding_boxes, propagatebound.cu:10-141 --------------------------------
template <class T> void bboxes(int n, const T* X, const T* Y, const T* Z, T* bb) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        const T* s[3] = {X + (int64_t)c * S, Y + (int64_t)c * S, Z + (int64_t)c * S};
        for (int a = 0; a < 3; ++a) {
            T mn = s[a][0], mx = s[a][0];
            for (int t = 1; t < NN; ++t) {
                if (mn > s[a][t]) mn = s[a][t];
                if (mx < s[a][t]) mx = s[a][t];
            }
            bb[6 * (int64_t)c + 2 * a] = mn;
            bb[6 * (int64_t)c + 2 * a + 1] = mx;
        }
    }
}

// ---- minmaxpre + minmaxpost, propagatebound.cu:145-204, 266-323 ---------------
template <class T> void extent(int n, const T* bb, T* o) {
    T mn[3] = {T(1.0E8f), T(1.0E8f), T(1.0E8f)}, mx[3] = {T(-1.0E8f), T(-1.0E8f), T(-1.0E8f)};
    for (int c = 0; c < n; ++c)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::fmin(mn[a], bb[6 * (int64_t)c + 2 * a]);
            mx[a] = std::fmax(mx[a], bb[6 * (int64_t)c + 2 * a + 1]);
        }
    for (int a = 0; a < 3; ++a) { o[2 * a] = mn[a]; o[2 * a + 1] = mx[a]; }
}

// ---- host grid dims (GPUbounce.cu:1331-1334) + makeNNlist (propagatebound.cu:209-263)
// Lists are filled in ascending cell order, so every list is ascending.
template <class T>
uint32_t grid_build(int n, const cs3d_params* p, const T* bb, int32_t budget, int32_t cap, T* gmin, int* dims,
                    int32_t* counts, int32_t* lists) {
    T e[6];
    extent(n, bb, e);
    const T DL = (T)p->grid_dl, ar = (T)p->attraction_range;
    int xd = (int)((e[1] - e[0]) / DL + T(1));
    int yd = (int)((e[3] - e[2]) / DL + T(1));
    int zd = (int)((e[5] - e[4]) / DL + T(1));
    gmin[0] = e[0]; gmin[1] = e[2]; gmin[2] = e[4];
    dims[0] = xd; dims[1] = yd; dims[2] = zd;
    int64_t nv = (int64_t)xd * yd * zd;
    if (nv > budget || xd <= 0 || yd <= 0 || zd <= 0) {
        // flagged; the grid falls back to one voxel (as the engine's), so later
        // voxel lookups stay inside the budget
        dims[0] = dims[1] = dims[2] = 1;
        std::fill(counts, counts + budget, 0);
        return CS3D_SYS_GRID_TOO_LARGE;
    }
    std::fill(counts, counts + nv, 0);
    uint32_t st = 0;
    for (int c = 0; c < n; ++c) {
        const T* b = bb + 6 * (int64_t)c;
        int sx = (int)((b[0] - ar - gmin[0]) / DL); if (sx < 0) sx = 0;
        int ex = (int)((b[1] + ar - gmin[0]) / DL); if (ex >= xd) ex = xd - 1;
        int sy = (int)((b[2] - ar - gmin[1]) / DL); if (sy < 0) sy = 0;
        int ey = (int)((b[3] + ar - gmin[1]) / DL); if (ey >= yd) ey = yd - 1;
        int sz = (int)((b[4] - ar - gmin[2]) / DL); if (sz < 0) sz = 0;
        int ez = (int)((b[5] + ar - gmin[2]) / DL); if (ez >= zd) ez = zd - 1;
        for (int j1 = sx; j1 <= ex; ++j1)
            for (int j2 = sy; j2 <= ey; ++j2)
                for (int j3 = sz; j3 <= ez; ++j3) {
                    int v = j3 * xd * yd + j2 * xd + j1;
                    int slot = counts[v]++;
                    if (slot < cap) lists[(int64_t)v * cap + slot] = c;
                    else st |= CS3D_SYS_VOXEL_OVERFLOW;
                }
    }
    return st;
}

// ---- PressureUpdate, PressureKernels.cu:4-41; rGrowth GPUbounce.cu:1077-1088 --------
template <class T> T rgrowth(const cs3d_params* p, int n) {
    if (p->do_pop_model == 1) {
        float r = (float)((double)p->growth_rate * (1 - (n * 1.0 / (double)p->max_pop)));
        if (r < 0) r = 0;
        return (T)r;
    }
    return (T)p->growth_rate;
}

template <class T>
void pressure_update(int n, const cs3d_params* p, T* P, const T* stiff, const T* scale, int64_t step, T inc0) {
    const T mx = (T)p->max_pressure, s1 = (T)p->stiffness1;
    for (int c = 0; c < n; ++c) {
        T inc = scale ? inc0 * scale[c] : inc0;
        if (p->use_different_stiffnesses) {
            if (step < p->phase_count) {
                if (P[c] < mx) P[c] += inc;
            } else if (stiff[c] == s1) {
                if (P[c] < mx) P[c] += inc;
            } else if (P[c] != T(0)) {
                P[c] = T(0);
            }
        } else if (P[c] < mx) {
            P[c] += inc;
        }
    }
}

}  // namespace

// ============================== RANMAR ==========================================
// rmarin, marsaglia.h:54-113
extern "C" void or_acos(const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] = cs3d_acos(x[i]);
}

extern "C" int or_ranmar_init(cs3d_ranmar_state* st, int ij, int kl) {
    if (ij < 0 || ij > 31328 || kl < 0 || kl > 30081) return 1;
    int i = (ij / 177) % 177 + 2;
    int j = ij % 177 + 2;
    int k = (kl / 169) % 178 + 1;
    int l = kl % 169;
    for (int ii = 1; ii <= 97; ii++) {
        float s = 0.0, t = 0.5;
        for (int jj = 1; jj <= 24; jj++) {
            int m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t *= 0.5;
        }
        st->u[ii] = s;
    }
    st->u[0] = 0;
    st->c = 362436.0 / 16777216.0;
    st->cd = 7654321.0 / 16777216.0;
    st->cm = 16777213.0 / 16777216.0;
    st->i97 = 97;
    st->j97 = 33;
    st->ready = 1;
    return 0;
}

// ranmar, marsaglia.h:115-144 (double literals kept; all values are 24-bit exact)
extern "C" void or_ranmar(cs3d_ranmar_state* st, float* rvec, int len) {
    for (int iv = 0; iv < len; iv++) {
        float uni = st->u[st->i97] - st->u[st->j97];
        if (uni < 0.0) uni += 1.0;
        st->u[st->i97] = uni;
        st->i97--;
        if (st->i97 == 0) st->i97 = 97;
        st->j97--;
        if (st->j97 == 0) st->j97 = 97;
        st->c -= st->cd;
        if (st->c < 0.0) st->c += st->cm;
        uni -= st->c;
        if (uni < 0.0) uni += 1.0;
        rvec[iv] = uni;
    }
}

// GetRandomVector, RandomVector.h:8-24 (float overload of sqrt)
extern "C" void or_random_vector(cs3d_ranmar_state* st, float* n) {
    float ran2[2];
    float s = 0;
    do {
        or_ranmar(st, ran2, 2);
        ran2[0] = 2.0f * ran2[0] - 1.0f;
        ran2[1] = 2.0f * ran2[1] - 1.0f;
        s = ran2[0] * ran2[0] + ran2[1] * ran2[1];
    } while (s >= 1.0f);
    float x1 = ran2[0], x2 = ran2[1];
    n[0] = 2 * x1 * sqrtf(1 - x1 * x1 - x2 * x2);
    n[1] = 2 * x2 * sqrtf(1 - x1 * x1 - x2 * x2);
    n[2] = 1 - 2 * (x1 * x1 + x2 * x2);
}

// GetRandomVectorBasis, RandomVector.h:26-78 (orthogonalisation typo w[2]*w[2] kept)
extern "C" void or_random_vector_basis(cs3d_ranmar_state* st, float* n, const float* basis) {
    float ran[1];
    or_ranmar(st, ran, 1);
    float Nv = (float)std::cos(2 * 3.14159 * (double)ran[0]);
    float Nw = (float)std::sin(2 * 3.14159 * (double)ran[0]);
    float v[3], w[3];
    if (basis[1] != 0) {
        v[0] = 0; v[1] = basis[2]; v[2] = -1 * basis[1];
        w[0] = basis[1]; w[1] = -1 * basis[0]; w[2] = 0;
    } else {
        v[0] = 0; v[1] = 1; v[2] = 0;
        w[0] = basis[2]; w[1] = 0; w[2] = -1 * basis[0];
    }
    float f = (w[0] * v[0] + w[1] * v[1] + w[2] * w[2]) / (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    w[0] = w[0] - f * v[0];
    w[1] = w[1] - f * v[1];
    w[2] = w[2] - f * v[2];
    f = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] / f; v[1] = v[1] / f; v[2] = v[2] / f;
    f = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    w[0] = w[0] / f; w[1] = w[1] / f; w[2] = w[2] / f;
    n[0] = Nv * v[0] + Nw * w[0];
    n[1] = Nv * v[1] + Nw * w[1];
    n[2] = Nv * v[2] + Nw * w[2];
}

// ============================== mesh rest state =================================
extern "C" void or_mesh_rest(const int32_t* nn, const float* X, const float* Y, const float* Z, float* R0,
                             float* theta0) {
    std::memset(R0, 0, sizeof(float) * 3 * S);
    std::memset(theta0, 0, sizeof(float) * 3 * S);
    for (int i = 0; i < NN; ++i) {  // GPUbounce.cu:1892-1908
        V3<float> a = mk(X[i], Y[i], Z[i]);
        for (int k = 0; k < 3; ++k) {
            int N = nn[k * S + i];
            R0[k * S + i] = mag(a - mk(X[N], Y[N], Z[N]));
        }
    }
    for (int i = 0; i < NN; ++i) {  // GPUbounce.cu:912-933 (float acos overload)
        V3<float> p = mk(X[i], Y[i], Z[i]);
        auto P = [&](int k) { int N = nn[k * S + i]; return mk(X[N], Y[N], Z[N]) - p; };
        V3<float> ni = P(0), nj = P(1), nk = P(2);
        theta0[3 * i + 0] = acosf(dot(ni, nj) / (mag(ni) * mag(nj)));
        theta0[3 * i + 1] = acosf(dot(nj, nk) / (mag(nj) * mag(nk)));
        theta0[3 * i + 2] = acosf(dot(ni, nk) / (mag(ni) * mag(nk)));
    }
}

// ============================== C wrappers, float mode ==========================
extern "C" void or_com(int n, const float* X, const float* Y, const float* Z, float* CMx, float* CMy, float* CMz) {
    com<float>(n, X, Y, Z, CMx, CMy, CMz);
}
extern "C" uint32_t or_volumes(int n, const cs3d_params* p, const int32_t* faces, const float* X, const float* Y,
                               const float* Z, const float* CMx, const float* CMy, const float* CMz,
                               const float* stiffness, int /*step*/, float* vol, float* area, int8_t* flag) {
    return volumes<float>(n, p, faces, X, Y, Z, CMx, CMy, CMz, stiffness, vol, area, flag);
}
extern "C" void or_bboxes(int n, const float* X, const float* Y, const float* Z, float* bbox) {
    bboxes<float>(n, X, Y, Z, bbox);
}
extern "C" void or_extent(int n, const float* bbox, float* out6) { extent<float>(n, bbox, out6); }
extern "C" uint32_t or_grid_build(int n, const cs3d_params* p, const float* bbox, int32_t budget, int32_t* grid,
                                  int32_t* counts, int32_t* lists) {
    float gmin[3];
    int dims[3];
    uint32_t st = grid_build<float>(n, p, bbox, budget, p->voxel_capacity, gmin, dims, counts, lists);
    std::memcpy(&grid[0], &gmin[0], 4);
    std::memcpy(&grid[1], &gmin[1], 4);
    std::memcpy(&grid[2], &gmin[2], 4);
    grid[3] = dims[0]; grid[4] = dims[1]; grid[5] = dims[2];
    grid[6] = budget; grid[7] = p->voxel_capacity;
    return st;
}
extern "C" uint32_t or_con_force(int n, const cs3d_params* p, const cs3d_mesh* m, const float* X, const float* Y,
                                 const float* Z, const float* CMx, const float* CMy, const float* CMz,
                                 const float* P, const float* stiffness, const float* bbox, const int32_t* grid,
                                 const int32_t* counts, const int32_t* lists, int bonded_only, float* FX,
                                 float* FY, float* FZ) {
    Grid<float> g = grid_from_meta(grid, counts, lists);
    return con_force<float>(n, p, m->nn, m->sign, m->R0, m->theta0, X, Y, Z, CMx, CMy, CMz, P, stiffness, bbox, g,
                            bonded_only != 0, FX, FY, FZ);
}
extern "C" uint32_t or_dis_force(int n, const cs3d_params* p, const cs3d_mesh* m, const float* X, const float* Y,
                                 const float* Z, const float* VX, const float* VY, const float* VZ,
                                 const float* bbox, const int32_t* grid, const int32_t* counts,
                                 const int32_t* lists, float* FX, float* FY, float* FZ) {
    Grid<float> g = grid_from_meta(grid, counts, lists);
    return dis_force<float>(n, p, m->nn, m->sign, X, Y, Z, VX, VY, VZ, bbox, g, FX, FY, FZ);
}
extern "C" void or_integrate(int n, const cs3d_params* p, float* X, float* Y, float* Z, float* VX, float* VY,
                             float* VZ, const float* FCX, const float* FCY, const float* FCZ, const float* FDX,
                             const float* FDY, const float* FDZ) {
    integrate<float>(n, p, X, Y, Z, VX, VY, VZ, FCX, FCY, FCZ, FDX, FDY, FDZ);
}
extern "C" void or_velocity_ab(int n, const cs3d_params* p, float* VX, float* VY, float* VZ, const float* FCX,
                               const float* FCY, const float* FCZ, const float* FDX, const float* FDY,
                               const float* FDZ) {
    velocity_ab<float>(n, p, VX, VY, VZ, FCX, FCY, FCZ, FDX, FDY, FDZ);
}
extern "C" float or_rgrowth(const cs3d_params* p, int n_cells) { return rgrowth<float>(p, n_cells); }
extern "C" void or_philox(uint32_t* c, uint32_t k0, uint32_t k1) { philox10(c, k0, k1); }
extern "C" float or_ln_unit(float s) { return ln_unit(s); }
extern "C" void or_normals3(uint32_t seed, uint32_t node, int64_t step, float* z) { normals3(seed, node, step, z); }
extern "C" void or_pressure_update(int n, const cs3d_params* p, float* P, const float* stiffness,
                                   const float* scale, int step, float rg) {
    pressure_update<float>(n, p, P, stiffness, scale, step, rg);
}

// ============================== full simulation =================================
namespace {

struct SimBase {
    virtual ~SimBase() {}
    virtual uint32_t init() = 0;
    virtual uint32_t step(int k) = 0;
    // wall seconds spent per part of the step since construction (CPU baseline breakdown):
    // 0 Integrate + PressureUpdate, 1 CalculateConForce, 2 both CalculateDisForce,
    // 3 random force + VelocityUpdateA/B, 4 CenterOfMass + volumes + division,
    // 5 grid rebuild + CorrectCoMMotion
    double phase_s[6] = {0, 0, 0, 0, 0, 0};
    virtual size_t get(int f, void* out, size_t bytes) const = 0;
    virtual size_t set(int f, const void* in, size_t bytes) = 0;
    virtual size_t get64(int f, double* out, size_t count) const = 0;
    int n = 0, cap = 0;
    int64_t stepno = 0;
};

template <class T> struct Sim : SimBase {
    cs3d_params p;
    cs3d_mesh m;
    std::vector<T> R0, th0;
    std::vector<T> X, Y, Z, VX, VY, VZ, FCX, FCY, FCZ, FDX, FDY, FDZ, VAX, VAY, VAZ, FRX, FRY, FRZ;
    std::vector<T> CMx, CMy, CMz, bb, P, E, vol, area, gscale, patch;
    std::vector<int8_t> flag;
    int32_t budget = 0;
    T gmin[3] = {0, 0, 0};
    int dims[3] = {1, 1, 1};
    std::vector<int32_t> counts, lists;
    cs3d_ranmar_state rng;
    uint32_t status = 0;
    float sys_cm[3] = {0.f, 0.f, 0.f};

    Sim(const cs3d_params& pp, const cs3d_mesh& mm, int n0, int cap0, int32_t bud, const float* x, const float* y,
        const float* z) : p(pp), m(mm) {
        n = n0; cap = cap0; budget = bud;
        size_t ns = (size_t)cap * S;
        for (auto* v : {&X, &Y, &Z, &VX, &VY, &VZ, &FCX, &FCY, &FCZ, &FDX, &FDY, &FDZ, &VAX, &VAY, &VAZ, &FRX, &FRY,
                        &FRZ})
            v->assign(ns, T(0));
        for (auto* v : {&CMx, &CMy, &CMz, &P, &E, &vol, &area}) v->assign(cap, T(0));
        gscale.assign(cap, T(1));
        bb.assign((size_t)cap * 6, T(0));
        flag.assign(cap, 0);
        counts.assign(budget, 0);
        lists.assign((size_t)budget * p.voxel_capacity, 0);
        R0.assign(m.R0, m.R0 + 3 * S);
        th0.assign(m.theta0, m.theta0 + 3 * S);
        if (x)
            for (size_t i = 0; i < (size_t)n * S; ++i) { X[i] = x[i]; Y[i] = y[i]; Z[i] = z[i]; }
        for (int c = 0; c < cap; ++c) { P[c] = (T)p.min_pressure; E[c] = (T)p.stiffness1; }
        or_ranmar_init(&rng, p.ranmar_ij, p.ranmar_kl);
    }
    Grid<T> grid() const {
        Grid<T> g;
        g.minx = gmin[0]; g.miny = gmin[1]; g.minz = gmin[2];
        g.xdiv = dims[0]; g.ydiv = dims[1]; g.zdiv = dims[2];
        g.budget = budget; g.cap = p.voxel_capacity;
        g.counts = counts.data(); g.lists = lists.data();
        return g;
    }
    uint32_t rebuild_grid() {
        bboxes<T>(n, X.data(), Y.data(), Z.data(), bb.data());
        return grid_build<T>(n, &p, bb.data(), budget, p.voxel_capacity, gmin, dims, counts.data(), lists.data());
    }
    uint32_t conforce() {
        return con_force<T>(n, &p, m.nn, m.sign, R0.data(), th0.data(), X.data(), Y.data(), Z.data(), CMx.data(),
                            CMy.data(), CMz.data(), P.data(), E.data(), bb.data(), grid(), false, FCX.data(),
                            FCY.data(), FCZ.data());
    }
    uint32_t disforce() {
        return dis_force<T>(n, &p, m.nn, m.sign, X.data(), Y.data(), Z.data(), VX.data(), VY.data(), VZ.data(),
                            bb.data(), grid(), FDX.data(), FDY.data(), FDZ.data());
    }
    // start-up of main: GPUbounce.cu:739-763 (CoM, bbox, extent, grid), 867-872 (P = Pmin), 978-1004 (forces)
    uint32_t init() override {
        uint32_t st = 0;
        com<T>(n, X.data(), Y.data(), Z.data(), CMx.data(), CMy.data(), CMz.data());
        st |= rebuild_grid();
        for (int c = 0; c < n; ++c) P[c] = (T)p.min_pressure;
        st |= conforce();
        st |= disforce();
        status |= st;
        return st;
    }
    // division phase, GPUbounce.cu:1222-1286 + cell_division postscriptinit.cu:148-256
    uint32_t divide(bool& any) {
        uint32_t st = 0;
        std::vector<int> ev;
        for (int c = 0; c < n; ++c)
            if (flag[c] == 1) { flag[c] = 0; ev.push_back(c); }  // count_and_get_div :2489-2500
        any = !ev.empty();
        std::vector<int> reset;
        const T half = (T)(float)(0.5 * (double)p.repulsion_range);
        for (size_t e = 0; e < ev.size(); ++e) {
            int rank = ev[e];
            float nrm[3] = {0, 1, 0};
            if (p.use_div_plane_basis) or_random_vector_basis(&rng, nrm, p.div_plane_basis);
            else or_random_vector(&rng, nrm);
            if (n >= cap) { st |= CS3D_SYS_CAPACITY; break; }
            int d = n;
            T Nx = nrm[0], Ny = nrm[1], Nz = nrm[2];
            if (is_f32<T>()) {
                if ((double)std::fabs(sqrtf(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]) - 1) > 1e-3)
                    st |= CS3D_SYS_BAD_NORMAL;
            }
            T cmx = CMx[rank], cmy = CMy[rank], cmz = CMz[rank];
            for (int atom = 0; atom < NN; ++atom) {
                int64_t pi = (int64_t)rank * S + atom, di = (int64_t)d * S + atom;
                T x0 = X[pi], y0 = Y[pi], z0 = Z[pi];
                for (int side = 0; side < 2; ++side) {
                    T nx = side ? T(-1) * Nx : Nx, ny = side ? T(-1) * Ny : Ny, nz = side ? T(-1) * Nz : Nz;
                    T x = x0 - cmx - half * nx, y = y0 - cmy - half * ny, z = z0 - cmz - half * nz;
                    T pd = x * nx + y * ny + z * nz;
                    if (pd < T(0) || pd == 0) { x = x - pd * nx; y = y - pd * ny; z = z - pd * nz; }
                    int64_t o = side ? di : pi;
                    X[o] = x + (cmx + half * nx);
                    Y[o] = y + (cmy + half * ny);
                    Z[o] = z + (cmz + half * nz);
                }
                VX[di] = VX[pi]; VY[di] = VY[pi]; VZ[di] = VZ[pi];
            }
            E[d] = p.daught_same_stiffness ? E[rank] : (T)p.stiffness1;
            gscale[d] = gscale[rank];
            reset.push_back(rank);
            reset.push_back(d);
            ++n;
        }
        for (int c : reset) P[c] = (T)p.min_pressure;  // PressureReset, PressureKernels.cu:45-65
        if (p.daughter_com_fix && !reset.empty()) {
            std::vector<T> cx(n), cy(n), cz(n);
            com<T>(n, X.data(), Y.data(), Z.data(), cx.data(), cy.data(), cz.data());
            for (int c : reset) { CMx[c] = cx[c]; CMy[c] = cy[c]; CMz[c] = cz[c]; }
        }
        return st;
    }
    // one pass of the step loop, GPUbounce.cu:1055-1506
    uint32_t step(int k) override {
        uint32_t st = 0;
        using clk = std::chrono::steady_clock;
        auto mark = clk::now();
        auto lap = [&](int q) {
            const auto t = clk::now();
            phase_s[q] += std::chrono::duration<double>(t - mark).count();
            mark = t;
        };
        for (int s = 0; s < k; ++s) {
            ++stepno;
            const bool rnd = p.add_rands != 0;  // Integrate uses the previous step's draw (d_fRanList)
            integrate<T>(n, &p, X.data(), Y.data(), Z.data(), VX.data(), VY.data(), VZ.data(), FCX.data(), FCY.data(),
                         FCZ.data(), FDX.data(), FDY.data(), FDZ.data(), rnd ? FRX.data() : nullptr,
                         rnd ? FRY.data() : nullptr, rnd ? FRZ.data() : nullptr);
            VAX = VX; VAY = VY; VAZ = VZ;
            T rg = rgrowth<T>(&p, n);
            pressure_update<T>(n, &p, P.data(), E.data(), gscale.data(), stepno, rg);
            lap(0);
            st |= conforce();
            lap(1);
            st |= disforce();
            lap(2);
            if (rnd) {  // CalculateRanForce, GPUbounce.cu:1133-1137
                for (int c = 0; c < n; ++c)
                    for (int i = 0; i < NN; ++i) {
                        float zr[3];
                        normals3(p.noise_seed, (uint32_t)(c * S + i), stepno, zr);
                        const int64_t gi = (int64_t)c * S + i;
                        FRX[gi] = (T)(p.rand_scale_factor * zr[0]);
                        FRY[gi] = (T)(p.rand_scale_factor * zr[1]);
                        FRZ[gi] = (T)(p.rand_scale_factor * zr[2]);
                    }
            }
            velocity_ab<T>(n, &p, VX.data(), VY.data(), VZ.data(), FCX.data(), FCY.data(), FCZ.data(), FDX.data(),
                           FDY.data(), FDZ.data(), rnd ? FRX.data() : nullptr, rnd ? FRY.data() : nullptr,
                           rnd ? FRZ.data() : nullptr);
            lap(3);
            st |= disforce();
            lap(2);
            com<T>(n, X.data(), Y.data(), Z.data(), CMx.data(), CMy.data(), CMz.data());
            bool divided = false;
            if (stepno <= p.div_time_steps && rg > T(0)) {
                st |= volumes<T>(n, &p, m.faces, X.data(), Y.data(), Z.data(), CMx.data(), CMy.data(), CMz.data(),
                                 E.data(), vol.data(), area.data(), flag.data());
                st |= divide(divided);
            }
            lap(4);
            int gi = p.grid_interval > 0 ? p.grid_interval : 1;
            if (divided || stepno % gi == 0) st |= rebuild_grid();
            if (p.correct_com) {
                // calc_sys_CM (GPUbounce.cu:2504-2520): sums in cell order, then one
                // division by the count; CorrectCoMMotion (propagate.cu:766-776) on the
                // mesh nodes (GPUbounce.cu:1431-1441)
                T c3[3] = {T(0), T(0), T(0)};
                for (int c = 0; c < n; ++c) { c3[0] += CMx[c]; c3[1] += CMy[c]; c3[2] += CMz[c]; }
                for (int a = 0; a < 3; ++a) c3[a] = c3[a] / (T)n;
                for (int c = 0; c < n; ++c)
                    for (int i = 0; i < NN; ++i) {
                        const int64_t g = (int64_t)c * S + i;
                        X[g] = X[g] - c3[0]; Y[g] = Y[g] - c3[1]; Z[g] = Z[g] - c3[2];
                    }
                for (int a = 0; a < 3; ++a) sys_cm[a] = (float)c3[a];
            }
            lap(5);
            status |= st;
            if (st & CS3D_SYS_FATAL_MASK) break;
        }
        return st;
    }
    const std::vector<T>* node_field(int f) const {
        switch (f) {
            case CS3D_F_X: return &X; case CS3D_F_Y: return &Y; case CS3D_F_Z: return &Z;
            case CS3D_F_VX: return &VX; case CS3D_F_VY: return &VY; case CS3D_F_VZ: return &VZ;
            case CS3D_F_FCX: return &FCX; case CS3D_F_FCY: return &FCY; case CS3D_F_FCZ: return &FCZ;
            case CS3D_F_FDX: return &FDX; case CS3D_F_FDY: return &FDY; case CS3D_F_FDZ: return &FDZ;
            case CS3D_F_VAX: return &VAX; case CS3D_F_VAY: return &VAY; case CS3D_F_VAZ: return &VAZ;
            case CS3D_F_CMX: return &CMx; case CS3D_F_CMY: return &CMy; case CS3D_F_CMZ: return &CMz;
            case CS3D_F_BBOX: return &bb; case CS3D_F_PRESSURE: return &P; case CS3D_F_STIFFNESS: return &E;
            case CS3D_F_VOLUME: return &vol; case CS3D_F_AREA: return &area;
            case CS3D_F_GROWTH_SCALE: return &gscale;
            case CS3D_F_FRX: return &FRX; case CS3D_F_FRY: return &FRY; case CS3D_F_FRZ: return &FRZ;
            default: return nullptr;
        }
    }
    size_t get(int f, void* out, size_t bytes) const override {
        if (const std::vector<T>* v = node_field(f)) {
            size_t cnt = std::min(bytes / sizeof(float), v->size());
            float* o = (float*)out;
            for (size_t i = 0; i < cnt; ++i) o[i] = (float)(*v)[i];
            return cnt * sizeof(float);
        }
        if (f == CS3D_F_SYSCM) {
            size_t cnt = std::min(bytes, sizeof(sys_cm));
            std::memcpy(out, sys_cm, cnt);
            return cnt;
        }
        if (f == CS3D_F_DIVFLAG) {
            size_t cnt = std::min(bytes, flag.size());
            std::memcpy(out, flag.data(), cnt);
            return cnt;
        }
        if (f == CS3D_F_GRID) {
            int32_t g[8];
            float gm[3] = {(float)gmin[0], (float)gmin[1], (float)gmin[2]};
            std::memcpy(g, gm, 12);
            g[3] = dims[0]; g[4] = dims[1]; g[5] = dims[2]; g[6] = budget; g[7] = p.voxel_capacity;
            size_t cnt = std::min(bytes, sizeof(g));
            std::memcpy(out, g, cnt);
            return cnt;
        }
        if (f == CS3D_F_VOXCOUNT || f == CS3D_F_VOXLIST) {
            const std::vector<int32_t>& v = f == CS3D_F_VOXCOUNT ? counts : lists;
            size_t cnt = std::min(bytes, v.size() * 4);
            std::memcpy(out, v.data(), cnt);
            return cnt;
        }
        if (f == CS3D_F_RANMAR) {
            size_t cnt = std::min(bytes, sizeof(rng));
            std::memcpy(out, &rng, cnt);
            return cnt;
        }
        return 0;
    }
    size_t set(int f, const void* in, size_t bytes) override {
        if (std::vector<T>* v = const_cast<std::vector<T>*>(node_field(f))) {
            size_t cnt = std::min(bytes / sizeof(float), v->size());
            const float* s = (const float*)in;
            for (size_t i = 0; i < cnt; ++i) (*v)[i] = (T)s[i];
            return cnt * sizeof(float);
        }
        if (f == CS3D_F_DIVFLAG) {
            size_t cnt = std::min(bytes, flag.size());
            std::memcpy(flag.data(), in, cnt);
            return cnt;
        }
        if (f == CS3D_F_RANMAR) {
            size_t cnt = std::min(bytes, sizeof(rng));
            std::memcpy(&rng, in, cnt);
            return cnt;
        }
        return 0;
    }
    size_t get64(int f, double* out, size_t count) const override {
        const std::vector<T>* v = node_field(f);
        if (!v) return 0;
        size_t cnt = std::min(count, v->size());
        for (size_t i = 0; i < cnt; ++i) out[i] = (double)(*v)[i];
        return cnt;
    }
};

}  // namespace

struct or_sim {
    SimBase* impl;
};

extern "C" or_sim* or_sim_create(const cs3d_params* p, const cs3d_mesh* m, int n, int max_cells, int32_t budget,
                                 const float* X, const float* Y, const float* Z, int precision) {
    if (!p || !m || n < 0 || max_cells < n || budget <= 0) return nullptr;
    or_sim* s = new or_sim;
    if (precision == 1) s->impl = new Sim<double>(*p, *m, n, max_cells, budget, X, Y, Z);
    else s->impl = new Sim<float>(*p, *m, n, max_cells, budget, X, Y, Z);
    return s;
}
extern "C" void or_sim_destroy(or_sim* s) {
    if (!s) return;
    delete s->impl;
    delete s;
}
extern "C" uint32_t or_sim_init(or_sim* s) { return s->impl->init(); }
extern "C" uint32_t or_sim_step(or_sim* s, int k) { return s->impl->step(k); }
extern "C" void or_sim_phase_seconds(const or_sim* s, double* out6) {
    for (int q = 0; q < 6; ++q) out6[q] = s->impl->phase_s[q];
}
extern "C" int or_sim_ncells(const or_sim* s) { return s->impl->n; }
extern "C" int64_t or_sim_stepno(const or_sim* s) { return s->impl->stepno; }
extern "C" size_t or_sim_get(const or_sim* s, int f, void* out, size_t bytes) { return s->impl->get(f, out, bytes); }
extern "C" size_t or_sim_set(or_sim* s, int f, const void* in, size_t bytes) { return s->impl->set(f, in, bytes); }
extern "C" size_t or_sim_get_f64(const or_sim* s, int f, double* out, size_t count) {
    return s->impl->get64(f, out, count);
}
extern "C" void or_sim_set_counts(or_sim* s, int n, int64_t step) {
    s->impl->n = n;
    s->impl->stepno = step;
}
extern "C" void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
extern "C" int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
