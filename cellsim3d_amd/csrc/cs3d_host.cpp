// cs3d_host.cpp - host-side parts of libcs3d that need no GPU: status strings,
// parameter defaults and validation, mesh rest state, host RANMAR and the
// binary trajectory writer.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/cs3d.h"
#include "cs3d_dev.h"
#include "cs3d_internal.h"

namespace cs3d {
thread_local std::string g_last_error;

cs3d_status fail(cs3d_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}
}  // namespace cs3d

using cs3d::fail;

extern "C" const char* cs3d_status_str(cs3d_status s) {
    switch (s) {
        case CS3D_OK: return "ok";
        case CS3D_INVALID_ARG: return "invalid argument";
        case CS3D_OUT_OF_MEMORY: return "out of memory";
        case CS3D_CAPACITY: return "capacity exceeded";
        case CS3D_NAN_DETECTED: return "non-finite state detected";
        case CS3D_LIST_OVERFLOW: return "list overflow";
        case CS3D_HIP_ERROR: return "HIP runtime error";
        case CS3D_NO_DEVICE: return "no usable GPU";
        case CS3D_UNSUPPORTED: return "unsupported option";
    }
    return "unknown status";
}

extern "C" const char* cs3d_last_error(void) { return cs3d::g_last_error.c_str(); }
extern "C" const char* cs3d_version(void) { return "cellsim3d-mi355x 0.1 (gfx950)"; }

// Defaults of the reference's inp.json (inp.json:1-108) plus the constants the
// reference hard-codes (angle k 1000 propagate.cu:335, DL 2.9 GPUbounce.cu:756,
// 32 cells per voxel propagatebound.cu:259).
extern "C" cs3d_status cs3d_params_default(cs3d_params* p) {
    if (!p) return fail(CS3D_INVALID_ARG, "params is NULL");
    std::memset(p, 0, sizeof(*p));
    p->mass = 0.04f;
    p->repulsion_range = 0.2f;
    p->attraction_range = 0.3f;
    p->repulsion_strength = 100.0f;
    p->attraction_strength = 0.5f;
    p->youngs_mod = 1000.0f;
    p->stiffness1 = 1.0f * 1000.0f;
    p->stiffness2 = 0.8f * 1000.0f;
    p->viscotic_damping = 1.0f;
    p->internal_damping = 100.0f;
    p->gamma_visc = 1.0f;
    p->division_vol = 2.9f;
    p->dt = 0.0001f;
    p->max_pressure = 65.0f;
    p->min_pressure = 50.0f;
    p->growth_rate = 0.002f;
    p->check_sphericity = 1;
    p->constrain_angles = 1;
    p->div_time_steps = 50000;
    p->non_div_time_steps = 50000;
    p->phase_count = 2000000000;
    p->use_rigid_box = 1;
    p->box_max[0] = 500.0f;
    p->box_max[1] = 500.0f;
    p->box_max[2] = 2.0f;
    p->thresh_dist = 0.01f;
    p->daught_same_stiffness = 1;
    p->div_plane_basis[2] = 1.0f;
    p->ranmar_ij = 1;
    p->ranmar_kl = 100;
    p->angle_k = 1000.0f;
    p->grid_dl = 2.9f;
    p->voxel_capacity = 32;
    p->grid_interval = 1;
    p->pair_skin = 0.05f;
    p->rand_scale_factor = 0.001f;
    p->noise_seed = 1;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_params_check(const cs3d_params* p) {
    if (!p) return fail(CS3D_INVALID_ARG, "params is NULL");
    if (p->add_rands && !(p->rand_scale_factor >= 0)) return fail(CS3D_INVALID_ARG, "rand_scale_factor must be >= 0");
    if (!(p->mass > 0)) return fail(CS3D_INVALID_ARG, "particle_mass must be > 0");
    if (!(p->dt > 0)) return fail(CS3D_INVALID_ARG, "time_interval must be > 0");
    if (!(p->grid_dl > 0)) return fail(CS3D_INVALID_ARG, "grid_dl must be > 0");
    if (!(p->attraction_range >= 0) || !(p->repulsion_range >= 0))
        return fail(CS3D_INVALID_ARG, "interaction ranges must be >= 0");
    if (p->voxel_capacity < 1 || p->voxel_capacity > 1024)
        return fail(CS3D_INVALID_ARG, "voxel_capacity must be in [1, 1024]");
    if (p->grid_interval < 1) return fail(CS3D_INVALID_ARG, "grid_interval must be >= 1");
    if (!(p->pair_skin >= 0) || p->pair_skin > 10) return fail(CS3D_INVALID_ARG, "pair_skin must be in [0, 10]");
    if (p->ranmar_ij < 0 || p->ranmar_ij > 31328 || p->ranmar_kl < 0 || p->ranmar_kl > 30081)
        return fail(CS3D_INVALID_ARG, "RANMAR seeds out of range (marsaglia.h:79-83)");
    return CS3D_OK;
}

// R0 (GPUbounce.cu:1892-1908) and theta0 (GPUbounce.cu:912-933, float acos).
extern "C" cs3d_status cs3d_mesh_build(const int32_t* nn, const int32_t* sign, const int32_t* faces, const float* x,
                                       const float* y, const float* z, cs3d_mesh* out) {
    if (!nn || !sign || !faces || !x || !y || !z || !out) return fail(CS3D_INVALID_ARG, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    for (int i = 0; i < 3 * CS3D_SLOTS; ++i) out->nn[i] = nn[i];
    for (int i = 0; i < CS3D_NODES; ++i) out->sign[i] = sign[i];
    for (int i = 0; i < CS3D_FACES * 7; ++i) out->faces[i] = faces[i];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < CS3D_NODES; ++i) {
            const int n = nn[k * CS3D_SLOTS + i];
            if (n < 0 || n >= CS3D_NODES) return fail(CS3D_INVALID_ARG, "neighbour index out of range");
        }
    for (int i = 0; i < CS3D_FACES * 7; ++i)
        if (faces[i] < 0 || faces[i] >= CS3D_NODES) return fail(CS3D_INVALID_ARG, "face index out of range");
    for (int i = 0; i < CS3D_NODES; ++i) {
        for (int k = 0; k < 3; ++k) {
            const int n = nn[k * CS3D_SLOTS + i];
            const float dx = x[i] - x[n], dy = y[i] - y[n], dz = z[i] - z[n];
            out->R0[k * CS3D_SLOTS + i] = sqrtf(dx * dx + dy * dy + dz * dz);
        }
        const int a = nn[i], b = nn[CS3D_SLOTS + i], c = nn[2 * CS3D_SLOTS + i];
        const float ix = x[a] - x[i], iy = y[a] - y[i], iz = z[a] - z[i];
        const float jx = x[b] - x[i], jy = y[b] - y[i], jz = z[b] - z[i];
        const float kx = x[c] - x[i], ky = y[c] - y[i], kz = z[c] - z[i];
        const float mi = sqrtf(ix * ix + iy * iy + iz * iz);
        const float mj = sqrtf(jx * jx + jy * jy + jz * jz);
        const float mk = sqrtf(kx * kx + ky * ky + kz * kz);
        out->theta0[3 * i + 0] = acosf((ix * jx + iy * jy + iz * jz) / (mi * mj));
        out->theta0[3 * i + 1] = acosf((jx * kx + jy * ky + jz * kz) / (mj * mk));
        out->theta0[3 * i + 2] = acosf((ix * kx + iy * ky + iz * kz) / (mi * mk));
    }
    return CS3D_OK;
}

// rmarin, marsaglia.h:54-113
// The device CenterOfMass tree: slots t and t+128 (t < 128, slots >= 180 count as
// zero), then t and t+64, then pairs at distance 32, 16, ..., 1; one division by 180.
extern "C" cs3d_status cs3d_cell_centres(const float* x, const float* y, const float* z, int32_t n_cells,
                                         float* cmx, float* cmy, float* cmz) {
    if (!x || !y || !z || !cmx || !cmy || !cmz || n_cells < 0) return fail(CS3D_INVALID_ARG, "bad cell_centres args");
    const float* in[3] = {x, y, z};
    float* out[3] = {cmx, cmy, cmz};
    for (int32_t c = 0; c < n_cells; ++c) {
        for (int a = 0; a < 3; ++a) {
            const float* v = in[a] + (int64_t)c * CS3D_SLOTS;
            float t[128];
            for (int i = 0; i < 128; ++i) t[i] = v[i] + (i + 128 < CS3D_NODES ? v[i + 128] : 0.0f);
            for (int i = 0; i < 64; ++i) t[i] = t[i] + t[i + 64];
            for (int o = 32; o > 0; o >>= 1)
                for (int i = 0; i < o; ++i) t[i] = t[i] + t[i + o];
            out[a][c] = t[0] / 180.0f;
        }
    }
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_ranmar_init(cs3d_ranmar_state* st, int32_t ij, int32_t kl) {
    if (!st) return fail(CS3D_INVALID_ARG, "state is NULL");
    if (ij < 0 || ij > 31328 || kl < 0 || kl > 30081) return fail(CS3D_INVALID_ARG, "RANMAR seeds out of range");
    int i = (ij / 177) % 177 + 2, j = ij % 177 + 2, k = (kl / 169) % 178 + 1, l = kl % 169;
    std::memset(st, 0, sizeof(*st));
    for (int ii = 1; ii <= 97; ii++) {
        float s = 0.0f, t = 0.5f;
        for (int jj = 1; jj <= 24; jj++) {
            const int m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t *= 0.5f;
        }
        st->u[ii] = s;
    }
    st->c = (float)(362436.0 / 16777216.0);
    st->cd = (float)(7654321.0 / 16777216.0);
    st->cm = (float)(16777213.0 / 16777216.0);
    st->i97 = 97;
    st->j97 = 33;
    st->ready = 1;
    return CS3D_OK;
}

// ranmar, marsaglia.h:115-144
extern "C" cs3d_status cs3d_ranmar(cs3d_ranmar_state* st, float* out, int32_t n) {
    if (!st || !out || n < 0) return fail(CS3D_INVALID_ARG, "bad argument");
    if (!st->ready) return fail(CS3D_INVALID_ARG, "RANMAR state not initialised");
    for (int v = 0; v < n; ++v) {
        float uni = st->u[st->i97] - st->u[st->j97];
        if (uni < 0.0f) uni += 1.0f;
        st->u[st->i97] = uni;
        if (--st->i97 == 0) st->i97 = 97;
        if (--st->j97 == 0) st->j97 = 97;
        st->c -= st->cd;
        if (st->c < 0.0f) st->c += st->cm;
        uni -= st->c;
        if (uni < 0.0f) uni += 1.0f;
        out[v] = uni;
    }
    return CS3D_OK;
}

// ---- binary trajectory, GPUbounce.cu:1007-1018 (header), 2446-2477 (frames),
// 1508-1511 (final cell count patched into the header)
struct cs3d_traj {
    FILE* f;
    int32_t variable_stiffness;
};

extern "C" cs3d_status cs3d_traj_open(const char* path, int32_t max_cells, int32_t variable_stiffness,
                                      int32_t max_frames, cs3d_traj** out) {
    if (!path || !out) return fail(CS3D_INVALID_ARG, "NULL argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CS3D_INVALID_ARG, std::string("cannot open ") + path);
    const int32_t hdr[3] = {max_cells, variable_stiffness ? 1 : 0, max_frames};
    std::fwrite(hdr, sizeof(int32_t), 3, f);
    *out = new cs3d_traj{f, variable_stiffness ? 1 : 0};
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_traj_append(cs3d_traj* t, int32_t step, int32_t frame, int32_t n_cells, const float* x,
                                        const float* y, const float* z, const int32_t* types) {
    if (!t || !t->f || n_cells < 0 || (n_cells > 0 && (!x || !y || !z)))
        return fail(CS3D_INVALID_ARG, "bad trajectory append");
    const int32_t h[3] = {step, frame, n_cells};
    std::fwrite(h, sizeof(int32_t), 3, t->f);
    for (int32_t c = 0; c < n_cells; ++c) {
        std::fwrite(&c, sizeof(int32_t), 1, t->f);
        std::fwrite(x + (size_t)c * CS3D_SLOTS, sizeof(float), CS3D_SLOTS, t->f);
        std::fwrite(y + (size_t)c * CS3D_SLOTS, sizeof(float), CS3D_SLOTS, t->f);
        std::fwrite(z + (size_t)c * CS3D_SLOTS, sizeof(float), CS3D_SLOTS, t->f);
        if (t->variable_stiffness) {
            const int32_t ty = types ? types[c] : 0;
            std::fwrite(&ty, sizeof(int32_t), 1, t->f);
        }
    }
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_traj_close(cs3d_traj* t, int32_t final_cells) {
    if (!t) return fail(CS3D_INVALID_ARG, "NULL trajectory");
    cs3d_status st = CS3D_OK;
    if (t->f) {
        if (std::fseek(t->f, 0, SEEK_SET) == 0) std::fwrite(&final_cells, sizeof(int32_t), 1, t->f);
        else st = fail(CS3D_INVALID_ARG, "cannot rewrite trajectory header");
        std::fclose(t->f);
    }
    delete t;
    return st;
}

// ---- exact products for the velocity kicks' divisions by the particle mass ----
// The reference divides in double: 0.5 * (double)s / (double)m with s a float
// (Integrate, VelocityUpdateA, propagate.cu:698-705, 721-723) and
// 0.5 * (double)dt * (double)f / (double)m with f a float (VelocityUpdateB,
// :732-734).  Both numerators carry at most a float's 24-bit significand (0.5 * s
// is exact; 0.5 * dt * f is an exact product of two floats), so each quotient is
// a float times a fixed real c (1 / (2 m), or dt / (2 m)).  With c = ch + cl as a
// double-double (ch = RN(c), cl = RN(c - ch)), fma(x, ch, x * cl) approximates
// x * c to far below half an ulp of double and rounds to the same double as the
// division except where x * c lies extremely close to a midpoint of two doubles;
// scaling x by a power of two scales everything exactly (no double overflow or
// underflow for float x), so whether that ever happens depends only on the 2^23
// significands of x, and they are all checked here, once per (m, dt).  Only a
// system whose check passes takes the products (kick_fast); otherwise the device
// divides.  Zeros keep their sign on the device by a select; a non-finite force is
// not covered (such a state is flagged NaN).
namespace cs3d {
namespace {
std::mutex g_kick_mu;
std::map<std::pair<uint32_t, uint32_t>, bool> g_kick_ok;

bool kick_check(double m, double dt, double ah, double al, double bh, double bl) {
    const double hdt = 0.5 * dt;
    for (uint32_t M = 0; M < (1u << 23); ++M) {
        const uint32_t bits = 0x3f800000u | M;
        float f;
        std::memcpy(&f, &bits, 4);
        const double x = (double)f;
        if (std::fma(x, ah, x * al) != 0.5 * x / m) return false;
        if (std::fma(x, bh, x * bl) != hdt * x / m) return false;
    }
    return true;
}
}  // namespace

void kick_consts(float m, float dt, SysMeta& M) {
    const double md = (double)m, dd = (double)dt;
    const double m2 = 2.0 * md;
    M.kick_ah = 1.0 / m2;
    M.kick_al = std::fma(-M.kick_ah, m2, 1.0) / m2;
    M.kick_bh = 0.5 * dd / md;
    M.kick_bl = std::fma(-M.kick_bh, md, 0.5 * dd) / md;
    M.kick_fast = 0;
    if (!(md > 0.0) || !(dd > 0.0) || !std::isfinite(M.kick_ah) || !std::isfinite(M.kick_bh)) return;
    uint32_t km, kd;
    std::memcpy(&km, &m, 4);
    std::memcpy(&kd, &dt, 4);
    std::lock_guard<std::mutex> lock(g_kick_mu);
    auto it = g_kick_ok.find({km, kd});
    if (it == g_kick_ok.end())
        it = g_kick_ok.emplace(std::make_pair(km, kd), kick_check(md, dd, M.kick_ah, M.kick_al, M.kick_bh,
                                                                  M.kick_bl)).first;
    M.kick_fast = it->second ? 1 : 0;
}
}  // namespace cs3d
