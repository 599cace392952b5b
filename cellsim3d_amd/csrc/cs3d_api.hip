// cs3d_api.hip - C ABI of libcs3d: context and pool management, systems,
// stepping (eager or HIP-graph replay), per-kernel entry points and field I/O.
// Replaces the host driver's device-side orchestration in main()
// (src/GPUbounce.cu:358-1506) for one or many systems per GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "cs3d_dev.h"
#include "cs3d_div.h"
#include "cs3d_internal.h"

namespace cs3d {
void launch_integrate(const Pool& p, const Range& r, int flags, int64_t step_override, hipStream_t st);
void launch_search(const Pool& p, const Range& r, int always, hipStream_t st);
void launch_con(const Pool& p, const Range& r, int mode, bool use_va, hipStream_t st);
void launch_dis(const Pool& p, const Range& r, int mode, int64_t step_override, int clear_bins, bool use_va,
                hipStream_t st);
void launch_bonded(const Pool& p, const Range& r, int angles, hipStream_t st);
void launch_divide(const Pool& p, const Range& r, int mode, hipStream_t st);
extern thread_local std::vector<std::pair<hipEvent_t, hipEvent_t>>* kernel_timer;
void launch_com_shift(const Pool& p, const Range& r, hipStream_t st);
void launch_live(const Pool& p, const Range& r, hipStream_t st);
void launch_set_nsys(const Pool& p, int n, hipStream_t st);
void launch_grid(const Pool& p, const Range& r, int vox_blocks, hipStream_t st);
void launch_vox_lists(const Pool& p, const Range& r, int vox_blocks, hipStream_t st);
void launch_vrange(const Pool& p, const Range& r, hipStream_t st);
void launch_velocity_ab(const Pool& p, const Range& r, int which, hipStream_t st);
void launch_cellbox(const Pool& p, const Range& r, hipStream_t st);
void launch_extent(const Pool& p, const Range& r, float* out, hipStream_t st);
void launch_pressure_reset(const Pool& p, const Range& r, const int32_t* cells, int count, float value,
                           hipStream_t st);
void launch_set_flags(const Pool& p, const Range& r, int rebuild, int search, hipStream_t st);
void launch_copy_calibrate(const void* src, void* dst, int64_t n4, hipStream_t st);
void launch_check_sqrt(unsigned long long* bad, hipStream_t st);
void launch_check_div(unsigned long long* bad, uint32_t a_count, uint64_t scaled_pairs, hipStream_t st);
void launch_acos(const double* x, double* y, int64_t n, hipStream_t st);
}  // namespace cs3d

using namespace cs3d;

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(CS3D_HIP_ERROR, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct Block {
    int64_t off, len;
};

// first-fit allocator over [0, total)
struct RangeAlloc {
    int64_t total = 0;
    std::vector<Block> used;  // sorted by off
    bool alloc(int64_t len, int64_t& off) {
        int64_t cur = 0;
        size_t i = 0;
        for (; i < used.size(); ++i) {
            if (used[i].off - cur >= len) break;
            cur = used[i].off + used[i].len;
        }
        if (i == used.size() && total - cur < len) return false;
        off = cur;
        used.insert(used.begin() + i, Block{off, len});
        return true;
    }
    void release(int64_t off) {
        for (size_t i = 0; i < used.size(); ++i)
            if (used[i].off == off) { used.erase(used.begin() + i); return; }
    }
    int64_t high() const { return used.empty() ? 0 : used.back().off + used.back().len; }
};

struct SysHost {
    bool used = false;
    int32_t base = 0, cap = 0;
    int64_t vreg = 0;  // voxel region offset
    int32_t vbudget = 0, vcap = 0;
    int64_t breg = 0;  // search-bin region offset
    int32_t bbudget = 0;
    cs3d_params prm;
};

}  // namespace

struct cs3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t cap_cells = 0;
    int64_t vox_total = 0;
    int64_t bin_total = 0;
    Pool pool;
    std::vector<void*> allocs;
    RangeAlloc cells, voxels, bins;
    std::vector<SysHost> sys;
    int nsys_high = 0;
    bool use_graph = true;
    int graph_steps = 10;
    struct GraphRec {
        hipGraphExec_t exec;
        int steps;
        bool shift;  // the captured steps include CorrectCoMMotion
    };
    std::map<int, GraphRec> graphs;  // target -> captured steps
    cs3d_mesh mesh_host;  // the tables the context was created with (checkpoint check)
    void* scratch = nullptr;  // small device buffer of the per-kernel entry points (grown on demand)
    size_t scratch_bytes = 0;
};

namespace {

template <class T> cs3d_status dalloc(cs3d_ctx* c, T*& p, int64_t count) {
    void* q = nullptr;
    const size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(CS3D_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(q, 0, bytes);
    if (e != hipSuccess) return fail(CS3D_HIP_ERROR, std::string("hipMemset: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    p = (T*)q;
    return CS3D_OK;
}

// Drops the captured steps of one target (-2: all).  The all-systems capture
// reads the systems in use from the device, so adding or removing a system
// leaves it valid.
void drop_graphs(cs3d_ctx* c, int target_only = -2) {
    for (auto it = c->graphs.begin(); it != c->graphs.end();) {
        if (target_only == -2 || it->first == target_only) {
            (void)hipGraphExecDestroy(it->second.exec);
            it = c->graphs.erase(it);
        } else {
            ++it;
        }
    }
}

bool valid_sys(cs3d_ctx* c, cs3d_sys s) { return c && s >= 0 && s < (int)c->sys.size() && c->sys[s].used; }

Range range_of(cs3d_ctx* c, cs3d_sys s) {
    Range r;
    if (s == CS3D_ALL) {
        // the whole pool and every system id: the kernels read the systems in use
        // from the device, so a capture of this range survives adds and removes
        r.slot_lo = 0;
        r.slot_hi = (int32_t)c->cap_cells;
        r.sys_lo = 0;
        r.sys_hi = kMaxSys;
        r.one_sys = 0;
        r.live = 0;
    } else {
        r.slot_lo = c->sys[s].base;
        r.slot_hi = c->sys[s].base + c->sys[s].cap;
        r.sys_lo = s;
        r.sys_hi = s + 1;
        r.one_sys = 1;
        r.live = 0;
    }
    return r;
}

int vox_blocks_of(cs3d_ctx* c, cs3d_sys s) {
    int64_t vb = 0;
    if (s == CS3D_ALL) {
        for (auto& h : c->sys)
            if (h.used) vb = std::max<int64_t>(vb, h.vbudget);
    } else {
        vb = c->sys[s].vbudget;
    }
    int64_t b = vb / 4096;
    return (int)std::min<int64_t>(64, std::max<int64_t>(1, b));
}

cs3d_status read_meta(cs3d_ctx* c, cs3d_sys s, SysMeta& m) {
    HIPCHK(hipMemcpyAsync(&m, c->pool.meta + s, sizeof(SysMeta), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

cs3d_status write_meta(cs3d_ctx* c, cs3d_sys s, const SysMeta& m) {
    HIPCHK(hipMemcpyAsync(c->pool.meta + s, &m, sizeof(SysMeta), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

bool any_correct_com(const cs3d_ctx* c, const Range& r) {
    for (int s = r.sys_lo; s < r.sys_hi; ++s)
        if (c->sys[s].used && c->sys[s].prm.correct_com) return true;
    return false;
}

// One step of the loop in GPUbounce.cu:1055-1506 as five launches:
//   integrate  Integrate + ForwardTime + PressureUpdate; per-cell part of a grid
//              rebuilt at the end of the last step; search bins (search steps)
//   search     contact candidates (steps flagged by the previous step)
//   con_dis    CalculateConForce + CalculateDisForce (post-Integrate velocities) +
//              VelocityUpdateA/B
//   dis_post   CalculateDisForce (end velocities) + CenterOfMass [+ volumes] +
//              bounding_boxes + next-step search decision
//   divide     division phase, step bookkeeping, extent and grid geometry
//              [+ CorrectCoMMotion, a sixth launch, for systems with correct_com]
//
// A step of several systems first lists their live cell slots (k_live); the four
// cell kernels of the step walk that list.  The division phase adds daughters after
// them, which the next step's list holds.
void enqueue_step_phase(cs3d_ctx* c, const Range& r, int phase) {
    const Pool& p = c->pool;
    Range rl = r;
    rl.live = r.one_sys ? 0 : 1;
    switch (phase) {
        case CS3D_PH_INTEGRATE:
            if (rl.live) launch_live(p, r, c->stream);
            launch_integrate(p, rl, I_KICK | I_PRESSURE | I_BINS | I_IFSEARCH | I_GRIDCOUNT, -1, c->stream);
            break;
        case CS3D_PH_SEARCH: launch_search(p, rl, 0, c->stream); break;
        case CS3D_PH_CON_DIS: launch_con(p, rl, M_STEP, true, c->stream); break;
        case CS3D_PH_DIS_POST: launch_dis(p, rl, M_STEP, -1, 1, false, c->stream); break;
        case CS3D_PH_DIVIDE:
            launch_divide(p, r, M_STEP, c->stream);
            if (any_correct_com(c, r)) launch_com_shift(p, r, c->stream);
            break;
        default: break;
    }
}

void enqueue_step(cs3d_ctx* c, const Range& r, int /*vb*/) {
    for (int q = 0; q < CS3D_PHASES; ++q) enqueue_step_phase(c, r, q);
}

cs3d_status run_startup(cs3d_ctx* c, cs3d_sys s) {
    const Pool& p = c->pool;
    Range r = range_of(c, s);
    const int vb = vox_blocks_of(c, s);
    // CenterOfMass, bounding_boxes, minmaxpre/post, makeNNlist (GPUbounce.cu:739-763)
    launch_dis(p, r, M_COM, 0, 0, false, c->stream);
    launch_set_flags(p, r, 1, -1, c->stream);
    launch_grid(p, r, vb, c->stream);
    launch_set_flags(p, r, 0, -1, c->stream);
    // initial CalculateConForce and CalculateDisForce (GPUbounce.cu:978-1004)
    launch_integrate(p, r, I_BINS, 0, c->stream);
    launch_search(p, r, 1, c->stream);
    launch_con(p, r, M_INIT, false, c->stream);
    launch_dis(p, r, M_DIS_V, 0, 1, false, c->stream);
    launch_set_flags(p, r, -1, 0, c->stream);  // candidate lists are current
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

struct FieldLoc {
    void* ptr = nullptr;
    size_t bytes = 0;
};

FieldLoc field_loc(cs3d_ctx* c, cs3d_sys s, cs3d_field f) {
    FieldLoc L;
    const SysHost& h = c->sys[s];
    const Pool& p = c->pool;
    const int64_t nb = (int64_t)h.base * kSlots;
    const size_t nbytes = (size_t)h.cap * kSlots * sizeof(float);
    float* node[] = {p.X, p.Y, p.Z, p.VX, p.VY, p.VZ, p.FCX, p.FCY, p.FCZ, p.FDX, p.FDY, p.FDZ, p.VAX, p.VAY, p.VAZ};
    if (f >= CS3D_F_X && f <= CS3D_F_VAZ) {
        L.ptr = node[f] + nb;
        L.bytes = nbytes;
        return L;
    }
    switch (f) {
        case CS3D_F_FRX: L.ptr = p.FRX + nb; L.bytes = nbytes; break;
        case CS3D_F_FRY: L.ptr = p.FRY + nb; L.bytes = nbytes; break;
        case CS3D_F_FRZ: L.ptr = p.FRZ + nb; L.bytes = nbytes; break;
        case CS3D_F_CMX: L.ptr = p.CMX + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_CMY: L.ptr = p.CMY + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_CMZ: L.ptr = p.CMZ + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_BBOX: L.ptr = p.BBOX + (int64_t)h.base * 6; L.bytes = (size_t)h.cap * 24; break;
        case CS3D_F_PRESSURE: L.ptr = p.P + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_STIFFNESS: L.ptr = p.E + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_VOLUME: L.ptr = p.VOL + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_AREA: L.ptr = p.AREA + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_DIVFLAG: L.ptr = p.FLAG + h.base; L.bytes = h.cap; break;
        case CS3D_F_VOXCOUNT: L.ptr = p.vox + h.vreg; L.bytes = (size_t)h.vbudget * 4; break;
        case CS3D_F_VOXLIST: L.ptr = p.vox + h.vreg + h.vbudget; L.bytes = (size_t)h.vbudget * h.vcap * 4; break;
        case CS3D_F_PAIRCOUNT: L.ptr = p.NPAIR + nb; L.bytes = nbytes; break;
        case CS3D_F_CANDCOUNT: L.ptr = p.NCAND + nb; L.bytes = nbytes; break;
        case CS3D_F_PAIRS: L.ptr = p.PAIRS + (int64_t)h.base * kPMax * kSlots; L.bytes = (size_t)h.cap * kPMax * kSlots * 4; break;
        case CS3D_F_GROWTH_SCALE: L.ptr = p.GS + h.base; L.bytes = h.cap * 4; break;
        case CS3D_F_GRID: L.bytes = 32; break;
        case CS3D_F_RANMAR: L.bytes = sizeof(cs3d_ranmar_state); break;
        case CS3D_F_SYSCM: L.bytes = 12; break;
        default: break;
    }
    return L;
}

}  // namespace

// ---- staging array of the allocation-free PressureReset -------------------------
// cs3d_k_pressure_reset stages the caller's cell list in the context's scratch
// buffer, which it allocates (hipMalloc) on first use and when a longer list
// comes.  cs3d_k_pressure_reset_noalloc stages it in `scratch` instead: a
// file-scope device array of kNoAllocCells entries, which every GPU the library
// runs on holds until the process ends.  Its device address differs per GPU; it is
// looked up once per device, with that device current, at context creation
// (hipGetSymbolAddress) and kept in g_scratch_dev.  Host code passes only that
// address to the device, never the array's host-side name.
constexpr int32_t kNoAllocCells = 65536;
static __device__ __align__(256) int32_t scratch[kNoAllocCells];

namespace {
std::mutex g_scratch_mu;
std::map<int, int32_t*> g_scratch_dev;  // device -> address of `scratch` on it

// with `device` current
cs3d_status lookup_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev.count(device)) return CS3D_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CS3D_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<int32_t*>(ptr);
    return CS3D_OK;
}

int32_t* scratch_of_device(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_dev.find(device);
    return it == g_scratch_dev.end() ? nullptr : it->second;
}
}  // namespace

extern "C" cs3d_status cs3d_ctx_create(int device, const cs3d_mesh* mesh, int64_t capacity_cells, int64_t voxel_budget,
                                       int64_t bin_budget, cs3d_ctx** out) {
    if (!mesh || !out || capacity_cells <= 0 || capacity_cells > (1ll << 30) / kSlots || voxel_budget <= 0 ||
        bin_budget <= 0)
        return fail(CS3D_INVALID_ARG, "bad context arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CS3D_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(CS3D_NO_DEVICE, "device index out of range");
    // angle gather table and mesh consistency (every bond listed from both ends)
    MeshDev md;
    std::memset(&md, 0, sizeof(md));
    std::memcpy(md.nn, mesh->nn, sizeof(md.nn));
    std::memcpy(md.sign, mesh->sign, sizeof(md.sign));
    std::memcpy(md.faces, mesh->faces, sizeof(md.faces));
    std::memcpy(md.R0, mesh->R0, sizeof(md.R0));
    // reciprocals of the rest lengths for the kernels' quotients (IEEE division here;
    // NaN where quot_ok fails sends those bonds to the kernels' division path)
    for (int i = 0; i < 3 * kSlots; ++i)
        md.R0inv[i] = quot_ok(mesh->R0[i]) ? 1.0f / mesh->R0[i] : std::numeric_limits<float>::quiet_NaN();
    std::memcpy(md.theta0, mesh->theta0, sizeof(md.theta0));
    for (int p = 0; p < kNodes; ++p) {
        for (int q = 0; q < 3; ++q) {
            const int n = mesh->nn[q * kSlots + p];
            if (n < 0 || n >= kNodes) return fail(CS3D_INVALID_ARG, "mesh neighbour out of range");
            const int a0 = mesh->nn[n], a1 = mesh->nn[kSlots + n], a2 = mesh->nn[2 * kSlots + n];
            int e0, e1;  // NeighNeighs (propagate.cu:63-78)
            if (a0 == p) { e0 = (n * 3 + 0) * 6 + 0; e1 = (n * 3 + 2) * 6 + 0; }
            else if (a1 == p) { e0 = (n * 3 + 0) * 6 + 3; e1 = (n * 3 + 1) * 6 + 0; }
            else if (a2 == p) { e0 = (n * 3 + 2) * 6 + 3; e1 = (n * 3 + 1) * 6 + 3; }
            else return fail(CS3D_INVALID_ARG, "mesh neighbour lists are not symmetric");
            md.gather[p * 6 + 2 * q] = e0;
            md.gather[p * 6 + 2 * q + 1] = e1;
        }
    }
    for (int i = 0; i < CS3D_FACES * 7; ++i)
        if (mesh->faces[i] < 0 || mesh->faces[i] >= kNodes) return fail(CS3D_INVALID_ARG, "mesh face index out of range");

    cs3d_ctx* c = new cs3d_ctx();
    c->mesh_host = *mesh;
    c->device = device;
    c->cap_cells = capacity_cells;
    c->vox_total = voxel_budget;
    c->bin_total = bin_budget;
    c->cells.total = capacity_cells;
    c->voxels.total = voxel_budget;
    c->bins.total = bin_budget;
    c->sys.resize(kMaxSys);
    auto cleanup = [&](cs3d_status st) {
        cs3d_ctx_destroy(c);
        return st;
    };
    if (hipSetDevice(device) != hipSuccess) return cleanup(fail(CS3D_HIP_ERROR, "hipSetDevice failed"));
    if (cs3d_status ls = lookup_scratch(device); ls != CS3D_OK) return cleanup(ls);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(CS3D_HIP_ERROR, "hipStreamCreate failed"));
    Pool& p = c->pool;
    p.wg_per_cu = 16;
    const int64_t ns = capacity_cells * kSlots;
    cs3d_status st = CS3D_OK;
    float** nodes[] = {&p.X, &p.Y, &p.Z, &p.VX, &p.VY, &p.VZ, &p.VAX, &p.VAY, &p.VAZ,
                       &p.FCX, &p.FCY, &p.FCZ, &p.FDX, &p.FDY, &p.FDZ, &p.FRX, &p.FRY, &p.FRZ};
    for (float** a : nodes)
        if ((st = dalloc(c, *a, ns)) != CS3D_OK) return cleanup(st);
    float** cells[] = {&p.CMX, &p.CMY, &p.CMZ, &p.P, &p.E, &p.VOL, &p.AREA, &p.GS};
    for (float** a : cells)
        if ((st = dalloc(c, *a, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.FLAG, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.BBOX, capacity_cells * 6)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.BBOXN, capacity_cells * 6)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.VRANGE, capacity_cells * 6)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.PAIRS, capacity_cells * kPMax * kSlots)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.NPAIR, capacity_cells * kSlots)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.CAND, capacity_cells * kPMax * kSlots)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.NCAND, capacity_cells * kSlots)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.XBX, ns)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.XBY, ns)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.XBZ, ns)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.SBOX, capacity_cells * 6)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.RESET, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.RSL, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.LIVE, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.n_live, 1)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.bins, bin_budget)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.EVL, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.slot_sys, capacity_cells)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.nsys, 1)) != CS3D_OK) return cleanup(st);
    if ((st = dalloc(c, p.meta, kMaxSys)) != CS3D_OK) return cleanup(st);
    cs3d_params* prm = nullptr;
    if ((st = dalloc(c, prm, kMaxSys)) != CS3D_OK) return cleanup(st);
    p.prm = prm;
    if ((st = dalloc(c, p.vox, voxel_budget)) != CS3D_OK) return cleanup(st);
    MeshDev* dm = nullptr;
    if ((st = dalloc(c, dm, 1)) != CS3D_OK) return cleanup(st);
    p.mesh = dm;
    if (hipMemcpy(dm, &md, sizeof(md), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(CS3D_HIP_ERROR, "mesh upload failed"));
    if (hipMemset(p.slot_sys, 0xFF, capacity_cells * sizeof(int32_t)) != hipSuccess)
        return cleanup(fail(CS3D_HIP_ERROR, "slot table init failed"));
    *out = c;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_ctx_destroy(cs3d_ctx* c) {
    if (!c) return CS3D_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void* q : c->allocs) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_sync(cs3d_ctx* c) {
    if (!c) return fail(CS3D_INVALID_ARG, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipGetLastError());
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_add_raw(cs3d_ctx* c, const cs3d_params* prm, int32_t n_cells, int32_t max_cells,
                                           int64_t voxels, cs3d_sys* out) {
    if (!c || !prm || !out) return fail(CS3D_INVALID_ARG, "NULL argument");
    cs3d_status st = cs3d_params_check(prm);
    if (st != CS3D_OK) return st;
    if (n_cells < 0 || max_cells < 1 || n_cells > max_cells) return fail(CS3D_INVALID_ARG, "bad cell counts");
    int sid = -1;
    for (int i = 0; i < kMaxSys; ++i)
        if (!c->sys[i].used) { sid = i; break; }
    if (sid < 0) return fail(CS3D_CAPACITY, "too many systems in this context");
    int64_t base = 0, vreg = 0;
    if (!c->cells.alloc(max_cells, base)) return fail(CS3D_CAPACITY, "cell pool exhausted");
    const int32_t vcap = prm->voxel_capacity;
    int64_t vb = voxels > 0 ? voxels : std::max<int64_t>(4096, 32 * (int64_t)max_cells);
    if (vb > (1ll << 30)) vb = 1ll << 30;
    if (!c->voxels.alloc(vb * (1 + vcap), vreg)) {
        c->cells.release(base);
        return fail(CS3D_CAPACITY, "voxel pool exhausted");
    }
    // search bins: enough for a tissue at moderate density; coarser bins beyond that
    // (counts, then float4 entries: 1 + 4 * kBCap words per bin; entries 16-byte aligned)
    int64_t bb = std::max<int64_t>(16384, 640 * (int64_t)max_cells);
    bb = std::min<int64_t>(bb, c->bin_total / (1 + 4 * kBCap));
    bb = bb / 4 * 4;
    int64_t breg = 0;
    while (bb >= 4096 && !c->bins.alloc(bb * (1 + 4 * kBCap) + 4, breg)) bb /= 2;
    if (bb < 4096) {
        c->cells.release(base);
        c->voxels.release(vreg);
        return fail(CS3D_CAPACITY, "search-bin pool exhausted");
    }
    SysHost& h = c->sys[sid];
    h.used = true;
    h.base = (int32_t)base;
    h.cap = max_cells;
    h.vreg = vreg;
    h.vbudget = (int32_t)vb;
    h.vcap = vcap;
    h.breg = breg;
    h.bbudget = (int32_t)bb;
    h.prm = *prm;
    c->nsys_high = std::max(c->nsys_high, sid + 1);
    HIPCHK(hipSetDevice(c->device));
    const Pool& p = c->pool;
    const int64_t nb = base * kSlots, nbytes = (int64_t)max_cells * kSlots * 4;
    float* node[] = {p.X,   p.Y,   p.Z,   p.VX,  p.VY,  p.VZ,  p.VAX, p.VAY, p.VAZ,
                     p.FCX, p.FCY, p.FCZ, p.FDX, p.FDY, p.FDZ, p.FRX, p.FRY, p.FRZ};
    for (float* a : node) HIPCHK(hipMemsetAsync(a + nb, 0, nbytes, c->stream));
    float* cellz[] = {p.CMX, p.CMY, p.CMZ, p.VOL, p.AREA};
    for (float* a : cellz) HIPCHK(hipMemsetAsync(a + base, 0, (size_t)max_cells * 4, c->stream));
    HIPCHK(hipMemsetAsync(p.FLAG + base, 0, max_cells, c->stream));
    HIPCHK(hipMemsetAsync(p.BBOX + base * 6, 0, (size_t)max_cells * 24, c->stream));
    HIPCHK(hipMemsetAsync(p.BBOXN + base * 6, 0, (size_t)max_cells * 24, c->stream));
    HIPCHK(hipMemsetAsync(p.NPAIR + nb, 0, (size_t)max_cells * kSlots * 4, c->stream));
    HIPCHK(hipMemsetAsync(p.vox + vreg, 0, (size_t)vb * (1 + vcap) * 4, c->stream));
    HIPCHK(hipMemsetAsync(p.bins + breg, 0, (size_t)bb * 4, c->stream));
    std::vector<float> pv(max_cells, prm->min_pressure), ev(max_cells, prm->stiffness1), gv(max_cells, 1.0f);
    std::vector<int32_t> owner(max_cells, sid);
    HIPCHK(hipMemcpyAsync(p.P + base, pv.data(), max_cells * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p.E + base, ev.data(), max_cells * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p.GS + base, gv.data(), max_cells * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(p.slot_sys + base, owner.data(), max_cells * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void*)(p.prm + sid), prm, sizeof(cs3d_params), hipMemcpyHostToDevice, c->stream));
    SysMeta m;
    std::memset(&m, 0, sizeof(m));
    m.n = n_cells;
    m.cap = max_cells;
    m.base = (int32_t)base;
    m.active = 1;
    m.step = 0;
    m.horizon = -1;
    m.vbase = vreg;
    m.lbase = vreg + vb;
    m.vbudget = (int32_t)vb;
    m.vcap = vcap;
    m.dims[0] = m.dims[1] = m.dims[2] = 1;
    m.bbudget = (int32_t)bb;
    m.bbase = breg;
    m.ebase = (breg + bb + 3) / 4 * 4;
    m.bh = prm->attraction_range > 0 ? prm->attraction_range * 1.001f : 1.0f;
    m.bdims[0] = m.bdims[1] = m.bdims[2] = 3;  // at least 3 per axis: the 27 periodic neighbours are distinct
    m.need_search = 1;
    for (int a = 0; a < 3; ++a) {  // minmaxpre sentinels (propagatebound.cu:161-166)
        float lo = 1.0E8f, hi = -1.0E8f;
        int32_t il, ih;
        std::memcpy(&il, &lo, 4);
        std::memcpy(&ih, &hi, 4);
        m.ext[2 * a] = il >= 0 ? il : il ^ 0x7FFFFFFF;
        m.ext[2 * a + 1] = ih >= 0 ? ih : ih ^ 0x7FFFFFFF;
    }
    cs3d_ranmar_init(&m.rng, prm->ranmar_ij, prm->ranmar_kl);
    kick_consts(prm->mass, prm->dt, m);
    launch_set_nsys(c->pool, c->nsys_high, c->stream);
    st = write_meta(c, sid, m);
    if (st != CS3D_OK) return st;
    // a captured all-systems step without CorrectCoMMotion cannot serve this system
    auto g = c->graphs.find(CS3D_ALL);
    if (prm->correct_com && g != c->graphs.end() && !g->second.shift) drop_graphs(c, CS3D_ALL);
    *out = sid;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_add(cs3d_ctx* c, const cs3d_params* prm, int32_t n_cells, int32_t max_cells,
                                       int64_t voxels, const float* x, const float* y, const float* z,
                                       const float* vx, const float* vy, const float* vz, cs3d_sys* out) {
    if (!x || !y || !z) return fail(CS3D_INVALID_ARG, "positions are required");
    cs3d_sys s;
    cs3d_status st = cs3d_system_add_raw(c, prm, n_cells, max_cells, voxels, &s);
    if (st != CS3D_OK) return st;
    const size_t bytes = (size_t)n_cells * kSlots * 4;
    const int64_t nb = (int64_t)c->sys[s].base * kSlots;
    const Pool& p = c->pool;
    auto up = [&](float* d, const float* h) -> cs3d_status {
        if (!h || bytes == 0) return CS3D_OK;
        HIPCHK(hipMemcpyAsync(d + nb, h, bytes, hipMemcpyHostToDevice, c->stream));
        return CS3D_OK;
    };
    if ((st = up(p.X, x)) || (st = up(p.Y, y)) || (st = up(p.Z, z)) || (st = up(p.VX, vx)) || (st = up(p.VY, vy)) ||
        (st = up(p.VZ, vz))) {
        cs3d_system_remove(c, s);
        return st;
    }
    if ((st = run_startup(c, s)) != CS3D_OK) {
        cs3d_system_remove(c, s);
        return st;
    }
    *out = s;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_startup(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    return run_startup(c, s);
}

extern "C" cs3d_status cs3d_system_remove(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    HIPCHK(hipStreamSynchronize(c->stream));
    SysHost& h = c->sys[s];
    std::vector<int32_t> owner(h.cap, -1);
    HIPCHK(hipMemcpy(c->pool.slot_sys + h.base, owner.data(), h.cap * 4, hipMemcpyHostToDevice));
    SysMeta m;
    std::memset(&m, 0, sizeof(m));
    HIPCHK(hipMemcpy(c->pool.meta + s, &m, sizeof(m), hipMemcpyHostToDevice));
    c->cells.release(h.base);
    c->voxels.release(h.vreg);
    c->bins.release(h.breg);
    drop_graphs(c, s);
    h = SysHost();
    int hi = 0;
    for (int i = 0; i < (int)c->sys.size(); ++i)
        if (c->sys[i].used) hi = i + 1;
    c->nsys_high = hi;
    launch_set_nsys(c->pool, hi, c->stream);
    HIPCHK(hipGetLastError());
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_set_horizon(cs3d_ctx* c, cs3d_sys s, int64_t horizon) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    m.horizon = horizon;
    return write_meta(c, s, m);
}

extern "C" cs3d_status cs3d_systems_set_horizon(cs3d_ctx* c, int32_t first, int32_t count, const int64_t* horizons) {
    if (!c || first < 0 || count < 0 || (count > 0 && !horizons) || (int64_t)first + count > (int64_t)c->sys.size())
        return fail(CS3D_INVALID_ARG, "bad horizon range");
    if (count == 0) return CS3D_OK;
    // the stream is idle while the metadata of the range makes its round trip
    std::vector<SysMeta> m(count);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpyAsync(m.data(), c->pool.meta + first, sizeof(SysMeta) * count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t i = 0; i < count; ++i)
        if (c->sys[first + i].used) m[i].horizon = horizons[i];
    HIPCHK(hipMemcpyAsync(c->pool.meta + first, m.data(), sizeof(SysMeta) * count, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_info(cs3d_ctx* c, cs3d_sys s, int32_t* n_cells, int64_t* step, uint32_t* status) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    if (n_cells) *n_cells = m.n;
    if (step) *step = m.step;
    if (status) *status = m.status;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_system_set_counts(cs3d_ctx* c, cs3d_sys s, int32_t n_cells, int64_t step) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (n_cells < 0 || n_cells > c->sys[s].cap) return fail(CS3D_INVALID_ARG, "cell count out of range");
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    m.n = n_cells;
    m.step = step;
    return write_meta(c, s, m);
}

extern "C" size_t cs3d_field_bytes(cs3d_ctx* c, cs3d_sys s, cs3d_field f) {
    if (!valid_sys(c, s)) return 0;
    return field_loc(c, s, f).bytes;
}

extern "C" cs3d_status cs3d_get(cs3d_ctx* c, cs3d_sys s, cs3d_field f, void* host, size_t bytes) {
    if (!valid_sys(c, s) || !host) return fail(CS3D_INVALID_ARG, "bad get arguments");
    FieldLoc L = field_loc(c, s, f);
    if (L.bytes == 0) return fail(CS3D_INVALID_ARG, "unknown field");
    if (bytes > L.bytes) return fail(CS3D_INVALID_ARG, "requested more bytes than the field holds");
    if (f == CS3D_F_GRID || f == CS3D_F_RANMAR || f == CS3D_F_SYSCM) {
        SysMeta m;
        cs3d_status st = read_meta(c, s, m);
        if (st != CS3D_OK) return st;
        if (f == CS3D_F_GRID) {
            int32_t g[8];
            std::memcpy(g, m.gmin, 12);
            g[3] = m.dims[0]; g[4] = m.dims[1]; g[5] = m.dims[2]; g[6] = m.vbudget; g[7] = m.vcap;
            std::memcpy(host, g, bytes);
        } else if (f == CS3D_F_RANMAR) {
            std::memcpy(host, &m.rng, bytes);
        } else {
            std::memcpy(host, m.sys_cm, bytes);
        }
        return CS3D_OK;
    }
    if (f == CS3D_F_BBOX || f == CS3D_F_VOXCOUNT || f == CS3D_F_VOXLIST) {
        // the fused step keeps voxel counts only and leaves the per-cell part of a
        // grid rebuilt in the last step to the next step: complete it and fill the
        // sorted voxel lists of the current grid
        const Range r = range_of(c, s);
        launch_vox_lists(c->pool, r, vox_blocks_of(c, s), c->stream);
        launch_set_flags(c->pool, r, 0, -1, c->stream);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(host, L.ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_get_state(cs3d_ctx* c, cs3d_sys s, float* xyz, float* v, float* P, float* CM, float* V,
                                      int32_t* n_cells, uint32_t* status) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    const size_t nn = (size_t)m.n * kSlots, nc = (size_t)m.n;
    auto blocks = [&](float* dst, const cs3d_field* f, int k, size_t count) -> cs3d_status {
        if (!dst || count == 0) return CS3D_OK;
        for (int a = 0; a < k; ++a) {
            const cs3d_status e = cs3d_get(c, s, f[a], dst + a * count, count * sizeof(float));
            if (e != CS3D_OK) return e;
        }
        return CS3D_OK;
    };
    const cs3d_field fx[3] = {CS3D_F_X, CS3D_F_Y, CS3D_F_Z};
    const cs3d_field fv[3] = {CS3D_F_VX, CS3D_F_VY, CS3D_F_VZ};
    const cs3d_field fc[3] = {CS3D_F_CMX, CS3D_F_CMY, CS3D_F_CMZ};
    const cs3d_field fp[1] = {CS3D_F_PRESSURE};
    const cs3d_field fo[1] = {CS3D_F_VOLUME};
    if ((st = blocks(xyz, fx, 3, nn)) != CS3D_OK || (st = blocks(v, fv, 3, nn)) != CS3D_OK ||
        (st = blocks(P, fp, 1, nc)) != CS3D_OK || (st = blocks(CM, fc, 3, nc)) != CS3D_OK ||
        (st = blocks(V, fo, 1, nc)) != CS3D_OK)
        return st;
    if (n_cells) *n_cells = m.n;
    if (status) *status = m.status;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_set(cs3d_ctx* c, cs3d_sys s, cs3d_field f, const void* host, size_t bytes) {
    if (!valid_sys(c, s) || !host) return fail(CS3D_INVALID_ARG, "bad set arguments");
    FieldLoc L = field_loc(c, s, f);
    if (L.bytes == 0) return fail(CS3D_INVALID_ARG, "unknown field");
    if (bytes > L.bytes) return fail(CS3D_INVALID_ARG, "more bytes than the field holds");
    if (f == CS3D_F_SYSCM) return fail(CS3D_INVALID_ARG, "the system centre is read-only");
    if (f == CS3D_F_CANDCOUNT) return fail(CS3D_INVALID_ARG, "the candidate counts are read-only");
    if (f == CS3D_F_GRID || f == CS3D_F_RANMAR) {
        SysMeta m;
        cs3d_status st = read_meta(c, s, m);
        if (st != CS3D_OK) return st;
        if (f == CS3D_F_GRID) {
            if (bytes < 24) return fail(CS3D_INVALID_ARG, "grid needs origin and dims");
            int32_t g[8];
            std::memcpy(g, host, std::min<size_t>(bytes, 32));
            std::memcpy(m.gmin, g, 12);
            m.dims[0] = g[3]; m.dims[1] = g[4]; m.dims[2] = g[5];
            if ((int64_t)g[3] * g[4] * g[5] > m.vbudget || g[3] < 1 || g[4] < 1 || g[5] < 1)
                return fail(CS3D_INVALID_ARG, "grid dims exceed the system's voxel budget");
        } else {
            std::memcpy(&m.rng, host, bytes);
        }
        cs3d_status st2 = write_meta(c, s, m);
        if (st2 != CS3D_OK || f != CS3D_F_GRID) return st2;
        launch_vrange(c->pool, range_of(c, s), c->stream);  // voxel ranges follow the new grid
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return CS3D_OK;
    }
    HIPCHK(hipMemcpyAsync(L.ptr, host, bytes, hipMemcpyHostToDevice, c->stream));
    if (f == CS3D_F_BBOX) launch_vrange(c->pool, range_of(c, s), c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f == CS3D_F_X || f == CS3D_F_Y || f == CS3D_F_Z || f == CS3D_F_BBOX) {  // stored candidates are stale
        SysMeta m;
        cs3d_status st = read_meta(c, s, m);
        if (st != CS3D_OK) return st;
        m.need_search = 1;
        return write_meta(c, s, m);
    }
    return CS3D_OK;
}

// ---- checkpoint / restart -------------------------------------------------------
// File layout (little-endian): "CS3DCKP1", u32 version, u32 sizeof(cs3d_params),
// cs3d_params, i32 n, i32 saved_cap, i64 step, i64 horizon, i64 cell steps, u32 status,
// cs3d_ranmar_state, grid geometry {f32 gmin[3], i32 dims[3], f32 bmin[3], f32 bh,
// i32 bdims[3]}, f32 system centre[3], u32 mesh checksum, then for cells 0..n-1: 15 node arrays
// (x, y, z, v, F_con, F_dis, F_ran; 192 floats per cell each), 8 cell arrays (CM, P,
// stiffness, growth scale, volume, area) and the contact-rule bbox and its voxel range
// (6 each).  Volume and area are kept so that a restarted system reports the last
// computed values, as an uninterrupted run does on steps without a volume pass.
namespace {
const char kCkptMagic[8] = {'C', 'S', '3', 'D', 'C', 'K', 'P', '1'};
constexpr uint32_t kCkptVersion = 4;

uint32_t mesh_checksum(const cs3d_ctx* c) {
    // FNV-1a over the mesh tables the context was built with
    uint32_t h = 2166136261u;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&c->mesh_host);
    for (size_t i = 0; i < sizeof(cs3d_mesh); ++i) h = (h ^ b[i]) * 16777619u;
    return h;
}

struct CkptGrid {
    float gmin[3];
    int32_t dims[3];
    float bmin[3];
    float bh;
    int32_t bdims[3];
};

template <class T> bool put(FILE* f, const T& v) { return std::fwrite(&v, sizeof(T), 1, f) == 1; }
template <class T> bool get(FILE* f, T& v) { return std::fread(&v, sizeof(T), 1, f) == 1; }
}  // namespace

extern "C" cs3d_status cs3d_system_save(cs3d_ctx* c, cs3d_sys s, const char* path) {
    if (!valid_sys(c, s) || !path) return fail(CS3D_INVALID_ARG, "bad save arguments");
    const Pool& p = c->pool;
    const Range r = range_of(c, s);
    // a grid rebuilt at the end of the last step is completed first (its bboxes
    // become current), as a read of the grid fields does
    launch_vox_lists(p, r, vox_blocks_of(c, s), c->stream);
    launch_set_flags(p, r, 0, -1, c->stream);
    HIPCHK(hipGetLastError());
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    const SysHost& h = c->sys[s];
    const int64_t n = m.n, base = h.base;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CS3D_INVALID_ARG, std::string("cannot open ") + path);
    bool ok = std::fwrite(kCkptMagic, 1, 8, f) == 8 && put(f, kCkptVersion) && put(f, (uint32_t)sizeof(cs3d_params)) &&
              put(f, h.prm) && put(f, (int32_t)n) && put(f, (int32_t)h.cap) && put(f, m.step) && put(f, m.horizon) &&
              put(f, m.cell_steps) && put(f, m.status) && put(f, m.rng);
    CkptGrid g;
    std::memcpy(g.gmin, m.gmin, sizeof(g.gmin));
    std::memcpy(g.dims, m.dims, sizeof(g.dims));
    std::memcpy(g.bmin, m.bmin, sizeof(g.bmin));
    g.bh = m.bh;
    std::memcpy(g.bdims, m.bdims, sizeof(g.bdims));
    ok = ok && put(f, g) && put(f, m.sys_cm) && put(f, mesh_checksum(c));
    const float* node[] = {p.X,   p.Y,   p.Z,   p.VX,  p.VY,  p.VZ,  p.FCX, p.FCY, p.FCZ,
                           p.FDX, p.FDY, p.FDZ, p.FRX, p.FRY, p.FRZ};
    std::vector<float> buf((size_t)std::max<int64_t>(n, 1) * kSlots);
    for (const float* a : node) {
        if (!ok) break;
        if (hipMemcpy(buf.data(), a + base * kSlots, (size_t)n * kSlots * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            std::fclose(f);
            return fail(CS3D_HIP_ERROR, "checkpoint read");
        }
        ok = std::fwrite(buf.data(), 4, (size_t)n * kSlots, f) == (size_t)n * kSlots;
    }
    const float* cell[] = {p.CMX, p.CMY, p.CMZ, p.P, p.E, p.GS, p.VOL, p.AREA};
    for (const float* a : cell) {
        if (!ok) break;
        if (hipMemcpy(buf.data(), a + base, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            std::fclose(f);
            return fail(CS3D_HIP_ERROR, "checkpoint read");
        }
        ok = std::fwrite(buf.data(), 4, (size_t)n, f) == (size_t)n;
    }
    if (ok) {
        std::vector<float> bb((size_t)std::max<int64_t>(n, 1) * 6);
        std::vector<int32_t> vr((size_t)std::max<int64_t>(n, 1) * 6);
        ok = hipMemcpy(bb.data(), p.BBOX + base * 6, (size_t)n * 24, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(vr.data(), p.VRANGE + base * 6, (size_t)n * 24, hipMemcpyDeviceToHost) == hipSuccess &&
             std::fwrite(bb.data(), 4, (size_t)n * 6, f) == (size_t)n * 6 &&
             std::fwrite(vr.data(), 4, (size_t)n * 6, f) == (size_t)n * 6;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? CS3D_OK : fail(CS3D_INVALID_ARG, std::string("cannot write ") + path);
}

extern "C" cs3d_status cs3d_system_load(cs3d_ctx* c, const char* path, int32_t max_cells, int64_t voxels,
                                        cs3d_sys* out) {
    if (!c || !path || !out) return fail(CS3D_INVALID_ARG, "bad load arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CS3D_INVALID_ARG, std::string("cannot open ") + path);
    char magic[8];
    uint32_t ver = 0, psize = 0, mcheck = 0;
    cs3d_params prm;
    int32_t n = 0, scap = 0;
    int64_t step = 0, horizon = -1, cell_steps = 0;
    uint32_t status = 0;
    cs3d_ranmar_state rng;
    CkptGrid g;
    float sys_cm[3];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCkptMagic, 8) == 0 && get(f, ver) &&
              ver == kCkptVersion && get(f, psize) && psize == sizeof(cs3d_params) && get(f, prm) && get(f, n) &&
              get(f, scap) && get(f, step) && get(f, horizon) && get(f, cell_steps) && get(f, status) && get(f, rng) && get(f, g) &&
              get(f, sys_cm) && get(f, mcheck);
    if (!ok) {
        std::fclose(f);
        return fail(CS3D_INVALID_ARG, std::string("not a checkpoint of this engine: ") + path);
    }
    if (mcheck != mesh_checksum(c)) {
        std::fclose(f);
        return fail(CS3D_INVALID_ARG, "checkpoint was written with a different mesh");
    }
    if (max_cells <= 0) max_cells = scap;
    cs3d_sys s = -1;
    cs3d_status st = cs3d_system_add_raw(c, &prm, n, max_cells, voxels, &s);
    if (st != CS3D_OK) {
        std::fclose(f);
        return st;
    }
    const Pool& p = c->pool;
    const int64_t base = c->sys[s].base;
    std::vector<float> buf((size_t)std::max(n, 1) * kSlots);
    float* node[] = {p.X, p.Y, p.Z, p.VX, p.VY, p.VZ, p.FCX, p.FCY, p.FCZ, p.FDX, p.FDY, p.FDZ, p.FRX, p.FRY, p.FRZ};
    for (float* a : node) {
        if (!ok) break;
        ok = std::fread(buf.data(), 4, (size_t)n * kSlots, f) == (size_t)n * kSlots &&
             hipMemcpy(a + base * kSlots, buf.data(), (size_t)n * kSlots * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    float* cell[] = {p.CMX, p.CMY, p.CMZ, p.P, p.E, p.GS, p.VOL, p.AREA};
    for (float* a : cell) {
        if (!ok) break;
        ok = std::fread(buf.data(), 4, (size_t)n, f) == (size_t)n &&
             hipMemcpy(a + base, buf.data(), (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (ok) {
        std::vector<float> bb((size_t)std::max(n, 1) * 6);
        std::vector<int32_t> vr((size_t)std::max(n, 1) * 6);
        ok = std::fread(bb.data(), 4, (size_t)n * 6, f) == (size_t)n * 6 &&
             std::fread(vr.data(), 4, (size_t)n * 6, f) == (size_t)n * 6 &&
             hipMemcpy(p.BBOX + base * 6, bb.data(), (size_t)n * 24, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(p.VRANGE + base * 6, vr.data(), (size_t)n * 24, hipMemcpyHostToDevice) == hipSuccess;
    }
    std::fclose(f);
    SysMeta m;
    if (ok) ok = read_meta(c, s, m) == CS3D_OK;
    if (ok) {
        m.step = step;
        m.horizon = horizon;
        m.cell_steps = cell_steps;
        m.status = status;
        m.rng = rng;
        std::memcpy(m.gmin, g.gmin, sizeof(g.gmin));
        std::memcpy(m.dims, g.dims, sizeof(g.dims));
        std::memcpy(m.bmin, g.bmin, sizeof(g.bmin));
        m.bh = g.bh;
        std::memcpy(m.bdims, g.bdims, sizeof(g.bdims));
        std::memcpy(m.sys_cm, sys_cm, sizeof(sys_cm));
        m.shift_pending = 0;
        m.rebuild = 0;
        m.need_search = 1;  // candidate lists are rebuilt at the next step
        m.search_next = 0;
        ok = (int64_t)m.dims[0] * m.dims[1] * m.dims[2] <= m.vbudget && write_meta(c, s, m) == CS3D_OK;
    }
    if (!ok) {
        cs3d_system_remove(c, s);
        return fail(CS3D_INVALID_ARG, std::string("truncated or inconsistent checkpoint: ") + path);
    }
    *out = s;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_set_graph(cs3d_ctx* c, int32_t enable) {
    if (!c) return fail(CS3D_INVALID_ARG, "NULL context");
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graphs(c);
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_set_launch(cs3d_ctx* c, int32_t wg_per_cu) {
    if (!c || wg_per_cu < 1 || wg_per_cu > 1024) return fail(CS3D_INVALID_ARG, "wg_per_cu must be in [1, 1024]");
    c->pool.wg_per_cu = wg_per_cu;
    drop_graphs(c);
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_systems_poll(cs3d_ctx* c, int32_t first, int32_t count, int32_t* n_cells, int64_t* steps,
                                         uint32_t* status) {
    if (!c || first < 0 || count < 0 || first + count > kMaxSys) return fail(CS3D_INVALID_ARG, "bad poll range");
    if (count == 0) return CS3D_OK;
    std::vector<SysMeta> m(count);
    HIPCHK(hipMemcpyAsync(m.data(), c->pool.meta + first, sizeof(SysMeta) * count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t i = 0; i < count; ++i) {
        const bool used = c->sys[first + i].used;
        if (n_cells) n_cells[i] = used ? m[i].n : -1;
        if (steps) steps[i] = used ? m[i].step : -1;
        if (status) status[i] = used ? m[i].status : 0u;
    }
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_systems_work(cs3d_ctx* c, int32_t first, int32_t count, int64_t* cell_steps) {
    if (!c || !cell_steps || first < 0 || count < 0 || first + count > kMaxSys)
        return fail(CS3D_INVALID_ARG, "bad work-query range");
    if (count == 0) return CS3D_OK;
    std::vector<SysMeta> m(count);
    HIPCHK(hipMemcpyAsync(m.data(), c->pool.meta + first, sizeof(SysMeta) * count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int32_t i = 0; i < count; ++i) cell_steps[i] = c->sys[first + i].used ? m[i].cell_steps : -1;
    return CS3D_OK;
}

// Moves a system to a larger slot range (with voxel and search-bin regions sized
// for it) and frees the old one.  Cells keep their local indices, so contact keys,
// event order and every result are unchanged; the next step searches all
// contact candidates again (which does not change any result either).
extern "C" cs3d_status cs3d_system_grow(cs3d_ctx* c, cs3d_sys s, int32_t max_cells, int64_t voxels) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    SysHost& h = c->sys[s];
    if (max_cells <= h.cap) return fail(CS3D_INVALID_ARG, "the new room must exceed the current one");
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    int64_t base = 0, vreg = 0, breg = 0;
    if (!c->cells.alloc(max_cells, base)) return fail(CS3D_CAPACITY, "cell pool exhausted");
    int64_t vb = voxels > 0 ? voxels : std::max<int64_t>(4096, 32 * (int64_t)max_cells);
    vb = std::min<int64_t>(std::max<int64_t>(vb, h.vbudget), 1ll << 30);
    if (!c->voxels.alloc(vb * (1 + h.vcap), vreg)) {
        c->cells.release(base);
        return fail(CS3D_CAPACITY, "voxel pool exhausted");
    }
    int64_t bb = std::max<int64_t>(16384, 640 * (int64_t)max_cells);
    bb = std::min<int64_t>(bb, c->bin_total / (1 + 4 * kBCap)) / 4 * 4;
    while (bb >= h.bbudget && !c->bins.alloc(bb * (1 + 4 * kBCap) + 4, breg)) bb /= 2;
    if (bb < h.bbudget) {  // the current bin grid must fit the new region
        c->cells.release(base);
        c->voxels.release(vreg);
        return fail(CS3D_CAPACITY, "search-bin pool exhausted");
    }
    HIPCHK(hipSetDevice(c->device));
    const Pool& p = c->pool;
    hipStream_t q = c->stream;
    const int64_t n = m.n, ob = h.base, ns = (int64_t)max_cells * kSlots;
    // the new range as cs3d_system_add leaves it, then the live cells copied over
    float* node[] = {p.X,   p.Y,   p.Z,   p.VX,  p.VY,  p.VZ,  p.VAX, p.VAY, p.VAZ, p.FCX, p.FCY,
                     p.FCZ, p.FDX, p.FDY, p.FDZ, p.FRX, p.FRY, p.FRZ, p.XBX, p.XBY, p.XBZ};
    for (float* a : node) {
        HIPCHK(hipMemsetAsync(a + base * kSlots, 0, ns * 4, q));
        HIPCHK(hipMemcpyAsync(a + base * kSlots, a + ob * kSlots, n * kSlots * 4, hipMemcpyDeviceToDevice, q));
    }
    int32_t* nodei[] = {p.NPAIR, p.NCAND};
    for (int32_t* a : nodei) {
        HIPCHK(hipMemsetAsync(a + base * kSlots, 0, ns * 4, q));
        HIPCHK(hipMemcpyAsync(a + base * kSlots, a + ob * kSlots, n * kSlots * 4, hipMemcpyDeviceToDevice, q));
    }
    int32_t* lists[] = {p.PAIRS, p.CAND};
    for (int32_t* a : lists)
        HIPCHK(hipMemcpyAsync(a + base * kPMax * kSlots, a + ob * kPMax * kSlots, n * kPMax * kSlots * 4,
                              hipMemcpyDeviceToDevice, q));
    std::vector<float> pv(max_cells, h.prm.min_pressure), ev(max_cells, h.prm.stiffness1), gv(max_cells, 1.0f);
    HIPCHK(hipMemcpyAsync(p.P + base, pv.data(), max_cells * 4, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(p.E + base, ev.data(), max_cells * 4, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(p.GS + base, gv.data(), max_cells * 4, hipMemcpyHostToDevice, q));
    float* cellz[] = {p.CMX, p.CMY, p.CMZ, p.VOL, p.AREA};
    for (float* a : cellz) HIPCHK(hipMemsetAsync(a + base, 0, (size_t)max_cells * 4, q));
    float* cells[] = {p.CMX, p.CMY, p.CMZ, p.P, p.E, p.VOL, p.AREA, p.GS};
    for (float* a : cells) HIPCHK(hipMemcpyAsync(a + base, a + ob, n * 4, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipMemsetAsync(p.FLAG + base, 0, max_cells, q));
    HIPCHK(hipMemcpyAsync(p.FLAG + base, p.FLAG + ob, n, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipMemsetAsync(p.RESET + base, 0, max_cells, q));
    float* boxes[] = {p.BBOX, p.BBOXN, p.SBOX};
    for (float* a : boxes) {
        HIPCHK(hipMemsetAsync(a + base * 6, 0, (size_t)max_cells * 24, q));
        HIPCHK(hipMemcpyAsync(a + base * 6, a + ob * 6, n * 24, hipMemcpyDeviceToDevice, q));
    }
    HIPCHK(hipMemcpyAsync(p.VRANGE + base * 6, p.VRANGE + ob * 6, n * 24, hipMemcpyDeviceToDevice, q));
    // voxel counts of the current grid (the voxel index depends on the grid dims only),
    // the rest of the voxel region and the bin counters cleared
    HIPCHK(hipMemsetAsync(p.vox + vreg, 0, (size_t)vb * (1 + h.vcap) * 4, q));
    HIPCHK(hipMemcpyAsync(p.vox + vreg, p.vox + h.vreg, (size_t)h.vbudget * 4, hipMemcpyDeviceToDevice, q));
    HIPCHK(hipMemsetAsync(p.bins + breg, 0, (size_t)bb * 4, q));
    std::vector<int32_t> owner(max_cells, s), none(h.cap, -1);
    HIPCHK(hipMemcpyAsync(p.slot_sys + base, owner.data(), max_cells * 4, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(p.slot_sys + ob, none.data(), h.cap * 4, hipMemcpyHostToDevice, q));
    m.base = (int32_t)base;
    m.cap = max_cells;
    m.vbase = vreg;
    m.lbase = vreg + vb;
    m.vbudget = (int32_t)vb;
    m.bbudget = (int32_t)bb;
    m.bbase = breg;
    m.ebase = (breg + bb + 3) / 4 * 4;
    m.need_search = 1;
    m.n_reset = 0;
    m.search_next = 0;
    st = write_meta(c, s, m);  // synchronises the stream (the host vectors above are still alive)
    if (st != CS3D_OK) return st;
    c->cells.release(h.base);
    c->voxels.release(h.vreg);
    c->bins.release(h.breg);
    h.base = (int32_t)base;
    h.cap = max_cells;
    h.vreg = vreg;
    h.vbudget = (int32_t)vb;
    h.breg = breg;
    h.bbudget = (int32_t)bb;
    drop_graphs(c, s);
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_set_graph_steps(cs3d_ctx* c, int32_t steps) {
    if (!c || steps < 1 || steps > 1000) return fail(CS3D_INVALID_ARG, "graph steps must be in [1, 1000]");
    c->graph_steps = steps;
    drop_graphs(c);
    return CS3D_OK;
}

namespace {
// Runs enqueue(q) for q in [0, groups) with every launch timed by its own
// start/stop events (the launch helper's kernel timer) and adds each group's
// summed kernel durations to ms[q].
template <class Enqueue>
cs3d_status timed_groups(cs3d_ctx* c, int groups, Enqueue enqueue, double* ms) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    std::vector<size_t> mark(groups + 1, 0);
    kernel_timer = &evs;
    for (int q = 0; q < groups; ++q) {
        mark[q] = evs.size();
        enqueue(q);
    }
    mark[groups] = evs.size();
    kernel_timer = nullptr;
    hipError_t e = hipStreamSynchronize(c->stream);
    for (int q = 0; q < groups && e == hipSuccess; ++q)
        for (size_t i = mark[q]; i < mark[q + 1]; ++i) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, evs[i].first, evs[i].second) == hipSuccess) ms[q] += t;
        }
    for (auto& pr : evs) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    return e == hipSuccess ? CS3D_OK : fail(CS3D_HIP_ERROR, hipGetErrorString(e));
}
}  // namespace

// Per-phase time of n_steps eager steps: each phase's kernel durations, from
// events the launches themselves record (no launch gaps).
extern "C" cs3d_status cs3d_step_timed(cs3d_ctx* c, cs3d_sys s, int32_t n_steps, double* phase_ms) {
    if (!c || !phase_ms) return fail(CS3D_INVALID_ARG, "NULL argument");
    if (s != CS3D_ALL && !valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    const Range r = range_of(c, s);
    if (r.slot_hi <= r.slot_lo || r.sys_hi <= r.sys_lo) return CS3D_OK;
    cs3d_status st = CS3D_OK;
    for (int k = 0; k < n_steps && st == CS3D_OK; ++k)
        st = timed_groups(c, CS3D_PHASES, [&](int q) { enqueue_step_phase(c, r, q); }, phase_ms);
    return st;
}

namespace {
// The captured graph_steps steps of a target, captured, instantiated and
// uploaded on first use.
cs3d_status graph_of(cs3d_ctx* c, cs3d_sys s, const cs3d_ctx::GraphRec** out) {
    auto it = c->graphs.find(s);
    if (it == c->graphs.end()) {
        const Range r = range_of(c, s);
        const int vb = vox_blocks_of(c, s);
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < c->graph_steps; ++k) enqueue_step(c, r, vb);
        HIPCHK(hipStreamEndCapture(c->stream, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(CS3D_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        e = hipGraphUpload(ex, c->stream);
        if (e != hipSuccess) {
            (void)hipGraphExecDestroy(ex);
            return fail(CS3D_HIP_ERROR, std::string("hipGraphUpload: ") + hipGetErrorString(e));
        }
        it = c->graphs.emplace(s, cs3d_ctx::GraphRec{ex, c->graph_steps, any_correct_com(c, r)}).first;
    }
    *out = &it->second;
    return CS3D_OK;
}
}  // namespace

extern "C" cs3d_status cs3d_graph_prepare(cs3d_ctx* c, cs3d_sys s) {
    if (!c) return fail(CS3D_INVALID_ARG, "NULL context");
    if (s != CS3D_ALL && !valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (!c->use_graph) return CS3D_OK;
    const cs3d_ctx::GraphRec* g = nullptr;
    cs3d_status st = graph_of(c, s, &g);
    if (st != CS3D_OK) return st;
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_step(cs3d_ctx* c, cs3d_sys s, int32_t n_steps) {
    if (!c) return fail(CS3D_INVALID_ARG, "NULL context");
    if (s != CS3D_ALL && !valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (n_steps <= 0) return CS3D_OK;
    if (s == CS3D_ALL && c->nsys_high == 0) return CS3D_OK;
    const Range r = range_of(c, s);
    const int vb = vox_blocks_of(c, s);
    int done = 0;
    if (c->use_graph && n_steps >= c->graph_steps) {
        const cs3d_ctx::GraphRec* g = nullptr;
        cs3d_status st = graph_of(c, s, &g);
        if (st != CS3D_OK) return st;
        for (; done + g->steps <= n_steps; done += g->steps) HIPCHK(hipGraphLaunch(g->exec, c->stream));
    }
    for (; done < n_steps; ++done) enqueue_step(c, r, vb);
    HIPCHK(hipGetLastError());
    return CS3D_OK;
}

// ---- per-kernel entry points -------------------------------------------------
namespace {
cs3d_status finish(cs3d_ctx* c) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return CS3D_OK;
}

// The context's scratch buffer with room for `bytes` (the stream is idle when an
// entry point asks for it: every entry point ends with finish()).
cs3d_status scratch_of(cs3d_ctx* c, size_t bytes, void** out) {
    if (c->scratch_bytes < bytes) {
        if (c->scratch) {
            c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), c->scratch));
            (void)hipFree(c->scratch);
            c->scratch = nullptr;
            c->scratch_bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes, 4096);
        hipError_t e = hipMalloc(&c->scratch, want);
        if (e != hipSuccess) {
            c->scratch = nullptr;
            return fail(CS3D_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        c->allocs.push_back(c->scratch);
        c->scratch_bytes = want;
    }
    *out = c->scratch;
    return CS3D_OK;
}
}  // namespace

extern "C" cs3d_status cs3d_k_integrate(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    const Range r = range_of(c, s);
    launch_integrate(c->pool, r, I_KICK | I_WRITEV, 0, c->stream);
    launch_set_flags(c->pool, r, -1, 1, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_pressure_update(cs3d_ctx* c, cs3d_sys s, int64_t step) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_integrate(c->pool, range_of(c, s), I_PRESSURE, step, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_con_force(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    const Range r = range_of(c, s);
    launch_integrate(c->pool, r, I_BINS, 0, c->stream);
    launch_search(c->pool, r, 1, c->stream);
    launch_con(c->pool, r, M_CON, false, c->stream);
    launch_integrate(c->pool, r, I_CLEAR, 0, c->stream);
    launch_set_flags(c->pool, r, -1, 1, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_con_force_bonded(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_bonded(c->pool, range_of(c, s), c->sys[s].prm.constrain_angles ? 1 : 0, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_dis_force(cs3d_ctx* c, cs3d_sys s, int32_t which) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    const Range r = range_of(c, s);
    launch_integrate(c->pool, r, I_BINS, 0, c->stream);
    launch_search(c->pool, r, 1, c->stream);
    launch_con(c->pool, r, M_PAIRS, false, c->stream);
    launch_dis(c->pool, r, which ? M_DIS_VA : M_DIS_V, 0, 1, which != 0, c->stream);
    launch_set_flags(c->pool, r, -1, 1, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_velocity_ab(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_velocity_ab(c->pool, range_of(c, s), 3, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_velocity_a(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_velocity_ab(c->pool, range_of(c, s), 1, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_velocity_b(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_velocity_ab(c->pool, range_of(c, s), 2, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_pressure_reset(cs3d_ctx* c, cs3d_sys s, const int32_t* cells, int32_t count,
                                             float value) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (count < 0 || (count > 0 && !cells)) return fail(CS3D_INVALID_ARG, "bad cell list");
    if (count == 0) return CS3D_OK;
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    for (int32_t i = 0; i < count; ++i)
        if (cells[i] < 0 || cells[i] >= m.n) return fail(CS3D_INVALID_ARG, "cell index out of range");
    void* buf = nullptr;
    if ((st = scratch_of(c, (size_t)count * sizeof(int32_t), &buf)) != CS3D_OK) return st;
    HIPCHK(hipMemcpyAsync(buf, cells, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    launch_pressure_reset(c->pool, range_of(c, s), (const int32_t*)buf, count, value, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_pressure_reset_noalloc(cs3d_ctx* c, cs3d_sys s, const int32_t* cells, int32_t count,
                                                     float value) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (count < 0 || (count > 0 && !cells)) return fail(CS3D_INVALID_ARG, "bad cell list");
    if (count > kNoAllocCells)
        return fail(CS3D_INVALID_ARG, "cell list longer than the allocation-free staging array (65536 cells)");
    if (count == 0) return CS3D_OK;
    SysMeta m;
    cs3d_status st = read_meta(c, s, m);
    if (st != CS3D_OK) return st;
    for (int32_t i = 0; i < count; ++i)
        if (cells[i] < 0 || cells[i] >= m.n) return fail(CS3D_INVALID_ARG, "cell index out of range");
    int32_t* buf = scratch_of_device(c->device);
    if (!buf) return fail(CS3D_HIP_ERROR, "no staging array on this device");
    HIPCHK(hipMemcpyAsync(buf, cells, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    launch_pressure_reset(c->pool, range_of(c, s), buf, count, value, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_com(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_dis(c->pool, range_of(c, s), M_COM, 0, 0, false, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_volumes(cs3d_ctx* c, cs3d_sys s, int64_t step) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_dis(c->pool, range_of(c, s), M_VOLUMES, step, 0, false, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_bboxes(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    SysMeta m0;
    cs3d_status st = read_meta(c, s, m0);
    if (st != CS3D_OK) return st;
    const Range r = range_of(c, s);
    launch_set_flags(c->pool, r, 1, -1, c->stream);  // the box kernel works on systems flagged for a rebuild
    launch_cellbox(c->pool, r, c->stream);
    if ((st = finish(c)) != CS3D_OK) return st;
    // restore the flag (a grid rebuilt at the end of the last step may still have
    // its per-cell part pending for the next step), and put the system extent the
    // box kernel accumulated back to the minmaxpre sentinels
    SysMeta m;
    if ((st = read_meta(c, s, m)) != CS3D_OK) return st;
    m.rebuild = m0.rebuild;
    for (int a = 0; a < 3; ++a) {
        float lo = 1.0E8f, hi = -1.0E8f;
        int32_t il, ih;
        std::memcpy(&il, &lo, 4);
        std::memcpy(&ih, &hi, 4);
        m.ext[2 * a] = il >= 0 ? il : il ^ 0x7FFFFFFF;
        m.ext[2 * a + 1] = ih >= 0 ? ih : ih ^ 0x7FFFFFFF;
    }
    return write_meta(c, s, m);
}

extern "C" cs3d_status cs3d_k_extent(cs3d_ctx* c, cs3d_sys s, float* out6) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    if (!out6) return fail(CS3D_INVALID_ARG, "NULL output");
    void* buf = nullptr;
    cs3d_status st = scratch_of(c, 6 * sizeof(float), &buf);
    if (st != CS3D_OK) return st;
    launch_extent(c->pool, range_of(c, s), (float*)buf, c->stream);
    HIPCHK(hipMemcpyAsync(out6, buf, 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

extern "C" cs3d_status cs3d_k_grid(cs3d_ctx* c, cs3d_sys s) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    const Range r = range_of(c, s);
    launch_set_flags(c->pool, r, 1, -1, c->stream);
    launch_grid(c->pool, r, vox_blocks_of(c, s), c->stream);
    launch_set_flags(c->pool, r, 0, -1, c->stream);
    return finish(c);
}

extern "C" cs3d_status cs3d_k_divide(cs3d_ctx* c, cs3d_sys s, int32_t* n_events) {
    if (!valid_sys(c, s)) return fail(CS3D_INVALID_ARG, "unknown system");
    launch_divide(c->pool, range_of(c, s), M_INIT, c->stream);
    cs3d_status st = finish(c);
    if (st != CS3D_OK) return st;
    if (n_events) {
        SysMeta m;
        st = read_meta(c, s, m);
        if (st != CS3D_OK) return st;
        *n_events = m.n_events;
    }
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_time_kernel(cs3d_ctx* c, cs3d_sys s, int32_t which, int32_t reps, double* avg_ms) {
    if (!valid_sys(c, s) || reps < 1 || !avg_ms || which < 0 || which > 2)
        return fail(CS3D_INVALID_ARG, "bad kernel timing arguments");
    const Range r = range_of(c, s);
    auto launch = [&]() {
        if (which == 0) launch_bonded(c->pool, r, c->sys[s].prm.constrain_angles ? 1 : 0, c->stream);
        else if (which == 1) launch_dis(c->pool, r, M_COM, 0, 0, false, c->stream);
        else launch_search(c->pool, r, 1, c->stream);
    };
    // the full search reads the search bins: filled from the current positions here,
    // cleared after, and the next step searches again (as after cs3d_k_con_force)
    if (which == 2) launch_integrate(c->pool, r, I_BINS, 0, c->stream);
    launch();  // warm-up
    double ms = 0.0;
    cs3d_status st = timed_groups(c, 1, [&](int) { for (int q = 0; q < reps; ++q) launch(); }, &ms);
    if (which == 2) {
        launch_integrate(c->pool, r, I_CLEAR, 0, c->stream);
        launch_set_flags(c->pool, r, -1, 1, c->stream);
        const cs3d_status st2 = finish(c);
        if (st == CS3D_OK) st = st2;
    }
    if (st != CS3D_OK) return st;
    *avg_ms = ms / reps;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_check_sqrt(cs3d_ctx* c, uint64_t* mismatches) {
    if (!c || !mismatches) return fail(CS3D_INVALID_ARG, "bad self-check arguments");
    unsigned long long* bad = nullptr;
    HIPCHK(hipMalloc(&bad, sizeof(*bad)));
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(*bad), c->stream);
    if (e == hipSuccess) {
        launch_check_sqrt(bad, c->stream);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(bad);
    if (e != hipSuccess) return fail(CS3D_HIP_ERROR, hipGetErrorString(e));
    *mismatches = (uint64_t)h;
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_check_div(cs3d_ctx* c, int32_t a_significands, int64_t scaled_pairs,
                                      uint64_t* mismatches) {
    if (!c || !mismatches || a_significands < 0 || a_significands > (1 << 23) || scaled_pairs < 0)
        return fail(CS3D_INVALID_ARG, "bad self-check arguments");
    unsigned long long* bad = nullptr;
    HIPCHK(hipMalloc(&bad, 6 * sizeof(*bad)));
    hipError_t e = hipMemsetAsync(bad, 0, 6 * sizeof(*bad), c->stream);
    if (e == hipSuccess) {
        launch_check_div(bad, (uint32_t)a_significands, (uint64_t)scaled_pairs, c->stream);
        e = hipGetLastError();
    }
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, bad, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(bad);
    if (e != hipSuccess) return fail(CS3D_HIP_ERROR, hipGetErrorString(e));
    for (int i = 0; i < 6; ++i) mismatches[i] = (uint64_t)h[i];
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_device_acos(cs3d_ctx* c, const double* x, double* y, int64_t n) {
    if (!c || n < 0 || (n > 0 && (!x || !y))) return fail(CS3D_INVALID_ARG, "bad self-check arguments");
    if (n == 0) return CS3D_OK;
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_acos(d, d + n, n, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, d + n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CS3D_HIP_ERROR, hipGetErrorString(e));
    return CS3D_OK;
}

extern "C" cs3d_status cs3d_copy_calibrate(cs3d_ctx* c, int64_t bytes, int32_t reps, double* avg_ms) {
    if (!c || bytes < 16 || reps < 1 || !avg_ms) return fail(CS3D_INVALID_ARG, "bad calibration arguments");
    const int64_t n4 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n4 * 16));
    if (hipMalloc(&b, n4 * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(CS3D_OUT_OF_MEMORY, "calibration buffer");
    }
    (void)hipMemsetAsync(a, 0, n4 * 16, c->stream);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch_copy_calibrate(a, b, n4, c->stream);  // warm-up
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch_copy_calibrate(a, b, n4, c->stream);
    (void)hipEventRecord(e1, c->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(CS3D_HIP_ERROR, hipGetErrorString(e));
    *avg_ms = ms / reps;
    return CS3D_OK;
}
