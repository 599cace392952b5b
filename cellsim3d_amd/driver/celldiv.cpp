// celldiv - host driver of the C180 cell-mechanics engine: a drop-in for the
// reference's CellDiv executable (src/GPUbounce.cu main, :289-1512).
//
//   celldiv <no_of_cells> <inp.json> <gpuid> [--grid-interval K] [--pair-skin S]
//           [--mesh-dir DIR] [--no-graph] [--daughter-com-fix] [--setup-only]
//           [--sync-output] [--dump-setup PREFIX] [--save-state FILE]
//
// Same command line, parameter file, input mesh files (C180, C180NN, C180C,
// C180_pentahexa in the working directory or --mesh-dir) and outputs: the
// trajectory (binary or ASCII, core.trajFileName), the forces CSV
// (core.forces_file when core.write_cont_force), and the empty psfil and
// velocity2.xyz the reference creates.  Frames are copied off the device at
// each trajectory step and written by a writer thread while the device steps on
// (two frame buffers; --sync-output writes in place).  The random force (add_rands) draws from a
// counter-based Philox stream instead of the reference's cuRAND states.  All
// simulation work runs on the GPU
// through the C ABI of libcs3d (include/cs3d.h); this file only parses,
// places, and writes.
//
// Exit codes follow the reference: -1 for I/O, parse or allocation errors
// (GPUbounce.cu:301, 1616-1886), 27 when the rigid box is too small (:1666-1670),
// -69 when the population outgrows MaxNoofC180s (:1265-1270), -60 for bad
// time-step parameters (:2079-2083).  Device-side failures (the reference's
// traps) print the system status and exit with 0, as CudaErrorCheck does
// (:39-45).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cs3d.h"
#include "json_lite.h"

namespace {

constexpr int S = CS3D_SLOTS;
constexpr int NN = CS3D_NODES;

// Everything read_json_params (GPUbounce.cu:1914-2247) reads that this driver uses.
struct Settings {
    cs3d_params p;
    int max_cells = 0;
    int time_steps = 0;  // core.div_time_steps (read with asFloat into an int, :1959)
    int equi_steps = 0;  // core.non_div_time_steps
    int traj_write_int = 0;
    std::string traj_file;
    bool binary = true;
    bool write_cont_force = false;
    std::string forces_file;
    int ran_z_offset = 0;
    float z_offset = 0.f;
    int restart = 0;
    bool do_adaptive_dt = false;
    float dt_max = 0.f, dt_tol = 0.f;
    bool correct_com = false;
    int new_cell_count_int = 0;
    bool count_only_internal = false;
    float rad_frac = 0.f;
    std::string mit_ind_file;
    bool overwrite_mit_ind = false;
    bool during_growth = false;
    bool choose_random_cell_indices = false;
    int use_walls = 0;
    std::string perp_axis;
    float d_axis = 0.f;
    float closeness_to_center = 0.f;
    float fraction_of_softer_cells = 0.f;
    int number_of_softer_cells = 0;
    bool flatbox = false;
    bool rand_pos = false;
    bool use_pbcs = false;
    int rand_seed = -1;
};

// read_json_params: parameter mapping (:1933-2076) and validation (:2077-2244).
int read_params(const char* path, Settings& s) {
    std::ifstream in(path);
    if (!in) {
        std::printf("Failed to parse %s\n", path);
        return -1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    jsonlite::Value root;
    try {
        root = jsonlite::parse(ss.str());
    } catch (const std::exception& e) {
        std::printf("Failed to parse %s\n%s\n", path, e.what());
        return -1;
    }
    std::printf("%s parsed successfully\n", path);
    cs3d_params_default(&s.p);
    cs3d_params& p = s.p;
    const jsonlite::Value& core = root["core"];
    if (core.is_null()) {
        std::printf("ERROR: Cannot load core simulation parameters\nExiting");
        return -1;
    }
    s.max_cells = core["MaxNoofC180s"].as_int();
    p.mass = core["particle_mass"].as_float();
    p.repulsion_range = core["repulsion_range"].as_float();
    p.attraction_range = core["attraction_range"].as_float();
    p.repulsion_strength = core["repulsion_strength"].as_float();
    p.attraction_strength = core["attraction_strength"].as_float();
    p.youngs_mod = core["Youngs_mod"].as_float();
    p.stiffness1 = core["stiffFactor1"].as_float() * p.youngs_mod;
    p.viscotic_damping = core["viscotic_damping"].as_float();
    p.internal_damping = core["internal_damping"].as_float();
    p.division_vol = core["division_Vol"].as_float();
    s.ran_z_offset = core["random_z_offset?"].as_int();
    s.z_offset = core["z_offset"].as_float();
    s.time_steps = (int)core["div_time_steps"].as_float();
    p.dt = core["time_interval"].as_float();
    s.restart = core["Restart"].as_int();
    s.traj_write_int = core["trajWriteInt"].as_int();
    s.equi_steps = core["non_div_time_steps"].as_int();
    s.traj_file = core["trajFileName"].as_string();
    s.binary = core["binaryOutput"].as_bool();
    p.max_pressure = core["maxPressure"].as_float();
    p.min_pressure = core["minPressure"].as_float();
    p.gamma_visc = core["gamma_visc"].as_float();
    p.growth_rate = core["growth_rate"].as_float();
    p.check_sphericity = core["checkSphericity"].as_bool();
    p.constrain_angles = core["constrainAngles"].as_bool();
    s.dt_max = core["dt_max"].as_float();
    s.dt_tol = core["dt_tol"].as_float();
    s.do_adaptive_dt = core["doAdaptive_dt"].as_bool();
    p.phase_count = core["phase_count"].as_int();
    s.write_cont_force = core["write_cont_force"].as_bool();
    s.forces_file = core["forces_file"].as_string();
    s.correct_com = core["correct_com"].as_bool();

    const jsonlite::Value& cnt = root["counting"];
    if (cnt.is_null()) {
        std::printf("ERROR: Cannot load counting parameters\nExiting");
        return -1;
    }
    s.mit_ind_file = cnt["mit-index_file_name"].as_string();
    s.count_only_internal = cnt["count_only_internal_cells?"].as_bool();
    s.rad_frac = cnt["radius_cutoff"].as_float();
    s.overwrite_mit_ind = cnt["overwrite_mit_ind_file?"].as_bool();
    s.new_cell_count_int = cnt["cell_count_int"].as_int();

    const jsonlite::Value& pop = root["population"];
    if (pop.is_null()) {
        std::printf("ERROR: Cannot load population parameters\nExiting");
        return -1;
    }
    p.do_pop_model = pop["doPopModel"].as_int();
    p.max_pop = pop["max_pop"].as_float();

    const jsonlite::Value& walls = root["walls"];
    if (walls.is_null()) {
        std::printf("ERROR: Cannot load wall parameters\nExiting");
        return -1;
    }
    s.use_walls = walls["useWalls"].as_int();
    s.perp_axis = walls["perpAxis"].as_string();
    s.d_axis = walls["dAxis"].as_float();
    p.thresh_dist = walls["threshDist"].as_float();

    const jsonlite::Value& div = root["divParams"];
    if (div.is_null()) {
        std::printf("ERROR: Cannot load division parameters\n");
        return -1;
    }
    p.use_div_plane_basis = div["useDivPlaneBasis"].as_int();
    p.div_plane_basis[0] = div["divPlaneBasisX"].as_float();
    p.div_plane_basis[1] = div["divPlaneBasisY"].as_float();
    p.div_plane_basis[2] = div["divPlaneBasisZ"].as_float();

    const jsonlite::Value& stf = root["stiffnessParams"];
    if (stf.is_null()) {
        std::printf("ERROR: Cannot load stiffness parameters\n");
        return -1;
    }
    p.use_different_stiffnesses = stf["useDifferentStiffnesses"].as_bool();
    p.stiffness2 = stf["softStiffFactor"].as_float() * p.youngs_mod;
    s.number_of_softer_cells = stf["numberOfSofterCells"].as_int();
    s.closeness_to_center = stf["closenessToCenter"].as_float();
    s.fraction_of_softer_cells = stf["fractionOfSofterCells"].as_float();
    s.during_growth = stf["duringGrowth"].as_bool();
    s.choose_random_cell_indices = stf["chooseRandomCellIndices"].as_bool();
    p.daught_same_stiffness = stf["daughtSameStiffness"].as_bool();
    p.recalc_r0 = stf["recalc_r0"].as_bool();

    const jsonlite::Value& box = root["boxParams"];
    if (box.is_null()) {
        std::printf("ERROR: Cannot load box parameters\n");
        return -1;
    }
    p.use_rigid_box = box["useRigidSimulationBox"].as_bool();
    s.use_pbcs = box["usePBCs"].as_bool();
    p.box_max[0] = box["box_len_x"].as_float();
    p.box_max[1] = box["box_len_y"].as_float();
    p.box_max[2] = box["box_len_z"].as_float();
    s.flatbox = box["flatbox"].as_bool();
    s.rand_pos = box["rand_pos"].as_bool();

    if (s.ran_z_offset == 0) s.z_offset = 0.0f;
    if (s.dt_tol > s.dt_max || s.dt_max <= 0 || s.dt_tol < 0) {
        std::printf("ERROR: Invalid time step parameters\nFATAL ERROR\n");
        std::exit(-60);
    }
    const jsonlite::Value& rnd = root["rand_params"];
    if (rnd.is_null()) {
        std::printf("ERROR: Cannot load randomness parameters\n");
        return -1;
    }
    p.add_rands = rnd["add_rands"].as_bool();
    s.rand_seed = rnd["rand_seed"].as_int();
    p.rand_scale_factor = rnd["rand_scale_factor"].as_float();
    if (rnd["rand_dist"].as_int() != 0) {
        std::printf("ERROR: rand_dist %d is not implemented (the reference implements 0 only)\n",
                    rnd["rand_dist"].as_int());
        return -1;
    }

    if (s.rad_frac < 0.4f || s.rad_frac > 0.8f) {
        std::printf("radFrac not in [0.4, 0.8] setting to 1.\n");
        s.count_only_internal = false;
    }
    if (s.traj_write_int == 0) s.traj_write_int = 1;
    if (s.new_cell_count_int == 0) s.new_cell_count_int = 1;
    if (s.traj_write_int > s.time_steps + s.equi_steps) {
        std::printf("Trajectory write interval is too large\n");
        return -1;
    }
    if (s.time_steps % s.traj_write_int != 0) {
        std::printf("Invalid trajectory write interval. Time steps must be divisible by it. \n");
        return -1;
    }
    if (s.new_cell_count_int > s.time_steps) {
        std::printf("New cell counting interval is too large. \n");
        return -1;
    }
    if (s.equi_steps <= 0) s.equi_steps = 0;
    if (p.do_pop_model != 1) p.do_pop_model = 0;
    if (p.max_pressure < 0) {
        std::printf("Invalid maximum pressure value of %f\nDisabling population modelling...", p.max_pressure);
        p.do_pop_model = 0;
    }
    if (!(s.closeness_to_center >= 0 && s.closeness_to_center <= 1)) {
        std::printf("ERROR: closenessToCenter is not in [0, 1]\nERROR: invalid input parameter\n");
        return -1;
    }
    if (s.use_walls && p.use_rigid_box) {
        std::printf("ERROR: Cannot use infinite XY walls and rigid simulation box simultaneously.\n"
                    "ERROR: Only use on or the other.\n");
        return -1;
    }
    if (s.fraction_of_softer_cells > 1.0f) {
        std::printf("ERROR: Softer cell fraction is > 1\n");
        return -1;
    }
    p.div_time_steps = s.time_steps;
    p.non_div_time_steps = s.equi_steps;
    return 0;
}

// Features of the reference's main that this driver does not provide; each is
// refused with a message rather than silently ignored.
int refuse_unsupported(const Settings& s) {
    const char* why = nullptr;
    if (s.do_adaptive_dt) why = "core.doAdaptive_dt = 1";
    else if (s.use_pbcs) why = "boxParams.usePBCs = 1 (unavailable in the reference as well)";
    if (!why) return 0;
    std::printf("ERROR: not supported by this driver: %s\n", why);
    return -1;
}

bool read_mesh(const std::string& dir, std::vector<float>& x, std::vector<float>& y, std::vector<float>& z,
               std::vector<int32_t>& nn, std::vector<int32_t>& sign, std::vector<int32_t>& faces) {
    auto path = [&](const char* f) { return dir.empty() || dir == "." ? std::string(f) : dir + "/" + f; };
    FILE* f = std::fopen(path("C180").c_str(), "r");  // GPUbounce.cu:1616-1626
    if (!f) {
        std::printf("Unable to open file C180\n");
        return false;
    }
    x.assign(NN, 0.f); y.assign(NN, 0.f); z.assign(NN, 0.f);
    for (int a = 0; a < NN; ++a)
        if (std::fscanf(f, "%f %f %f", &x[a], &y[a], &z[a]) != 3) {
            std::printf("   Unable to read file C180 on line %d\n", a + 1);
            std::fclose(f);
            return false;
        }
    std::fclose(f);
    f = std::fopen(path("C180NN").c_str(), "r");  // read_fullerene_nn, :1822-1838 (1-based)
    if (!f) {
        std::printf("Unable to open file C180NN\n");
        return false;
    }
    nn.assign(3 * S, 0);
    sign.assign(S, 0);
    for (int i = 0; i < NN; ++i) {
        int a, b, c, sg;
        if (std::fscanf(f, "%d, %d, %d, %d", &a, &b, &c, &sg) != 4) {
            std::printf("Unable to read line %d in file C180NN\n", i);
            std::fclose(f);
            return false;
        }
        nn[i] = a - 1; nn[S + i] = b - 1; nn[2 * S + i] = c - 1;
        sign[i] = sg;
    }
    std::fclose(f);
    f = std::fopen(path("C180C").c_str(), "r");  // :1840-1856, 270 bonds (read for validation)
    if (!f) {
        std::printf("Unable to open file C180C\n");
        return false;
    }
    for (int i = 0; i < 270; ++i) {
        int a, b;
        if (std::fscanf(f, "%d, %d", &a, &b) != 2) {
            std::printf("Unable to read line %d in file C180C\n", i);
            std::fclose(f);
            return false;
        }
    }
    std::fclose(f);
    f = std::fopen(path("C180_pentahexa").c_str(), "r");  // :1858-1886
    if (!f) {
        std::printf("Unable to open file C180_pentahexa\n");
        return false;
    }
    faces.assign(CS3D_FACES * 7, 0);
    for (int q = 0; q < 12; ++q) {
        int r[5];
        if (std::fscanf(f, "%d %d %d %d %d", &r[0], &r[1], &r[2], &r[3], &r[4]) != 5) {
            std::printf("Error: Unable to read line %d in file C180_pentahexa\n", q);
            std::fclose(f);
            return false;
        }
        const int row[7] = {r[0], r[1], r[2], r[3], r[4], r[0], r[0]};
        for (int k = 0; k < 7; ++k) faces[7 * q + k] = row[k];
    }
    for (int q = 0; q < 80; ++q) {
        int r[6];
        if (std::fscanf(f, "%d %d %d %d %d %d", &r[0], &r[1], &r[2], &r[3], &r[4], &r[5]) != 6) {
            std::printf("Error: Unable to read line %d in file C180_pentahexa\n", q);
            std::fclose(f);
            return false;
        }
        const int row[7] = {r[0], r[1], r[2], r[3], r[4], r[5], r[0]};
        for (int k = 0; k < 7; ++k) faces[84 + 7 * q + k] = row[k];
    }
    std::fclose(f);
    return true;
}

// initialize_C180s (GPUbounce.cu:1598-1774): the mesh centred on its node mean
// (float sums, divided by the double 180.0) and copied to the start positions.
int place_cells(const Settings& s, int n, int side, std::vector<float> ix, std::vector<float> iy,
                std::vector<float> iz, cs3d_ranmar_state* rng, std::vector<float>& X, std::vector<float>& Y,
                std::vector<float>& Z) {
    float sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < NN; ++i) { sx += ix[i]; sy += iy[i]; sz += iz[i]; }
    sx = (float)((double)sx / 180.0);
    sy = (float)((double)sy / 180.0);
    sz = (float)((double)sz / 180.0);
    for (int i = 0; i < NN; ++i) { ix[i] -= sx; iy[i] -= sy; iz[i] -= sz; }
    X.assign((size_t)s.max_cells * S, 0.f);
    Y.assign((size_t)s.max_cells * S, 0.f);
    Z.assign((size_t)s.max_cells * S, 0.f);
    const float L1 = 3.0f;  // lattice spacing (GPUbounce.cu:107)
    const float* bm = s.p.box_max;
    if (s.p.use_rigid_box) {
        float rcheck = powf(0.75 * (1.f / 3.14159) * 0.786, 1.f / 3.f);
        const float vol = bm[0] * bm[1] * bm[2];
        const int k = (int)std::floor(vol / 0.786);
        if (k < n) {
            std::fprintf(stderr, "ERROR: Simulation box is too small\n       Big enough for %d\n", k);
            return 27;
        }
        std::printf("Can fit upto %d cells\n", k);
        if (s.rand_pos) {
            std::vector<float> cx, cy, cz;
            while ((int)cx.size() < n) {
                float r[3];
                cs3d_ranmar(rng, r, 3);
                const float CMx = r[0] * (bm[0] - 1.f) + 1.f;
                const float CMy = r[1] * (bm[1] - 1.f) + 1.f;
                const float CMz = s.flatbox ? bm[2] / 2 : r[2] * (bm[2] - 1.f) + 1.f;
                bool far = !(CMx + rcheck > bm[0] || CMy + rcheck > bm[1] || CMz + rcheck > bm[2]);
                for (size_t q = 0; q < cx.size() && far; ++q) {
                    const float dx = cx[q] - CMx, dy = cy[q] - CMy, dz = cz[q] - CMz;
                    if (sqrtf(dx * dx + dy * dy + dz * dz) < 2 * rcheck) far = false;
                }
                if (far) { cx.push_back(CMx); cy.push_back(CMy); cz.push_back(CMz); }
            }
            for (int c = 0; c < n; ++c)
                for (int a = 0; a < NN; ++a) {
                    X[(size_t)c * S + a] = ix[a] + cx[c];
                    Y[(size_t)c * S + a] = iy[a] + cy[c];
                    Z[(size_t)c * S + a] = iz[a] + cz[c];
                }
        } else {
            const float ctr[3] = {0.5f * bm[0], 0.5f * bm[1], 0.5f * bm[2]};
            for (int c = 0; c < n; ++c) {
                const int ey = c % side, ex = c / side;
                for (int a = 0; a < NN; ++a) {
                    X[(size_t)c * S + a] = (float)(ix[a] + L1 * ex + 0.5 * L1 + ctr[0]);
                    Y[(size_t)c * S + a] = (float)(iy[a] + L1 * ey + 0.5 * L1 + ctr[1]);
                    Z[(size_t)c * S + a] = iz[a] + ctr[2];
                }
            }
        }
        for (int i = 0; i < n * S; ++i)
            if (X[i] > bm[0] || Y[i] > bm[1] || Z[i] > bm[2]) {
                std::printf("node outside the box\n%f %f %f\n", X[i], Y[i], Z[i]);
                break;
            }
    } else {
        for (int c = 0; c < n; ++c) {
            const int ey = c % side, ex = c / side;
            for (int a = 0; a < NN; ++a) {
                X[(size_t)c * S + a] = (float)(ix[a] + L1 * ex + 0.5 * L1);
                Y[(size_t)c * S + a] = (float)(iy[a] + L1 * ey + 0.5 * L1);
                Z[(size_t)c * S + a] = iz[a] + s.z_offset;
            }
        }
    }
    return 0;
}

// The host copy of the system centre, sysCMx/y/z (GPUbounce.cu:240, initial
// value 1): calc_sys_CM overwrites it; getRmax2 reads whatever it holds.
struct SysCentre {
    float x = 1.0f, y = 1.0f, z = 1.0f;
};

// calc_sys_CM (GPUbounce.cu:2504-2520): float sums in cell order, divided by the count
void calc_sys_cm(SysCentre& sc, const float* cx, const float* cy, const float* cz, int n) {
    sc.x = 0; sc.y = 0; sc.z = 0;
    for (int c = 0; c < n; ++c) { sc.x += cx[c]; sc.y += cy[c]; sc.z += cz[c]; }
    sc.x = sc.x / n; sc.y = sc.y / n; sc.z = sc.z / n;
}

// getRmax2 (GPUbounce.cu:2523-2535)
float get_rmax2(const SysCentre& sc, const float* cx, const float* cy, const float* cz, int n) {
    float r2 = 0;
    for (int c = 0; c < n; ++c) {
        const float dx = cx[c] - sc.x, dy = cy[c] - sc.y, dz = cz[c] - sc.z;
        r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
    }
    return r2;
}

// rGrowth of the step loop (GPUbounce.cu:1077-1088), from the cell count at the
// start of the step
float growth_rate_now(const cs3d_params& p, int n) {
    if (p.do_pop_model == 1) {
        float r = p.growth_rate * (1 - (n * 1.0 / p.max_pop));
        return r < 0 ? 0.f : r;
    }
    return p.growth_rate;
}

// Wall set-up (GPUbounce.cu:808-861): only the checks have an effect, since the
// contact kernel applies no wall force.  The axis must be z; the wall gap
// dAxis * 2 * (3 divVol / (4 pi))^(1/3) must not be smaller than the z spread the
// reference measures (its sort covers the first n entries of the z array only).
int check_walls(const Settings& s, int n, const std::vector<float>& Z) {
    if (s.use_walls != 1) return 0;
    if (s.perp_axis.empty() || (s.perp_axis[0] != 'Z' && s.perp_axis[0] != 'z')) {
        std::printf(" Invalid wall axis selection %s\nExiting...", s.perp_axis.c_str());
        return -1;
    }
    std::vector<float> tz(Z.begin(), Z.begin() + (size_t)n * S);
    std::sort(tz.begin(), tz.begin() + n);
    float radius = 3.0 * s.p.division_vol / 4.0;
    radius = radius / 3.14159;
    radius = std::pow(radius, 0.33333333333);
    const float gap = s.d_axis * 2 * radius;
    if (gap < (tz[n] - tz[0])) {
        std::printf("Distance between walls is too small\nExiting...");
        std::printf("Starting system size= %f\nWall gap = %f", tz[n] - tz[0], gap);
        return -1;
    }
    return 0;
}

// Softer cells chosen at set-up (GPUbounce.cu:615-692): within closenessToCenter of
// the largest centre distance, else numberOfSofterCells (or the fraction) cells,
// the first ones in a rigid box and RANMAR-drawn indices otherwise.  E holds one
// stiffness per cell slot.  Returns 12516 when more cells are asked for than exist.
int soft_cells_at_setup(const Settings& s, int n, const std::vector<float>& X, const std::vector<float>& Y,
                        const std::vector<float>& Z, cs3d_ranmar_state* rng, std::vector<float>& E) {
    E.assign(s.max_cells, s.p.stiffness1);
    if (!s.p.use_different_stiffnesses) return 0;
    std::printf("Making some cells softer... \n");
    if (s.closeness_to_center > 0.f && s.closeness_to_center < 1.f) {
        std::printf("Only making cells within %f of max radius softer\n", s.closeness_to_center);
        std::vector<float> cx(n), cy(n), cz(n);
        cs3d_cell_centres(X.data(), Y.data(), Z.data(), n, cx.data(), cy.data(), cz.data());
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int c = 0; c < n; ++c) { sx = sx + cx[c]; sy = sy + cy[c]; sz = sz + cz[c]; }
        const float inv = 1 / (float)n;  // float3 / float is (1/a)*b (inc/VectorFunctions.hpp:84-92)
        sx = inv * sx; sy = inv * sy; sz = inv * sz;
        std::printf("COM = (%f, %f, %f)\n", sx, sy, sz);
        float rmax = 0;
        std::vector<float> mags(n);
        for (int c = 0; c < n; ++c) {
            const float dx = cx[c] - sx, dy = cy[c] - sy, dz = cz[c] - sz;
            mags[c] = sqrtf(dx * dx + dy * dy + dz * dz);
            rmax = std::max(rmax, mags[c]);
        }
        int made = 0;
        for (int c = 0; c < n; ++c)
            if (mags[c] / rmax <= s.closeness_to_center) { E[c] = s.p.stiffness2; ++made; }
        std::printf("Made %d cells softer\n", made);
        return 0;
    }
    std::printf("Choosing softer cells randomly\n");
    int count = s.number_of_softer_cells;
    if (s.fraction_of_softer_cells > 0.f && s.fraction_of_softer_cells < 1.f)
        count = (int)std::round(s.fraction_of_softer_cells * (float)n);
    if (count > n) {
        std::printf("ERROR: Too many softer cells requested\n");
        return 12516;
    }
    if (s.p.use_rigid_box) {
        for (int c = 0; c < count; ++c) E[c] = s.p.stiffness2;
    } else {
        for (int c = 0; c < count; ++c) {
            float r;
            cs3d_ranmar(rng, &r, 1);
            const int ind = (int)std::round(r * n);  // may be n: a slot not yet in use
            if (ind < s.max_cells) E[ind] = s.p.stiffness2;
        }
    }
    std::printf("Made %d cells softer\n", count);
    return 0;
}

// correct_com at set-up (GPUbounce.cu:785-806): the mean of the cell centres is
// subtracted from every mesh node (the reference also shifts the 12 padding
// slots of each cell, which are no mesh nodes; they stay 0 here)
void correct_com_at_setup(int n, std::vector<float>& X, std::vector<float>& Y, std::vector<float>& Z,
                          SysCentre& sc) {
    std::vector<float> cx(n), cy(n), cz(n);
    cs3d_cell_centres(X.data(), Y.data(), Z.data(), n, cx.data(), cy.data(), cz.data());
    calc_sys_cm(sc, cx.data(), cy.data(), cz.data(), n);
    for (int c = 0; c < n; ++c)
        for (int a = 0; a < NN; ++a) {
            const size_t g = (size_t)c * S + a;
            X[g] = X[g] - sc.x; Y[g] = Y[g] - sc.y; Z[g] = Z[g] - sc.z;
        }
}

struct Output {
    FILE* ascii = nullptr;
    cs3d_traj* bin = nullptr;
    FILE* forces = nullptr;
};

bool get_field(cs3d_ctx* ctx, cs3d_sys sys, cs3d_field f, void* host, size_t bytes) {
    const cs3d_status st = cs3d_get(ctx, sys, f, host, bytes);
    if (st != CS3D_OK) {
        std::printf("ERROR: cs3d_get(%d): %s (%s)\n", (int)f, cs3d_status_str(st), cs3d_last_error());
        return false;
    }
    return true;
}

// One output frame copied off the device: positions (and stiffness types), and
// for the forces CSV the conservative force, velocity, pressure and volume.
struct Frame {
    int step = 0, frame = 0, n = 0;
    bool positions = true;  // false: forces CSV only (step 0, written after frame 0)
    std::vector<float> X, Y, Z, E, FX, FY, FZ, VX, VY, VZ, P, V;
};

// write_traj (GPUbounce.cu:2412-2444) / WriteBinaryTraj (:2446-2477) of one frame;
// with variable stiffness each cell carries its type (0/H: stiffness1, 1/C: other)
bool write_positions(const Settings& s, Output& out, int step, int frame, int n, const std::vector<float>& X,
                     const std::vector<float>& Y, const std::vector<float>& Z, const std::vector<float>& E) {
    const bool var = s.p.use_different_stiffnesses != 0;
    if (out.bin) {
        std::vector<int32_t> types;
        if (var)
            for (int c = 0; c < n; ++c) types.push_back(E[c] == s.p.stiffness1 ? 0 : 1);
        if (cs3d_traj_append(out.bin, step, frame, n, X.data(), Y.data(), Z.data(), var ? types.data() : nullptr) !=
            CS3D_OK)
            return false;
    } else if (out.ascii) {
        std::fprintf(out.ascii, "%d\n", n * 192);
        std::fprintf(out.ascii, "Step: %d frame: %d\n", step, step / s.traj_write_int);
        for (int c = 0; c < n; ++c) {
            if (!var) std::fprintf(out.ascii, "cell: %d\n", c);
            else if (E[c] == s.p.stiffness1) std::fprintf(out.ascii, "cell: %d H\n", c);
            else if (E[c] == s.p.stiffness2) std::fprintf(out.ascii, "cell: %d C\n", c);
            else std::fprintf(out.ascii, "cell: %d UnknownStiffness\n", c);
            for (int q = 0; q < S; ++q)
                std::fprintf(out.ascii, "%.7f,  %.7f,  %.7f\n", X[(size_t)c * S + q], Y[(size_t)c * S + q],
                             Z[(size_t)c * S + q]);
        }
    }
    return true;
}

// writeForces (GPUbounce.cu:2567-2598): total conservative force, velocity,
// position, pressure and volume of every node
void write_forces(Output& out, const Frame& f) {
    if (!out.forces) return;
    for (int c = 0; c < f.n; ++c)
        for (int q = 0; q < NN; ++q) {
            const size_t g = (size_t)c * S + q;
            const float fm = sqrtf(f.FX[g] * f.FX[g] + f.FY[g] * f.FY[g] + f.FZ[g] * f.FZ[g]);
            const float vm = sqrtf(f.VX[g] * f.VX[g] + f.VY[g] * f.VY[g] + f.VZ[g] * f.VZ[g]);
            std::fprintf(out.forces,
                         "%d,%d,%d,%d,%d,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f,%.8f\n", f.step,
                         f.n, c, q, c * 192 + q, f.FX[g], f.FY[g], f.FZ[g], fm, f.VX[g], f.VY[g], f.VZ[g], vm, f.X[g],
                         f.Y[g], f.Z[g], f.P[c], f.V[c]);
        }
}

bool write_out(const Settings& s, Output& out, const Frame& f) {
    if (f.positions && !write_positions(s, out, f.step, f.frame, f.n, f.X, f.Y, f.Z, f.E)) return false;
    write_forces(out, f);
    return true;
}

// Copies one frame's fields to the host (the device work stays queued behind it).
bool fetch_frame(cs3d_ctx* ctx, cs3d_sys sys, const Settings& s, const Output& out, int step, int frame, int n,
                 Frame& f) {
    f.step = step; f.frame = frame; f.n = n; f.positions = true;
    const size_t nn = (size_t)n * S;
    auto get = [&](cs3d_field fld, std::vector<float>& v, size_t count) {
        v.resize(std::max<size_t>(count, 1));
        return count == 0 || get_field(ctx, sys, fld, v.data(), count * 4);
    };
    bool ok = get(CS3D_F_X, f.X, nn) && get(CS3D_F_Y, f.Y, nn) && get(CS3D_F_Z, f.Z, nn);
    if (ok && s.p.use_different_stiffnesses) ok = get(CS3D_F_STIFFNESS, f.E, n);
    if (ok && out.forces)
        ok = get(CS3D_F_FCX, f.FX, nn) && get(CS3D_F_FCY, f.FY, nn) && get(CS3D_F_FCZ, f.FZ, nn) &&
             get(CS3D_F_VX, f.VX, nn) && get(CS3D_F_VY, f.VY, nn) && get(CS3D_F_VZ, f.VZ, nn) &&
             get(CS3D_F_PRESSURE, f.P, n) && get(CS3D_F_VOLUME, f.V, n);
    return ok;
}

// Output on a writer thread with two frame buffers: the main thread copies frame
// k + 1 off the device and steps on while frame k is formatted and written; frames
// reach the files in order.  With sync = true every frame is written in place.
class FrameWriter {
  public:
    FrameWriter(const Settings& s, Output& out, bool sync) : s_(s), out_(out), sync_(sync) {
        if (!sync_) th_ = std::thread([this] { run(); });
    }
    ~FrameWriter() { finish(); }
    // takes the frame's buffers; blocks while the previous frame is still pending
    void submit(Frame&& f) {
        if (sync_) {
            if (!write_out(s_, out_, f)) failed_ = true;
            return;
        }
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return !has_; });
        next_ = std::move(f);
        has_ = true;
        cv_.notify_all();
    }
    // waits until every submitted frame is written; true when all writes succeeded
    bool finish() {
        if (!sync_ && th_.joinable()) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                stop_ = true;
            }
            cv_.notify_all();
            th_.join();
        }
        return !failed_;
    }

  private:
    void run() {
        for (;;) {
            Frame f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return has_ || stop_; });
                if (!has_) return;
                f = std::move(next_);
                has_ = false;
            }
            cv_.notify_all();
            if (!write_out(s_, out_, f)) failed_ = true;
        }
    }
    const Settings& s_;
    Output& out_;
    bool sync_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    Frame next_;
    bool has_ = false, stop_ = false;
    std::atomic<bool> failed_{false};
};

bool get_centres(cs3d_ctx* ctx, cs3d_sys sys, int n, std::vector<float>& cx, std::vector<float>& cy,
                 std::vector<float>& cz) {
    cx.resize(std::max(n, 1)); cy.resize(std::max(n, 1)); cz.resize(std::max(n, 1));
    return n == 0 || (get_field(ctx, sys, CS3D_F_CMX, cx.data(), (size_t)n * 4) &&
                      get_field(ctx, sys, CS3D_F_CMY, cy.data(), (size_t)n * 4) &&
                      get_field(ctx, sys, CS3D_F_CMZ, cz.data(), (size_t)n * 4));
}

// num_cells_far (GPUbounce.cu:2537-2565): of the num_div newest cells, those whose
// centre lies beyond radius_cutoff of the largest centre distance (a daughter's
// centre is still 0 in the step that made it)
bool num_cells_far(cs3d_ctx* ctx, cs3d_sys sys, const Settings& s, int n, int num_div, SysCentre& sc, int& far) {
    far = 0;
    if (num_div == 0 || n < 50) return true;
    std::vector<float> cx, cy, cz;
    if (!get_centres(ctx, sys, n, cx, cy, cz)) return false;
    calc_sys_cm(sc, cx.data(), cy.data(), cz.data(), n);
    const float rmax2 = get_rmax2(sc, cx.data(), cy.data(), cz.data(), n);
    for (int c = n - num_div; c < n; ++c) {
        const float dx = cx[c] - sc.x, dy = cy[c] - sc.y, dz = cz[c] - sc.z;
        const float dr2 = dx * dx + dy * dy + dz * dz;
        if (dr2 > s.rad_frac * s.rad_frac * rmax2) ++far;
    }
    return true;
}

// Softer cells chosen when growth ends (GPUbounce.cu:1340-1429, useDifferentStiffnesses
// without duringGrowth): distinct RANMAR-drawn indices, or the first cells (in index
// order) within (1 - closenessToCenter) of the largest centre distance.  The
// largest distance is taken about the system centre held from the last
// calc_sys_CM, as the reference does, before the centre is recomputed.
int soft_cells_after_growth(cs3d_ctx* ctx, cs3d_sys sys, Settings& s, int n, SysCentre& sc) {
    std::printf("Now making some cells softer...\n");
    if (s.fraction_of_softer_cells > 0) s.number_of_softer_cells = (int)roundf(s.fraction_of_softer_cells * n);
    const int want = s.number_of_softer_cells;
    std::printf("Will make %d cells softer\n", want);
    std::vector<float> E(std::max(n, 1));
    if (n > 0 && !get_field(ctx, sys, CS3D_F_STIFFNESS, E.data(), (size_t)n * 4)) return -1;
    if (s.choose_random_cell_indices) {
        if (want > n + 1) {  // the reference's draw loop would never end
            std::printf("ERROR: %d distinct softer cells requested from %d indices\n", want, n + 1);
            return -1;
        }
        cs3d_ranmar_state rng;
        if (!get_field(ctx, sys, CS3D_F_RANMAR, &rng, sizeof(rng))) return -1;
        std::vector<int> chosen;
        std::printf("Make cells with indices ");
        while ((int)chosen.size() < want) {
            float r;
            cs3d_ranmar(&rng, &r, 1);
            const int ind = (int)roundf(r * n);
            if (std::find(chosen.begin(), chosen.end(), ind) == chosen.end()) {
                chosen.push_back(ind);
                std::printf("%d ", ind);
            }
        }
        for (int ind : chosen)
            if (ind < n) E[ind] = s.p.stiffness2;  // index n names no cell (the copy covers n cells)
        if (cs3d_set(ctx, sys, CS3D_F_RANMAR, &rng, sizeof(rng)) != CS3D_OK) return -1;
    } else {
        std::vector<float> cx, cy, cz;
        if (!get_centres(ctx, sys, n, cx, cy, cz)) return -1;
        const float rmax2 = get_rmax2(sc, cx.data(), cy.data(), cz.data(), n);
        calc_sys_cm(sc, cx.data(), cy.data(), cz.data(), n);
        const float f = 1 - s.closeness_to_center;
        std::printf("Made cells with indices ");
        int made = 0;
        for (int c = 0; made < want && c < n; ++c) {
            const float dx = cx[c] - sc.x, dy = cy[c] - sc.y, dz = cz[c] - sc.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= f * f * rmax2) {
                std::printf("%d, ", c);
                ++made;
                E[c] = s.p.stiffness2;
            }
        }
    }
    if (n > 0 && cs3d_set(ctx, sys, CS3D_F_STIFFNESS, E.data(), (size_t)n * 4) != CS3D_OK) return -1;
    std::printf("\b\b softer\n");
    return 0;
}

// mitotic-index file (GPUbounce.cu:1515-1547): one line per counting interval
// (Time_steps / cell_count_int + 1 lines), new cells over the cells that did not
// divide, 0 where either count is 0.  The reference reads its last line one past
// the end of its arrays; it is 0 here.
bool write_mit_index(const Settings& s, const std::vector<int>& dividing, const std::vector<int>& total) {
    FILE* f = std::fopen(s.mit_ind_file.c_str(), s.overwrite_mit_ind ? "w" : "a");
    if (!f) {
        std::printf("Failed to open mit-index.dat\n");
        std::exit(1);
    }
    for (size_t i = 0; i < dividing.size(); ++i) {
        if (dividing[i] != 0 && total[i] != 0) std::fprintf(f, "%f\n", (float)dividing[i] / total[i]);
        else std::fprintf(f, "%f\n", 0.0);
    }
    std::fclose(f);
    return true;
}

bool write_bytes(const std::string& path, const void* data, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
    std::printf("celldiv (C180 cell mechanics on AMD Instinct, libcs3d %s)\n", cs3d_version());
    if (argc < 4) {
        std::printf("Usage: celldiv no_of_cells inpFile.json gpuid [--grid-interval K] [--pair-skin S] "
                    "[--mesh-dir DIR] [--no-graph] [--daughter-com-fix] [--setup-only] [--sync-output] "
                    "[--dump-setup PREFIX] [--save-state FILE]\n");
        return 0;
    }
    const int n_cells = std::atoi(argv[1]);
    const int gpuid = std::atoi(argv[3]);
    int grid_interval = 1;
    float pair_skin = -1.f;
    std::string mesh_dir = ".", dump_setup, save_state;
    bool graph = true, com_fix = false, setup_only = false, sync_output = false;
    for (int a = 4; a < argc; ++a) {
        const std::string o = argv[a];
        if (o == "--grid-interval" && a + 1 < argc) grid_interval = std::atoi(argv[++a]);
        else if (o == "--pair-skin" && a + 1 < argc) pair_skin = (float)std::atof(argv[++a]);
        else if (o == "--mesh-dir" && a + 1 < argc) mesh_dir = argv[++a];
        else if (o == "--no-graph") graph = false;
        else if (o == "--daughter-com-fix") com_fix = true;
        else if (o == "--setup-only") setup_only = true;
        else if (o == "--sync-output") sync_output = true;
        else if (o == "--dump-setup" && a + 1 < argc) dump_setup = argv[++a];
        else if (o == "--save-state" && a + 1 < argc) save_state = argv[++a];
        else {
            std::printf("unknown option %s\n", o.c_str());
            return -1;
        }
    }
    Settings s;
    if (read_params(argv[2], s) != 0) return -1;
    std::printf("%d\n", s.max_cells);
    const int side = (int)(std::sqrt((double)n_cells) + 0.5);
    if (n_cells > s.max_cells || n_cells < 1) {
        std::printf("Usage: Celldiv no_of_threads\n       no_of_threads should be a square, n^2, < %d\n",
                    s.max_cells);
        return 0;
    }
    if (refuse_unsupported(s) != 0) return -1;
    s.p.grid_interval = grid_interval;
    if (pair_skin >= 0) s.p.pair_skin = pair_skin;
    s.p.daughter_com_fix = com_fix ? 1 : 0;

    // generate_random (GPUbounce.cu:1777-1810): RANMAR seeds
    int ij, kl;
    if (s.rand_seed < 0) {
        const time_t now = std::time(nullptr);
        ij = (int)(now % 31328);
        kl = (int)(now % 30081);
    } else {
        ij = s.rand_seed;
        kl = s.rand_seed * 100;
    }
    s.p.ranmar_ij = ij;
    s.p.ranmar_kl = kl;
    // random force: the reference seeds cuRAND MT19937 with rand_seed or the time
    // (GPUbounce.cu:714-728); this engine keys its counter-based stream the same way
    s.p.noise_seed = s.rand_seed > 0 ? (uint32_t)s.rand_seed : (uint32_t)std::time(nullptr);
    if (s.p.add_rands)
        std::printf("random force on: Philox4x32-10 normals (not the reference's cuRAND stream), scale %g\n",
                    s.p.rand_scale_factor);
    cs3d_ranmar_state rng;
    if (cs3d_ranmar_init(&rng, ij, kl) != CS3D_OK) {
        std::printf("RANMAR seeds out of range: %d, %d\n", ij, kl);
        std::exit(1);  // rmarin's exit (marsaglia.h:79-83)
    }
    std::printf("RNG seeds: %d, %d\n", ij, kl);
    if (cs3d_params_check(&s.p) != CS3D_OK) {
        std::printf("ERROR: %s\n", cs3d_last_error());
        return -1;
    }

    std::vector<float> ix, iy, iz;
    std::vector<int32_t> nn, sign, faces;
    if (!read_mesh(mesh_dir, ix, iy, iz, nn, sign, faces)) return -1;
    std::vector<float> X, Y, Z;
    const int rc = place_cells(s, n_cells, side, ix, iy, iz, &rng, X, Y, Z);
    if (rc != 0) return rc;
    cs3d_mesh mesh;
    if (cs3d_mesh_build(nn.data(), sign.data(), faces.data(), X.data(), Y.data(), Z.data(), &mesh) != CS3D_OK) {
        std::printf("ERROR: mesh: %s\n", cs3d_last_error());
        return -1;
    }
    std::vector<float> E;
    const int rs = soft_cells_at_setup(s, n_cells, X, Y, Z, &rng, E);
    if (rs != 0) return rs;
    SysCentre sc;
    if (s.correct_com) correct_com_at_setup(n_cells, X, Y, Z, sc);
    s.p.correct_com = s.correct_com ? 1 : 0;
    if (check_walls(s, n_cells, Z) != 0) return -1;

    // --dump-setup: the parameter block, mesh tables and initial positions exactly as
    // handed to the engine (cs3d_params and cs3d_mesh bytes; x, y and z of the placed
    // cells as float32, n_cells * 192 each), for tools that re-run the same system
    const size_t pos_bytes = (size_t)n_cells * S * sizeof(float);
    if (!dump_setup.empty() && (!write_bytes(dump_setup + ".params", &s.p, sizeof(s.p)) ||
                                !write_bytes(dump_setup + ".mesh", &mesh, sizeof(mesh)) ||
                                !write_bytes(dump_setup + ".x", X.data(), pos_bytes) ||
                                !write_bytes(dump_setup + ".y", Y.data(), pos_bytes) ||
                                !write_bytes(dump_setup + ".z", Z.data(), pos_bytes))) {
        std::printf("ERROR: cannot write %s.params / .mesh / .x / .y / .z\n", dump_setup.c_str());
        return -1;
    }

    // outputs, GPUbounce.cu:332-335, 768-778, 1007-1048
    if (FILE* f = std::fopen("psfil", "w")) std::fclose(f);
    Output out;
    if (s.binary) {
        const int max_frames = (s.time_steps + s.equi_steps + 1) / s.traj_write_int;
        if (cs3d_traj_open(s.traj_file.c_str(), s.max_cells, s.p.use_different_stiffnesses, max_frames, &out.bin) !=
            CS3D_OK) {
            std::printf("Failed to open %s \n", s.traj_file.c_str());
            return -1;
        }
    } else {
        out.ascii = std::fopen(s.traj_file.c_str(), "w");
        if (!out.ascii) {
            std::printf("Failed to open %s \n", s.traj_file.c_str());
            return -1;
        }
        std::fprintf(out.ascii, "Header Start:\nMaximum number of cells:\n%d\n", s.max_cells);
        std::fprintf(out.ascii, "Using variable stiffness:\n%s\n", s.p.use_different_stiffnesses ? "True" : "False");
        std::fprintf(out.ascii, "Maximum number of frames:\n%d\n", (s.time_steps + s.equi_steps + 1) / s.traj_write_int);
        std::fprintf(out.ascii, "Header End\n");
    }
    if (s.write_cont_force) {
        out.forces = std::fopen(s.forces_file.c_str(), "w");
        if (out.forces)
            std::fprintf(out.forces, "step,num_cells,cell_ind,node_ind,glob_node_ind,FX,FY,FZ,F,VX,VY,VZ,V,X,Y,Z,P,Vol\n");
    }
    if (FILE* f = std::fopen("velocity2.xyz", "w")) std::fclose(f);
    // frame 0: the start positions (the device start-up does not move nodes)
    int frame = 1;
    if (!write_positions(s, out, 0, frame, n_cells, X, Y, Z, E)) return -1;
    if (setup_only) {  // parameters, mesh, placement and frame 0 only; no device
        if (out.bin) cs3d_traj_close(out.bin, n_cells);
        if (out.ascii) std::fclose(out.ascii);
        if (out.forces) std::fclose(out.forces);
        std::printf("setup done: %d cells placed\n", n_cells);
        return 0;
    }

    // device pools: voxel lists for at least the 1024*1024/4 counters the reference
    // clears (GPUbounce.cu:1329), search bins sized like the engine's default
    const int64_t vcells = std::max<int64_t>(1 << 18, 32 * (int64_t)s.max_cells);
    const int64_t vwords = vcells * (1 + s.p.voxel_capacity);
    const int64_t bwords = std::max<int64_t>(1 << 22, (int64_t)s.max_cells * 640 * (1 + 4 * CS3D_BIN_CAP));
    cs3d_ctx* ctx = nullptr;
    cs3d_status st = cs3d_ctx_create(gpuid, &mesh, s.max_cells, vwords, bwords, &ctx);
    if (st != CS3D_OK) {
        std::printf("Could not set to divice %d: %s\n", gpuid, cs3d_last_error());
        return -1;
    }
    cs3d_set_graph(ctx, graph ? 1 : 0);
    cs3d_sys sys = -1;
    st = cs3d_system_add(ctx, &s.p, n_cells, s.max_cells, vcells, X.data(), Y.data(), Z.data(), nullptr, nullptr,
                         nullptr, &sys);
    if (st != CS3D_OK) {
        std::printf("ERROR: system: %s (%s)\n", cs3d_status_str(st), cs3d_last_error());
        cs3d_ctx_destroy(ctx);
        return -1;
    }
    // the RANMAR stream continues where placement and soft-cell selection left it
    // (one stream, as in the reference)
    cs3d_set(ctx, sys, CS3D_F_RANMAR, &rng, sizeof(rng));
    if (s.p.use_different_stiffnesses && n_cells > 0) {
        // soft cells chosen at set-up are on the device before the initial forces
        // (GPUbounce.cu:691 uploads d_Youngs_mod before the start-up CalculateConForce at :978-1004)
        st = cs3d_set(ctx, sys, CS3D_F_STIFFNESS, E.data(), (size_t)n_cells * 4);
        if (st == CS3D_OK && std::any_of(E.begin(), E.begin() + n_cells, [&](float e) { return e != s.p.stiffness1; }))
            st = cs3d_system_startup(ctx, sys);
        if (st != CS3D_OK) {
            std::printf("ERROR: soft cells at set-up: %s (%s)\n", cs3d_status_str(st), cs3d_last_error());
            cs3d_ctx_destroy(ctx);
            return -1;
        }
    }

    const int total = s.time_steps + s.equi_steps;
    const int I = s.new_cell_count_int;
    // growth ends at the first step past Time_steps + 1 (GPUbounce.cu:1340)
    const int growth_end = s.time_steps + 2;
    const bool select_after_growth = s.p.use_different_stiffnesses && !s.during_growth && growth_end <= total;
    std::vector<int> dividing(s.time_steps / I + 1, 0), totals(s.time_steps / I + 1, 0);
    int window_new = 0;  // cells counted since the last counting step
    bool growth_done = false;
    // stops: trajectory frames; k*I - 1 and k*I (the counting step needs the count
    // at its start); every growth step when only internal cells are counted; the
    // step before growth ends and the one that ends it
    auto next_stop = [&](int at) {
        int nx = std::min(total, (at / s.traj_write_int + 1) * s.traj_write_int);
        if (at < s.time_steps) {
            if (s.count_only_internal) nx = std::min(nx, at + 1);
            const int b = (at / I + 1) * I;
            if (b <= s.time_steps) nx = std::min(nx, b - 1 > at ? b - 1 : b);
        }
        if (select_after_growth && at < growth_end) nx = std::min(nx, at < growth_end - 1 ? growth_end - 1 : growth_end);
        return nx;
    };
    int n = n_cells;
    int rcode = 0;
    FrameWriter writer(s, out, sync_output);
    if (out.forces) {  // the forces of the start-up sequence, step 0
        Frame f0;
        if (!fetch_frame(ctx, sys, s, out, 0, frame, n, f0)) rcode = -1;
        f0.positions = false;
        if (rcode == 0) writer.submit(std::move(f0));
    }
    int step = 0;
    while (rcode == 0 && step < total) {
        const int next = next_stop(step);
        const int n_start = n;
        st = cs3d_step(ctx, sys, next - step);
        if (st == CS3D_OK) st = cs3d_sync(ctx);
        if (st != CS3D_OK) {
            std::printf("Error: %s (%s)\n", cs3d_status_str(st), cs3d_last_error());
            rcode = -1;
            break;
        }
        int32_t ncur = 0;
        int64_t done = 0;
        uint32_t status = 0;
        cs3d_system_info(ctx, sys, &ncur, &done, &status);
        step = (int)done;
        n = ncur;
        if (status & CS3D_SYS_CAPACITY) {
            std::printf("ERROR: Population is %d, only allocated enough memory for %d\n"
                        "ERROR: Fatal error, crashing...\n", n + 1, s.max_cells);
            rcode = -69;
            break;
        }
        if (status & CS3D_SYS_FATAL_MASK) {
            std::printf("Device failure at step %d: system status 0x%x\n", step, status);
            rcode = 0;
            break;
        }
        if (step < next) {
            std::printf("Error: the system stopped at step %d\n", step);
            rcode = -1;
            break;
        }
        if (step <= s.time_steps) {  // mitotic-index counting (GPUbounce.cu:1288-1302)
            int born = n - n_start;
            if (s.count_only_internal && born > 0 && growth_rate_now(s.p, n_start) > 0) {
                int far = 0;
                if (!num_cells_far(ctx, sys, s, n, born, sc, far)) { rcode = -1; break; }
                born -= far;
            }
            window_new += born;
            // the counting step runs only inside the division block (rGrowth > 0)
            if (step % I == 0) {
                if (growth_rate_now(s.p, n_start) > 0) {
                    dividing[(step - 1) / I] = window_new;
                    totals[(step - 1) / I] = n - window_new;
                }
                window_new = 0;
            }
        }
        if (select_after_growth && step == growth_end - 1 && s.correct_com) {
            // the system centre of the last correct_com, read by getRmax2 below
            float c3[3];
            if (!get_field(ctx, sys, CS3D_F_SYSCM, c3, sizeof(c3))) { rcode = -1; break; }
            sc.x = c3[0]; sc.y = c3[1]; sc.z = c3[2];
        }
        if (!growth_done && step >= growth_end) {
            std::printf("Cell growth halted.\nProceeding with MD simulation without growth...\n");
            growth_done = true;
        }
        if (select_after_growth && step == growth_end && soft_cells_after_growth(ctx, sys, s, n, sc) != 0) {
            rcode = -1;
            break;
        }
        if (step % s.traj_write_int == 0) {
            ++frame;
            Frame f;
            if (!fetch_frame(ctx, sys, s, out, step, frame, n, f)) rcode = -1;
            else writer.submit(std::move(f));
        }
    }
    if (!writer.finish()) {
        std::printf("ERROR: writing the output files failed\n");
        if (rcode == 0) rcode = -1;
    }
    if (out.bin) cs3d_traj_close(out.bin, n);
    if (out.ascii) std::fclose(out.ascii);
    if (out.forces) std::fclose(out.forces);
    // --save-state: the system's complete state after the last step (cs3d_system_save)
    if (rcode == 0 && !save_state.empty() && cs3d_system_save(ctx, sys, save_state.c_str()) != CS3D_OK) {
        std::printf("ERROR: saving the state: %s\n", cs3d_last_error());
        rcode = -1;
    }
    cs3d_ctx_destroy(ctx);
    if (rcode == 0) write_mit_index(s, dividing, totals);
    std::printf("%d cells after %d steps\n", n, step);
    return rcode;
}
