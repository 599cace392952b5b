/*
 * cs3d_oracle.h - CPU oracle of the CellSim3D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a C++/OpenMP restatement of the
 * reference's kernel bodies (SoftSimu/CellSim3D, the src/ .cu files) in their operation
 * order, used as the correctness checker for the HIP engine and as the CPU
 * baseline of bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the engine (libcs3d.so) never does.
 *
 * Modes: the float ("reference-exact") mode reproduces the reference's FP32
 * arithmetic including its double-literal promotions; the double mode runs the
 * same algorithm in FP64 and is used to state the FP32 error.
 *
 * Canonical orders where the reference is order-non-deterministic:
 *   - voxel lists are ascending in cell index (reference: atomicAdd order,
 *     propagatebound.cu:247);
 *   - the per-cell volume and area sums use a 128-slot pairwise tree over the
 *     92 faces, faces >= 92 contributing 0 (reference: shared-memory atomicAdd,
 *     volume.cu:120-124).
 */
#ifndef CS3D_ORACLE_H
#define CS3D_ORACLE_H
#include <stdint.h>
#include "../include/cs3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RANMAR (marsaglia.h:54-144) and plane normals (RandomVector.h:8-78). */
int or_ranmar_init(cs3d_ranmar_state* st, int ij, int kl);
void or_ranmar(cs3d_ranmar_state* st, float* out, int n);
void or_random_vector(cs3d_ranmar_state* st, float* n3);
void or_random_vector_basis(cs3d_ranmar_state* st, float* n3, const float* basis3);

/* The double acos of the angle force (propagate.cu:36), as the device kernels
   compute it (cellsim3d_amd/csrc/cs3d_acos.h), over n arguments. */
void or_acos(const double* x, double* y, int n);

/* R0 (GPUbounce.cu:1892-1908) and theta0 (GPUbounce.cu:909-938) of one placed cell. */
void or_mesh_rest(const int32_t* nn, const float* x, const float* y, const float* z,
                  float* R0, float* theta0);

/* Per-kernel functions, float reference-exact mode. Node arrays have 192-slot
 * stride and n cells. Return values are CS3D_SYS_* status bits. */
void or_com(int n, const float* X, const float* Y, const float* Z, float* CMx, float* CMy, float* CMz);
uint32_t or_volumes(int n, const cs3d_params* p, const int32_t* faces, const float* X, const float* Y,
                    const float* Z, const float* CMx, const float* CMy, const float* CMz,
                    const float* stiffness, int step, float* vol, float* area, int8_t* flag);
void or_bboxes(int n, const float* X, const float* Y, const float* Z, float* bbox);
void or_extent(int n, const float* bbox, float* out6);
/* grid[8]: minx, miny, minz as float bits; xdiv, ydiv, zdiv, budget, capacity. */
uint32_t or_grid_build(int n, const cs3d_params* p, const float* bbox, int32_t budget, int32_t* grid,
                       int32_t* counts, int32_t* lists);
uint32_t or_con_force(int n, const cs3d_params* p, const cs3d_mesh* m, const float* X, const float* Y,
                      const float* Z, const float* CMx, const float* CMy, const float* CMz,
                      const float* P, const float* stiffness, const float* bbox, const int32_t* grid,
                      const int32_t* counts, const int32_t* lists, int bonded_only, float* FX, float* FY,
                      float* FZ);
uint32_t or_dis_force(int n, const cs3d_params* p, const cs3d_mesh* m, const float* X, const float* Y,
                      const float* Z, const float* VX, const float* VY, const float* VZ, const float* bbox,
                      const int32_t* grid, const int32_t* counts, const int32_t* lists, float* FX,
                      float* FY, float* FZ);
void or_integrate(int n, const cs3d_params* p, float* X, float* Y, float* Z, float* VX, float* VY, float* VZ,
                  const float* FCX, const float* FCY, const float* FCZ, const float* FDX, const float* FDY,
                  const float* FDZ);
void or_velocity_ab(int n, const cs3d_params* p, float* VX, float* VY, float* VZ, const float* FCX,
                    const float* FCY, const float* FCZ, const float* FDX, const float* FDY, const float* FDZ);
float or_rgrowth(const cs3d_params* p, int n_cells);
void or_pressure_update(int n, const cs3d_params* p, float* P, const float* stiffness, const float* scale,
                        int step, float rgrowth);

/* Full simulation state (reference step loop, GPUbounce.cu:1055-1506). */
typedef struct or_sim or_sim;
/* precision 0 = float reference-exact, 1 = double. */
or_sim* or_sim_create(const cs3d_params* p, const cs3d_mesh* m, int n, int max_cells, int32_t budget,
                      const float* X, const float* Y, const float* Z, int precision);
void or_sim_destroy(or_sim* s);
/* Start-up sequence of main (GPUbounce.cu:739-1004). */
uint32_t or_sim_init(or_sim* s);
/* n steps; returns accumulated status bits. */
uint32_t or_sim_step(or_sim* s, int nsteps);
int or_sim_ncells(const or_sim* s);
int64_t or_sim_stepno(const or_sim* s);
/* Field access with the cs3d_field numbering; floats are converted from the
 * internal precision. Returns bytes copied (0 on unknown field). */
size_t or_sim_get(const or_sim* s, int field, void* out, size_t bytes);
size_t or_sim_set(or_sim* s, int field, const void* in, size_t bytes);
/* Double-precision node field access (field = X..VAZ), for error statements. */
size_t or_sim_get_f64(const or_sim* s, int field, double* out, size_t count);
/* Sets live cell count and step counter (state injection). */
void or_sim_set_counts(or_sim* s, int n, int64_t step);
/* Number of OpenMP threads the oracle uses (0 = runtime default). */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
