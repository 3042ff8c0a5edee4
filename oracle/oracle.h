/*
 * CPU oracle for the nonlinear magnetostatic solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is an FP64 CPU restatement of the
 * reference's host algorithms, used as the checker for the MI355X library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef FEMM_ORACLE_H
#define FEMM_ORACLE_H

#include "femm_model.h"

#ifdef __cplusplus
extern "C" {
#endif

/* B-H spline (nonlinear_model.cpp:13-113).  b,h: n raw knots.
 * Outputs (n each): processed knots and slopes; cold secant; smoothing passes.
 * Returns 1 on success. */
int oracle_bh_spline(const double *b, const double *h, int n, double *out_b, double *out_h,
                     double *out_slope, double *cold_secant, int *smoothing_passes);

/* B-H evaluation (nonlinear_model.cpp:115-148): out[4] = {H, dH/dB, nu, dnu/dB^2}. */
void oracle_bh_eval(const double *b, const double *h, const double *slope, int n, double field_b,
                    double out[4]);

/* Model validation (nonlinear_model.cpp:223-319). */
int oracle_validate(const femm_model_t *model);

/* Reduced Newton system from the host assembler (nonlinear_model.cpp:453-627 and
 * :321-394).  Handle-based: create, query sizes, copy out, destroy. */
typedef struct oracle_assembly oracle_assembly;
int oracle_assemble(const femm_model_t *model, const double *a_old, const double *currents,
                    int use_newton, oracle_assembly **out);
void oracle_assembly_sizes(const oracle_assembly *assembly, int32_t *free_count, int32_t *nnz);
void oracle_assembly_copy(const oracle_assembly *assembly, int32_t *row_offsets,
                          int32_t *column_indices, double *values, double *diagonal,
                          double *rhs_per_amp, double *rhs_offset, int32_t *free_nodes,
                          int32_t *node_to_free, int8_t *node_sign, double *boundary_values);
void oracle_assembly_destroy(oracle_assembly *assembly);

/* Symbolic CSR and ordered contributor lists (nonlinear_topology.cpp:12-268). */
typedef struct oracle_symbolic oracle_symbolic;
int oracle_prepare_symbolic(const femm_model_t *model, oracle_symbolic **out);
void oracle_symbolic_sizes(const oracle_symbolic *symbolic, int32_t *free_count, int32_t *nnz,
                           int32_t *matrix_entries, int32_t *rhs_entries);
/* matrix_entries: [entries][4] = (triangle, local_i, local_j, sign);
 * rhs_entries: [entries][3] = (triangle, local_i, sign);
 * circuit_rhs: [circuit][free_count]. */
void oracle_symbolic_copy(const oracle_symbolic *symbolic, int32_t *row_offsets,
                          int32_t *column_indices, int32_t *diagonal_slots,
                          int32_t *matrix_offsets, int32_t *matrix_entries, int32_t *rhs_offsets,
                          int32_t *rhs_entries, double *pm_rhs, double *circuit_rhs,
                          int32_t *free_nodes, int32_t *node_to_free, int8_t *node_sign,
                          double *boundary_values);
void oracle_symbolic_destroy(oracle_symbolic *symbolic);
/* The fixed air-gap element terms (nonlinear_topology.cpp:180-232): age_values[nnz] per slot
 * and age_rhs[free_count]. */
void oracle_symbolic_age(const oracle_symbolic *symbolic, double *age_values, double *age_rhs);

/* Air-gap element postprocess (ComputeAirGapElementPostprocess,
 * postprocess/airgap_postprocess.cpp:41-160) of the nodal potential a[node_count]:
 * out[3] = {Fx, Fy, torque}; radial_b[angle_count] from the first element's Fourier series. */
int oracle_age_postprocess(const femm_model_t *model, const double *a, const double *angles_rad,
                           int angle_count, double out[3], double *radial_b);

/* The 10 x 10 air-gap element cell matrix (BuildNativeFemmAirGapMatrix,
 * native_airgap_matrix.cpp:5-22), row-major into out[100]; 0 for invalid inputs. */
int oracle_age_matrix(double K, double Ki, double ci, double co, double out[100]);

/* Jacobi PCG on one CSR system (pcg_kernels.cu:10-344, sequential sums).
 * info[3] = {status, iterations, residual_l2}. */
void oracle_pcg(int n, const int32_t *row_offsets, const int32_t *column_indices,
                const double *values, const double *diagonal, const double *rhs,
                double relative_tolerance, int max_iterations, double *x, double info[3]);

/* Same PCG with the block-Jacobi preconditioner of block_size consecutive rows
 * (1 = point Jacobi). */
void oracle_pcg_block(int n, const int32_t *row_offsets, const int32_t *column_indices,
                      const double *values, const double *diagonal, const double *rhs,
                      double relative_tolerance, int max_iterations, int block_size, double *x,
                      double info[3]);

/* The block-Jacobi inverse: planes[block_size][n] (planes[j n + i] multiplies r of the j-th
 * row of i's block) and fallback[ceil(n / block_size)] (1 = block replaced by point Jacobi). */
void oracle_block_jacobi(int n, const int32_t *row_offsets, const int32_t *column_indices,
                         const double *values, const double *diagonal, int block_size,
                         double *planes, int8_t *fallback);

/* Preconditioner block size of every later Newton / safeguarded / batch solve (1 = point
 * Jacobi, the default). */
void oracle_set_block_size(int block_size);

/* Per-item result of a nonlinear solve. */
typedef struct oracle_result {
  int32_t status;
  int32_t newton_iterations;
  int32_t linear_iterations;
  int32_t reserved;
  double relative_step;           /* last undamped step */
  double equation_residual;       /* ||K(A)A - b|| / ||b|| at the returned state */
} oracle_result_t;

/* Single-item Newton driver (nonlinear_solver.cu:667-769) with recovery
 * (:515-528 -> :534-659), followed by the equation-residual check
 * (:549-558,575) and FinalizeResult (:16-83).  warm_start may be NULL.
 * Outputs: a[node_count], bx/by[triangle_count], flux[circuit_count]. */
int oracle_newton_solve(const femm_model_t *model, const double *currents,
                        const femm_newton_options_t *options, const double *warm_start,
                        double *a, double *bx, double *by, double *flux,
                        oracle_result_t *result);

/* Safeguarded Newton alone (SolveSafeguarded, nonlinear_solver.cu:534-659: full
 * steps with Armijo backtracking on the equation residual, exact linear tolerance).
 * warm_start (may be NULL) replaces the free nodes of the plain start, as the Solve
 * hook does (:667-698); the residual scale stays the plain start's. */
int oracle_safeguarded_solve(const femm_model_t *model, const double *currents,
                             const femm_newton_options_t *options, const double *warm_start,
                             double *a, oracle_result_t *result);
/* The same with inexact linear solves: tolerance forcing x (relative equation residual)
 * within [linear_relative_tolerance, 1e-3]; forcing = 0 is oracle_safeguarded_solve. */
int oracle_safeguarded_solve_forcing(const femm_model_t *model, const double *currents,
                                     const femm_newton_options_t *options, double forcing,
                                     const double *warm_start, double *a,
                                     oracle_result_t *result);

/* Batch of independent items on `threads` host threads (one item per thread at a
 * time, as the reference's host-assembly pool, nonlinear_solver.cu:371-396). */
int oracle_newton_batch(const femm_model_t *model, int items, const double *currents,
                        const femm_newton_options_t *options, int threads, double *a,
                        double *flux, oracle_result_t *results);

/* Equation residual ||K(A)A - rhs(A)|| / ||same at the boundary start||
 * (nonlinear_solver.cu:549-576). */
double oracle_equation_residual(const femm_model_t *model, const double *a,
                                const double *currents);

/* Element B and circuit flux (nonlinear_solver.cu:16-83). */
int oracle_finalize(const femm_model_t *model, const double *a, double *bx, double *by,
                    double *flux);

/* Weighted-stress force/torque and smoothed air-gap radial B
 * (weighted_stress.cu:24-183, field_postprocess.cpp:34-403).
 * selected/air select by group number when >= 0 in *_group, else by material label.
 * out[3] = {Fx, Fy, torque}; radial_b[angle_count]; mask_out[node_count] may be NULL. */
int oracle_weighted_stress(const femm_model_t *model, const double *a, const double *bx,
                           const double *by, int32_t selected_label, int32_t air_label,
                           int32_t selected_group, int32_t air_group, double radius_m,
                           const double *angles_rad, int angle_count, double out[3],
                           double *radial_b, double *mask_out);

/* Shared-operator tangent solves (NonlinearP1Solver::SolveTangentMultiRhs): at base_a with
 * base_currents, da[d][node_count] and dflux[d][circuit_count] for each of `directions`
 * circuit-current directions; results[d] carries status and linear_iterations. */
int oracle_tangent_multi_rhs(const femm_model_t *model, const double *base_a,
                             const double *base_currents, int directions,
                             const double *direction_currents,
                             const femm_newton_options_t *options, double *da, double *dflux,
                             oracle_result_t *results);

#ifdef __cplusplus
}
#endif

#endif
