/*
 * femm_gpu_amd: MI355X-native nonlinear magnetostatic solve path (C ABI).
 *
 * Plain pointers and sizes only.  Every function returns a status code of
 * femm_model.h (the reference's Status integers) and never throws.  A plan
 * owns all device memory it allocates; host arrays stay owned by the caller.
 * Calls on one plan are single-thread-affine; a plan may be created on any
 * thread.  Host-pointer entry points return after results are on the host.
 */
#ifndef FEMM_GPU_AMD_H
#define FEMM_GPU_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "femm_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct femm_plan femm_plan;

enum { FEMM_PRECISION_FP64 = 0, FEMM_PRECISION_MIXED = 1 };

/* How the batched Newton solve runs.  The FP64 default with pcg_warm_start = 0
 * is the reference algorithm (x0 = 0 in every Newton correction). */
typedef struct femm_solve_config {
  int32_t precision;           /* FEMM_PRECISION_FP64: FP64 Jacobi PCG.
                                  FEMM_PRECISION_MIXED: FP32 values / SpMV / Jacobi inside FP64
                                  iterative refinement (FP64 residual b - A x). */
  int32_t pcg_warm_start;      /* 1: each Newton correction starts PCG from the current state */
  int32_t verify_equation_residual; /* 1: compute ||K(A)A - b(A)|| / ||b(A0)|| per item */
  int32_t max_refinements;     /* FP64 refinement cap in the mixed mode */
  double inner_relative_tolerance; /* FP32 inner solve target relative to its right-hand side */
} femm_solve_config_t;

/* Per-item outcome (reference SolveInfo plus the Newton counters of
 * NonlinearSolveResult, nonlinear_model.h:116-131). */
typedef struct femm_item_result {
  int32_t status;              /* FEMM_OK, or the first failure of this item */
  int32_t newton_iterations;
  int32_t linear_iterations;   /* PCG iterations summed over Newton steps (FP32 + FP64) */
  int32_t refinements;         /* FP64 refinement residuals evaluated (mixed mode) */
  double relative_step;        /* last undamped ||A_cand - A|| / ||A_cand|| */
  double equation_residual;    /* relative equation residual at the returned state */
  double last_linear_tolerance;
  double linear_residual;      /* last verified ||b - A x|| of the final correction */
} femm_item_result_t;

/* Plan: validation, signed DOF map, symbolic CSR, ordered contributor lists and
 * device upload.  Replaces NonlinearP1Solver::Initialize + PrepareDeviceAssembly
 * (gpu_solver/src/solvers/nonlinear_solver.cuh:35-38) and
 * DeviceNonlinearAssemblyPlan::Initialize (physics/nonlinear_assembly.cuh:14-26). */
int femm_plan_create(const femm_model_t *model, femm_plan **out);
void femm_plan_destroy(femm_plan *plan);

/* sizes[6] = {node_count, triangle_count, free_count, nnz, circuit_count, quadrature_points}. */
int femm_plan_sizes(const femm_plan *plan, int32_t sizes[6]);

/* Symbolic CSR and DOF map (PreparedNonlinearAssembly::symbolic,
 * physics/nonlinear_topology.h:41-51): row_offsets[free+1], column_indices[nnz],
 * diagonal_slots[free], free_nodes[free], node_to_free[N], node_sign[N], boundary[N]. */
int femm_plan_copy_symbolic(const femm_plan *plan, int32_t *row_offsets, int32_t *column_indices,
                            int32_t *diagonal_slots, int32_t *free_nodes, int32_t *node_to_free,
                            int8_t *node_sign, double *boundary_values);

/* Batched Newton assembly on the device, copied back in CSR order:
 * values[items][nnz], rhs[items][free], diagonal[items][free].  states[items][N],
 * currents[items][C].  Replaces DeviceNonlinearAssemblyPlan::Assemble and
 * GpuCsrSolver::CopyPreparedDeviceAssembly (solvers/csr_solver.cuh:61-62). */
int femm_assemble_batch(femm_plan *plan, int32_t items, const double *states,
                        const double *currents, int32_t use_newton, double *values, double *rhs,
                        double *diagonal);

/* Batched Jacobi PCG on the plan's pattern with per-item values (x0 = 0), as
 * GpuCsrSolver::SolveBatch / LaunchDeterministicPcg (solvers/pcg_kernels.cuh:9-21).
 * info[items][3] = {status, iterations, residual_l2}. */
int femm_pcg_batch(femm_plan *plan, int32_t items, const double *values, const double *rhs,
                   double relative_tolerance, int32_t max_iterations, double *x, double *info);

/* Batched nonlinear Newton solve with host arrays.  Replaces
 * NonlinearP1Solver::SolveBatch (solvers/nonlinear_solver.cu:183-530) and
 * FinalizeResult (:16-83).  currents[items][C]; warm_starts[items][N] or NULL;
 * config NULL = FP64 reference algorithm.  Outputs: a[items][N],
 * bx/by[items][T] (may be NULL), flux[items][C], results[items]. */
int femm_newton_batch(femm_plan *plan, int32_t items, const double *currents,
                      const femm_newton_options_t *options, const femm_solve_config_t *config,
                      const double *warm_starts, double *a, double *bx, double *by, double *flux,
                      femm_item_result_t *results);

/* Same solve on device-resident buffers, enqueued on `stream` (a hipStream_t;
 * NULL = default stream) and not synchronised: every pointer is device memory
 * of the sizes above; warm_starts may be NULL.  For callers that keep inputs
 * and outputs in HBM (benchmarks, pipelines). */
int femm_newton_batch_device(femm_plan *plan, int32_t items, const double *currents,
                             const femm_newton_options_t *options,
                             const femm_solve_config_t *config, const double *warm_starts,
                             double *a, double *flux, femm_item_result_t *results, void *stream);

/* Element B and circuit flux for given states (FinalizeResult,
 * solvers/nonlinear_solver.cu:16-83) on the device. */
int femm_finalize_batch(femm_plan *plan, int32_t items, const double *states, double *bx,
                        double *by, double *flux);

/* Weighted-stress load and air-gap B selection (reference FrozenPostprocessOptions,
 * gpu_solver/src/postprocess/postprocess.h:20-32). */
typedef struct femm_postprocess_options {
  int32_t selected_label;      /* material label of the loaded body (used when selected_group < 0) */
  int32_t air_label;           /* material label of free space (used when air_group < 0) */
  int32_t selected_group;      /* >= 0: select the loaded body by FEMM group number */
  int32_t air_group;           /* >= 0: select free space by FEMM group number */
  double airgap_radius_m;      /* probe circle radius */
  const double *angles_rad;    /* probe angles */
  int32_t angle_count;
  int32_t max_mask_iterations; /* mask Laplace PCG cap (reference default 256) */
  double mask_relative_tolerance; /* mask PCG tolerance (reference 1e-12) */
} femm_postprocess_options_t;

/* Geometry-only postprocess plan: weighted-stress mask (Laplace solve on the GPU,
 * thresholded at 0.5), mask gradients and probe location.  Replaces
 * PrepareWeightedStressMask / SolveWeightedStressMask (postprocess/weighted_stress.cu:24-183)
 * and BuildFrozenPostprocessPlanWithMask (postprocess/field_postprocess.cpp:221-339). */
int femm_plan_set_postprocess(femm_plan *plan, const femm_postprocess_options_t *options);

/* The 0/1 mask of the current postprocess plan, mask[node_count]. */
int femm_postprocess_copy_mask(const femm_plan *plan, double *mask);

/* Weighted-stress force/torque and FEMM-smoothed radial air-gap B per item
 * (ComputeFrozenPostprocess, postprocess/field_postprocess.cpp:341-403):
 * states[items][N] -> loads[items][3] = {Fx, Fy, torque}, radial_b[items][angle_count]. */
int femm_postprocess_batch(femm_plan *plan, int32_t items, const double *states, double *loads,
                           double *radial_b);

/* Same on device pointers, enqueued on `stream` (hipStream_t) without synchronising. */
int femm_postprocess_batch_device(femm_plan *plan, int32_t items, const double *states,
                                  double *loads, double *radial_b, void *stream);

/* Stable status name (core/types.cpp:6-43). */
const char *femm_status_name(int32_t status);

/* Library / device description: writes a NUL-terminated string. */
int femm_device_info(char *text, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif
