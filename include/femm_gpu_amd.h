/*
 * femm_gpu_amd: MI355X-native nonlinear magnetostatic solve path (C ABI).
 *
 * Plain pointers and sizes only.  Every function returns a status code of
 * femm_model.h (the reference's Status integers) and never throws: a failed host
 * or device allocation returns FEMM_GPU_ALLOCATION_FAILED, any other internal
 * exception FEMM_INTERNAL_ERROR.  A plan
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
                                  FEMM_PRECISION_MIXED: FP32 values / SpMV / Jacobi / vectors with
                                  FP64 reliable updates (the FP64 residual b - A x replaces the
                                  FP32 one periodically; acceptance on the FP64 residual). */
  int32_t pcg_warm_start;      /* 1: each Newton correction starts PCG from the current state */
  int32_t verify_equation_residual; /* 1: compute ||K(A)A - b(A)|| / ||b(A0)|| per item */
  int32_t max_refinements;     /* cap on FP64 residual replacements per solve (mixed mode) */
  double inner_relative_tolerance; /* mixed mode: replace the residual once the FP32 one has
                                      dropped by this factor since the last replacement */
  int32_t partitions;          /* workgroups per item: 1 = one workgroup owns an item;
                                  > 1 = the item's rows, nodes and triangles are split over
                                  that many workgroups (small batches, large meshes);
                                  0 = chosen by the library (femm_plan_autotune) */
  int32_t block_size;          /* PCG preconditioner: 0 or 1 = point Jacobi (the reference
                                  default); 2, 4 or 8 = block-Jacobi on blocks of that many
                                  consecutive free rows (PcgPreconditioner::kBlockJacobi,
                                  solvers/options.h:8-10, block_jacobi_preconditioner.cu),
                                  FP32 planes in the mixed form, which takes 4 only */
  int32_t split_newton_iterations; /* learned solve, one workgroup per item: 0 = the warm
                                  attempt runs to its budget in one launch.  k > 0 = every
                                  item runs at most k safeguarded iterations in that launch;
                                  the items still iterating then continue (same state, the
                                  rest of the budget) as partitioned teams that share the
                                  whole GPU, so the batch's slowest items do not leave most
                                  compute units idle.  No counterpart in the reference. */
  double inexact_forcing;      /* safeguarded Newton (learned warm attempt, recovery): 0 = every
                                  linear solve to linear_relative_tolerance (the reference's
                                  SolveSafeguarded, nonlinear_solver.cu:534-659); c > 0 =
                                  inexact Newton (Eisenstat & Walker forcing terms): linear
                                  tolerance c x the relative equation residual within
                                  [linear_relative_tolerance, 1e-3], and a step that meets
                                  both gates is solved again to the exact tolerance before the
                                  item is accepted (at the exact candidate when its own
                                  equation residual passes). */
  double acceptance_linear_tolerance; /* safeguarded Newton: 0 = the step that meets both
                                  gates is accepted as the reference's SolveSafeguarded does
                                  (at the current state).  t > 0 (at most the options'
                                  linear_relative_tolerance) = the linearization that met the
                                  gates is solved again to t, starting from its last solution
                                  when pcg_warm_start is set, and the item is returned at that
                                  candidate when its own equation residual passes; the state
                                  is then fixed to about t times the operator's conditioning
                                  rather than to the gates' resolution.  No counterpart in the
                                  reference. */
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
  int32_t fp64_iterations;     /* of linear_iterations, those run in FP64: all of them in the
                                  FP64 forms; the FP64 finishes and acceptance re-solves in
                                  the mixed form */
  int32_t trial_assemblies;    /* Newton assemblies at trial states (the safeguarded policy's
                                  Armijo / acceptance candidates) */
} femm_item_result_t;

/* Plan: validation, signed DOF map, symbolic CSR, ordered contributor lists and
 * device upload.  Replaces NonlinearP1Solver::Initialize + PrepareDeviceAssembly
 * (gpu_solver/src/solvers/nonlinear_solver.cuh:35-38) and
 * DeviceNonlinearAssemblyPlan::Initialize (physics/nonlinear_assembly.cuh:14-26). */
int femm_plan_create(const femm_model_t *model, femm_plan **out);
void femm_plan_destroy(femm_plan *plan);

/* sizes[6] = {node_count, triangle_count, free_count, nnz, circuit_count, quadrature_points}. */
int femm_plan_sizes(const femm_plan *plan, int32_t sizes[6]);

/* The fixed air-gap element terms of the plan (PreparedNonlinearAssembly::age_values and
 * age_rhs, physics/nonlinear_topology.h:41-51; built at nonlinear_topology.cpp:180-232):
 * age_values[nnz] per CSR slot, added after the triangle terms of every assembly, and
 * age_rhs[free], the load of fixed nodes; zeros for a model without air-gap elements. */
int femm_plan_copy_air_gap_terms(const femm_plan *plan, double *age_values, double *age_rhs);

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

/* Same with the block-Jacobi preconditioner of block_size (2, 4 or 8) consecutive rows,
 * built on the device from each item's values (DeterministicPcgKernel<..., kBlock = true>,
 * solvers/pcg_kernels.cu:10-344, with BlockJacobiPlan::Prepare,
 * block_jacobi_preconditioner.cu:188-258); block_size 1 = femm_pcg_batch. */
int femm_pcg_batch_block(femm_plan *plan, int32_t items, const double *values, const double *rhs,
                         double relative_tolerance, int32_t max_iterations, int32_t block_size,
                         double *x, double *info);

/* Batched nonlinear Newton solve with host arrays.  Replaces
 * NonlinearP1Solver::SolveBatch (solvers/nonlinear_solver.cu:183-530) and
 * FinalizeResult (:16-83).  currents[items][C]; warm_starts[items][N] or NULL;
 * config NULL = FP64 reference algorithm.  Outputs: a[items][N],
 * bx/by[items][T] (may be NULL), flux[items][C], results[items].  items = 0 returns
 * FEMM_OK without touching any pointer or the device (an empty batch, :189-192). */
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

/* femm_newton_batch_device without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  Same arguments, kernels and results as
 * femm_newton_batch_device, but the solve workspace lives in a fixed array of
 * FEMM_NOALLOC_SCRATCH_BYTES (128 MiB) that the library holds on every GPU it runs on until
 * the process ends, instead of the plan's workspace.  Limits:
 *  - the plan must already be resident on the device (any earlier computing call on it
 *    uploads it; a fresh plan gives FEMM_INVALID_ARGUMENT);
 *  - a batch whose workspace exceeds the array gives FEMM_GPU_ALLOCATION_FAILED and
 *    launches nothing (per item about 96 T + 8 S + 48 n + 8 N + 17 KiB bytes in FP64, S the
 *    SELL-64 value slots, about 1.2 nnz; the mixed form adds 4 S + 20 n, block-Jacobi
 *    64 n (+ 32 n mixed): about 21 MB per item of the 58,682-node motor, so 6 items);
 *  - every call on a GPU shares that GPU's array and first clears the part it uses, on
 *    `stream`: calls that overlap on one GPU (other streams or threads) race. */
#define FEMM_NOALLOC_SCRATCH_BYTES (128ull << 20)
int femm_newton_batch_device_noalloc(femm_plan *plan, int32_t items, const double *currents,
                                     const femm_newton_options_t *options,
                                     const femm_solve_config_t *config,
                                     const double *warm_starts, double *a, double *flux,
                                     femm_item_result_t *results, void *stream);

/* Batched safeguarded Newton.  Replaces NonlinearP1Solver::SolveSafeguarded
 * (solvers/nonlinear_solver.cu:534-659): full Newton steps from the plain start (or
 * warm_starts, the hook of Solve, :667-698) with Armijo backtracking on the equation residual
 * and both acceptance gates, no driver and no recovery; config->inexact_forcing > 0 makes its
 * linear solves inexact.  Arguments and outputs as femm_newton_batch(_device). */
int femm_newton_batch_safeguarded(femm_plan *plan, int32_t items, const double *currents,
                                  const femm_newton_options_t *options,
                                  const femm_solve_config_t *config, const double *warm_starts,
                                  double *a, double *bx, double *by, double *flux,
                                  femm_item_result_t *results);
int femm_newton_batch_safeguarded_device(femm_plan *plan, int32_t items, const double *currents,
                                         const femm_newton_options_t *options,
                                         const femm_solve_config_t *config,
                                         const double *warm_starts, double *a, double *flux,
                                         femm_item_result_t *results, void *stream);

/* Measurement entry (no counterpart in the reference): `reps` products y = A x per item with
 * the operator values the last solve or assembly on this plan left in its workspace, by the
 * PCG's own SELL-64 SpMV, enqueued on `stream`.  precision FEMM_PRECISION_FP64: x, y are
 * double [items][free]; FEMM_PRECISION_MIXED: float [items][free] on the FP32 value copy
 * (present after a mixed-precision solve). */
int femm_spmv_batch_device(femm_plan *plan, int32_t items, int32_t precision, int32_t reps,
                           const void *x, void *y, void *stream);

/* Autotuner: times a short probe solve (two Newton iterations of 200 PCG iterations each)
 * of `items` items for partition counts 1, 2, 4, ... that fit the device, and records the
 * fastest in the plan; later solves of `items` items with config->partitions = 0 use it.
 * The reference picks its cooperative shard count by a fixed rule
 * (PcgCooperativeShardCount, solvers/options.cpp:43-53). */
int femm_plan_autotune(femm_plan *plan, int32_t items, const femm_solve_config_t *config,
                       int32_t *partitions);

/* One plan's batch inside a group solve: meshes of uneven size solved in one launch.
 * Pointers are device memory of the sizes femm_newton_batch_device takes for `plan`. */
typedef struct femm_group_entry {
  femm_plan *plan;
  int32_t items;
  int32_t partitions;            /* workgroups per item (>= 1) */
  const double *currents;        /* [items][C] */
  double *a;                     /* [items][N] */
  double *flux;                  /* [items][C] */
  femm_item_result_t *results;   /* [items] */
} femm_group_entry_t;

/* A fixed set of entries solved together (no counterpart in the reference, which batches
 * items of one mesh only: solvers/nonlinear_solver.cu:662-742).  The entries' pointers are
 * kept; their contents are read / written by each femm_newton_group call.  An entry's plan
 * may appear once per group.  The plans must outlive the group. */
typedef struct femm_group femm_group;

int femm_group_create(int32_t count, const femm_group_entry_t *entries, femm_group **out);
void femm_group_destroy(femm_group *group);

/* Workgroups per item for every entry (partitions[count]); all of the group's workgroups
 * must be co-resident on the device. */
int femm_group_set_partitions(femm_group *group, const int32_t *partitions);

/* Group autotuner.  Times the probe solve of femm_plan_autotune on every entry's plan at
 * partition counts 1, 2, 4, ... and chooses each entry's partitions so that the slowest
 * entry, priced as work[e] x its probe time (work = expected PCG iterations per item; NULL:
 * sqrt(free DOFs), the Jacobi-PCG growth on a 2D mesh), finishes as early as possible with
 * all workgroups co-resident.  Applies the choice and returns it in partitions[count]
 * (may be NULL) and the launch's workgroup count in *blocks (may be NULL). */
int femm_group_autotune(femm_group *group, const femm_solve_config_t *config,
                        const double *work, int32_t *partitions, int32_t *blocks);

/* The reference Newton driver (per item as femm_newton_batch) for every entry of the group
 * in one cooperative launch on `stream`, not synchronised; with options->recover_nonconvergence
 * a second group launch re-solves exhausted items as femm_newton_batch does.  All entries
 * share options and config (config->partitions is ignored: the group's own apply). */
int femm_newton_group(femm_group *group, const femm_newton_options_t *options,
                      const femm_solve_config_t *config, void *stream);
/* The same launch with the safeguarded Newton (NonlinearP1Solver::SolveSafeguarded,
 * nonlinear_solver.cu:534-659) from the plain start for every item of every entry, no driver
 * and no recovery; config->inexact_forcing > 0 makes its linear solves inexact. */
int femm_newton_group_safeguarded(femm_group *group, const femm_newton_options_t *options,
                                  const femm_solve_config_t *config, void *stream);

/* The launches of one femm_newton_group call (safeguarded = 0) or femm_newton_group_safeguarded
 * call (safeguarded != 0) captured once into a HIP graph that the group keeps; replacing an
 * earlier capture.  No counterpart in the reference, which launches per Newton step from the
 * host (solvers/nonlinear_solver.cu:183-530).  Capture synchronises the device, prepares every
 * entry's device state (this may allocate) and records only kernels, so a replay neither
 * allocates nor copies.  The graph keeps the options, config, partitions and the entries'
 * pointers of the capture; each replay re-reads the currents and rewrites a / flux / results
 * as the eager call would, with the same results bit for bit. */
int femm_group_capture(femm_group *group, const femm_newton_options_t *options,
                       const femm_solve_config_t *config, int32_t safeguarded);
/* One replay of the captured graph on `stream`, not synchronised.  FEMM_INVALID_ARGUMENT
 * without a capture, or once an entry's plan has released or reallocated the device arrays or
 * workspace the graph was captured on (capture again then). */
int femm_group_replay(femm_group *group, void *stream);

/* Per-item material variants on the plan's mesh.  Each variant is a full model with the
 * plan's nodes, triangles, Dirichlet nodes and values, node constraints, problem type,
 * depth, circuit count, material count and B-H knot layout; its material values, the
 * triangle-to-material map and the B-H knot values may differ (per-sample coercivity, B-H
 * scale, coil fill factors, a rotor position on a fixed mesh).  The reference solves one
 * model per NonlinearP1Solver (nonlinear_solver.cuh:32-38), so a variant item equals a
 * reference solve of that variant's model.  count = 0 clears the set; a variant that fails
 * validation leaves the set cleared and returns that variant's status. */
int femm_plan_set_variants(femm_plan *plan, int32_t count, const femm_model_t *variants);

/* femm_newton_batch(_device) with item b solved on variant variant_of_item[b] (an index
 * outside the set gives that item INVALID_ARGUMENT). */
int femm_newton_batch_variants(femm_plan *plan, int32_t items, const int32_t *variant_of_item,
                               const double *currents, const femm_newton_options_t *options,
                               const femm_solve_config_t *config, const double *warm_starts,
                               double *a, double *flux, femm_item_result_t *results);
int femm_newton_batch_variants_device(femm_plan *plan, int32_t items,
                                      const int32_t *variant_of_item, const double *currents,
                                      const femm_newton_options_t *options,
                                      const femm_solve_config_t *config,
                                      const double *warm_starts, double *a, double *flux,
                                      femm_item_result_t *results, void *stream);

/* Shared-operator tangent solves: NonlinearP1Solver::SolveTangentMultiRhs
 * (solvers/nonlinear_solver.cu:85-172) with GpuCsrSolver::SolveCurrentOperatorBatch
 * (solvers/csr_solver.cu:296-390).  At a converged state base_a[N] with circuit currents
 * base_currents[C], the Newton operator J = K + C is assembled once and
 * J dA_d = sum_c direction_currents[d][c] s_c is solved for each of `directions` right-hand
 * sides (s_c: circuit c's per-ampere source load), by Jacobi PCG from x0 = 0 with
 * options->linear_relative_tolerance / max_linear_iterations.  Outputs per direction: the
 * field derivative da[d][N] (zero on prescribed nodes), its element B dbx/dby[d][T] (may be
 * NULL) and flux derivative dflux[d][C] (FinalizeResult of the derivative; with a unit
 * direction, a column of the differential inductance per unit depth), and results[d]
 * (status, linear_iterations, linear_residual).  A non-finite direction makes every
 * result INVALID_ARGUMENT, as in the reference. */
int femm_tangent_multi_rhs(femm_plan *plan, const double *base_a, const double *base_currents,
                           int32_t directions, const double *direction_currents,
                           const femm_newton_options_t *options, double *da, double *dbx,
                           double *dby, double *dflux, femm_item_result_t *results);

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

/* Geometry-only postprocess plan, built on the GPU from the plan's device arrays:
 * weighted-stress mask (element rules, Laplace solve, thresholded at 0.5), mask gradients,
 * node incidence and probe location.  Replaces
 * PrepareWeightedStressMask / SolveWeightedStressMask (postprocess/weighted_stress.cu:24-183)
 * and BuildFrozenPostprocessPlanWithMask (postprocess/field_postprocess.cpp:221-339).
 * A plan whose model has air-gap elements gets their postprocess instead, as the reference's
 * adapters choose for sliding-band artifacts (app/sample_adapter.cpp:59-61): force and torque
 * from the Maxwell stress of every element's cells on its mean circle and radial B at the
 * angles from the first element's discrete Fourier series (ComputeAirGapElementPostprocess,
 * postprocess/airgap_postprocess.cpp:41-160); labels, groups, radius and mask settings are
 * ignored, non-finite angles are FEMM_INVALID_ARGUMENT, and a first element whose cells and
 * harmonics exceed 60 KiB of LDS (about 3,800 cells) is FEMM_UNSUPPORTED_FEATURE. */
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

/* ---------------------------------------------------------------------------
 * Learned warm start.  The reference has a warm-start hook but no learned part
 * (NonlinearP1Solver::Solve(..., warm_start), nonlinear_solver.cu:667-698: the start
 * replaces the initial A, Dirichlet values are re-imposed and Newton terms are used
 * from iteration 0).  The functions below have no counterpart in the reference; they
 * produce starts for that hook.
 * ------------------------------------------------------------------------- */

/* MLP inputs -> hidden -> hidden -> outputs (SiLU), products in BF16 MFMA with FP32
 * accumulation, FP32 master weights trained by Adam.  Parameters are one FP32 array:
 * W1 [in_p][hidden], b1 [hidden], W2 [hidden][hidden], b2 [hidden], W3 [hidden][out_p],
 * b3 [out_p], where in_p / out_p round inputs / outputs up to a multiple of 16 (padded
 * entries are zero and never updated); y = x W + b per layer. */
typedef struct femm_mlp femm_mlp;

int femm_mlp_create(int32_t inputs, int32_t hidden, int32_t outputs, uint64_t seed,
                    femm_mlp **out);
void femm_mlp_destroy(femm_mlp *mlp);
/* layout[10] = {inputs, hidden, outputs, in_p, out_p, parameter count, offset of b1,
 * offset of W2, offset of W3, offset of b3} (W1 at 0, b2 right after W2). */
int femm_mlp_layout(const femm_mlp *mlp, int64_t layout[10]);
int femm_mlp_get_parameters(const femm_mlp *mlp, float *params);
/* Sets FP32 master weights (BF16 copies re-rounded) and clears the Adam state. */
int femm_mlp_set_parameters(femm_mlp *mlp, const float *params);
/* Inputs are standardised as (x - mean) / stddev in FP64 before rounding to BF16. */
int femm_mlp_set_standardization(femm_mlp *mlp, const double *mean, const double *stddev);
/* y[rows][outputs] = MLP(x[rows][inputs]). */
int femm_mlp_forward(femm_mlp *mlp, int32_t rows, const double *x, float *y);
/* One forward + backward pass: loss = mean over rows x outputs of (y - target)^2 and its
 * gradient with respect to every parameter (same layout as the parameters). */
int femm_mlp_gradients(femm_mlp *mlp, int32_t rows, const double *x, const float *target,
                       float *loss, float *gradients);
/* Minibatch Adam (beta 0.9 / 0.999, eps 1e-8), learning rate cosine-decayed to 1 %,
 * epoch-wise shuffles from `seed`; loss_history[steps] (optional) gets each step's loss. */
int femm_mlp_train(femm_mlp *mlp, int32_t samples, const double *x, const float *target,
                   int32_t steps, int32_t batch, float learning_rate, uint64_t seed,
                   float *loss_history);

/* Warm-start model for one plan: POD basis of `modes` modes from solver states, an FP64
 * least-squares linear path and the MLP (inputs -> hidden -> hidden -> modes) fitted to
 * the remaining coefficients.  A0 = mean + sum_k c_k phi_k. */
typedef struct femm_warm_start femm_warm_start;

typedef struct femm_warm_start_train {
  int32_t steps;            /* Adam steps */
  int32_t batch;            /* minibatch rows */
  float learning_rate;      /* initial rate, cosine-decayed to 1 % */
  int32_t pod_iterations;   /* cap on subspace iterations for the POD eigenproblem */
  int32_t linear_skip;      /* 1: FP64 linear path, the MLP fits what it leaves */
  int32_t reserved;
  uint64_t seed;
} femm_warm_start_train_t;

typedef struct femm_warm_start_report {
  int32_t samples;
  int32_t pod_iterations;
  double pod_energy_fraction;       /* retained share of the centred snapshot energy */
  double pod_leading_sigma;
  double pod_last_sigma;
  double linear_residual_fraction;  /* ||c - linear(c)|| / ||c|| over the training set */
  double coefficient_scale;         /* MLP output unit */
  float first_loss;
  float final_loss;
} femm_warm_start_report_t;

/* The plan must outlive the warm start. */
int femm_warm_start_create(femm_plan *plan, int32_t inputs, int32_t modes, int32_t hidden,
                           uint64_t seed, femm_warm_start **out);
void femm_warm_start_destroy(femm_warm_start *ws);
/* features[samples][inputs] and states[samples][N] (converged solutions). */
int femm_warm_start_fit(femm_warm_start *ws, int32_t samples, const double *features,
                        const double *states, const femm_warm_start_train_t *train);
int femm_warm_start_fit_device(femm_warm_start *ws, int32_t samples, const double *features,
                               const double *states, const femm_warm_start_train_t *train,
                               void *stream);
int femm_warm_start_report(const femm_warm_start *ws, femm_warm_start_report_t *report);
/* a0[items][N] predicted starts. */
int femm_warm_start_predict(femm_warm_start *ws, int32_t items, const double *features,
                            double *a0);
/* Batched solve from the learned start: the reference's safeguarded Newton (full steps,
 * Armijo backtracking on the equation residual, SolveSafeguarded,
 * nonlinear_solver.cu:534-659) with at most warm_newton_iterations iterations and both
 * acceptance gates; every item not accepted there is solved again in the same call from
 * the plain start by the same safeguarded policy with options->max_newton_iterations (the
 * reference's own recovery path), so every item is accepted under one policy.  Counters of
 * the attempts add up per item. */
int femm_newton_batch_learned(femm_plan *plan, femm_warm_start *ws, int32_t items,
                              const int32_t *variant_of_item, const double *features,
                              const double *currents, const femm_newton_options_t *options,
                              const femm_solve_config_t *config, int32_t warm_newton_iterations,
                              double *a, double *flux, femm_item_result_t *results);
int femm_newton_batch_learned_device(femm_plan *plan, femm_warm_start *ws, int32_t items,
                                     const int32_t *variant_of_item, const double *features,
                                     const double *currents, const femm_newton_options_t *options,
                                     const femm_solve_config_t *config,
                                     int32_t warm_newton_iterations, double *a, double *flux,
                                     femm_item_result_t *results, void *stream);

#ifdef __cplusplus
}
#endif

#endif
