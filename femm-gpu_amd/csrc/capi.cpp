// C ABI of the MI355X solve path (declared in include/femm_gpu_amd.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <new>
#include <memory>
#include <string>
#include <vector>

#include "api_guard.hpp"
#include "device.hpp"
#include "femm_gpu_amd.h"
#include "handle.hpp"
#include "plan.hpp"
#include "post.hpp"

namespace
{

constexpr const char *kStatusNames[] = {
    "OK", "INPUT_IO", "INVALID_ARGUMENT", "UNSUPPORTED_FEATURE", "MESH_INVALID",
    "BOUNDARY_INVALID", "GPU_UNAVAILABLE", "GPU_ALLOCATION_FAILED", "ASSEMBLY_FAILED",
    "LINEAR_SOLVE_NOT_CONVERGED", "LINEAR_SOLVE_BREAKDOWN", "NUMERICAL_NONFINITE", "OUTPUT_IO",
    "INTERNAL_ERROR", "NONLINEAR_SOLVE_NOT_CONVERGED", "INVALID_MATERIAL"};

bool GpuPresent()
{
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

int EnsureIo(femm_plan *plan, size_t bytes)
{
  if (plan->io_bytes >= bytes)
    return FEMM_OK;
  if (plan->io) {
    (void)hipDeviceSynchronize();
    (void)hipFree(plan->io);
    plan->io = nullptr;
    plan->io_bytes = 0;
  }
  if (hipMalloc(&plan->io, bytes) != hipSuccess)
    return FEMM_GPU_ALLOCATION_FAILED;
  plan->io_bytes = bytes;
  return FEMM_OK;
}

size_t Round(size_t bytes) { return (bytes + 255) / 256 * 256; }

femm_newton_options_t DefaultOptions()
{
  femm_newton_options_t o;
  o.relative_tolerance = 1e-12;
  o.linear_relative_tolerance = 1e-13;
  o.max_newton_iterations = 32;
  o.max_linear_iterations = 128;
  o.recover_nonconvergence = 0;
  o.adaptive_linear_tolerance = 0;
  o.require_equation_residual = 0;
  o.reserved = 0;
  return o;
}

femm_solve_config_t DefaultConfig()
{
  femm_solve_config_t c;
  c.precision = FEMM_PRECISION_FP64;
  c.pcg_warm_start = 0;
  c.verify_equation_residual = 1;
  c.max_refinements = 64;
  c.inner_relative_tolerance = 0.1;
  c.partitions = 1;
  c.block_size = 1;
  c.split_newton_iterations = 0;
  c.inexact_forcing = 0.0;
  c.acceptance_linear_tolerance = 0.0;
  return c;
}

}  // namespace

namespace femm
{

int EnsurePlanDevice(femm_plan *plan)
{
  if (plan->dev.base != nullptr)
    return FEMM_OK;
  if (!GpuPresent())
    return FEMM_GPU_UNAVAILABLE;
  const int st = femm::UploadPlan(plan->host, &plan->dev);
  if (st != FEMM_OK)
    femm::ReleaseDevice(&plan->dev);
  return st;
}

// Probe solve of the autotuners: at most two Newton iterations of at most 200 PCG iterations
// (a tolerance no solve reaches, so the first linear solve ends at its cap and, as the
// reference driver does with a linear solve that fails, ends the item) of `items` items at
// unit circuit currents with `partitions` workgroups per item: one assembly and 200 PCG
// iterations; *ms is the second of two timed launches.
int ProbeSolve(femm_plan *plan, int32_t items, const femm_solve_config_t &config,
               int32_t partitions, float *ms)
{
  femm_solve_config_t c = config;
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = femm::EnsureWork(plan->host, items, c.precision == FEMM_PRECISION_MIXED, &plan->dev,
                          c.block_size > 1);
  if (st != FEMM_OK)
    return st;
  const femm::HostPlan &h = plan->host;
  femm_newton_options_t o = DefaultOptions();
  o.relative_tolerance = 1e-300;
  o.linear_relative_tolerance = 1e-300;
  o.max_newton_iterations = 2;
  o.max_linear_iterations = 200;
  o.adaptive_linear_tolerance = 0;
  c.verify_equation_residual = 0;
  c.partitions = partitions;
  const size_t B = static_cast<size_t>(items);
  const size_t sa = Round(sizeof(double) * B * h.N), sc = Round(sizeof(double) * B * std::max(h.C, 1));
  const size_t sr = Round(sizeof(femm_item_result_t) * B);
  if ((st = EnsureIo(plan, sa + 2 * sc + sr)) != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_a = reinterpret_cast<double *>(io);
  double *d_cur = reinterpret_cast<double *>(io + sa);
  double *d_flux = reinterpret_cast<double *>(io + sa + sc);
  auto *d_res = reinterpret_cast<femm_item_result_t *>(io + sa + 2 * sc);
  std::vector<double> ones(B * std::max(h.C, 1), 1.0);
  if (hipMemcpy(d_cur, ones.data(), sizeof(double) * B * h.C, hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return FEMM_INTERNAL_ERROR;
  }
  for (int rep = 0; rep < 2 && st == FEMM_OK; ++rep) {
    (void)hipEventRecord(e0, nullptr);
    st = femm::LaunchNewton(plan->dev, items, d_cur, o, c, nullptr, d_a, d_flux, d_res, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (st == FEMM_OK && (hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(ms, e0, e1) != hipSuccess))
      st = FEMM_INTERNAL_ERROR;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

bool ConfigValid(const femm_solve_config_t &c)
{
  // The mixed form's FP32 block-Jacobi keeps a 4-row block in one thread's registers.
  const bool block_ok = c.block_size == 0 || c.block_size == 1 || c.block_size == 4 ||
                        (c.precision == FEMM_PRECISION_FP64 &&
                         (c.block_size == 2 || c.block_size == 8));
  return (c.precision == FEMM_PRECISION_FP64 || c.precision == FEMM_PRECISION_MIXED) && block_ok &&
         c.split_newton_iterations >= 0 && c.inexact_forcing >= 0.0 && c.inexact_forcing < 1.0 &&
         c.acceptance_linear_tolerance >= 0.0 && std::isfinite(c.acceptance_linear_tolerance);
}

femm_newton_options_t DefaultNewtonOptions() { return DefaultOptions(); }
femm_solve_config_t DefaultSolveConfig() { return DefaultConfig(); }
bool DevicePresent() { return GpuPresent(); }
}  // namespace femm

using femm::EnsurePlanDevice;

namespace
{

// Batched Newton on device pointers, optionally per-item material variants: the reference
// driver followed by the bounded recovery when options ask for it, or (safeguarded) the
// safeguarded Newton alone.
int SolveDevice(femm_plan *plan, int32_t items, const int32_t *variant_of_item,
                const double *currents, const femm_newton_options_t *options,
                const femm_solve_config_t *config, const double *warm_starts, double *a,
                double *flux, femm_item_result_t *results, void *stream, bool safeguarded = false)
{
  // An empty batch is an empty result, as SolveBatch returns for empty currents
  // (solvers/nonlinear_solver.cu:189-192); nothing is launched.
  if (plan != nullptr && items == 0)
    return FEMM_OK;
  if (plan == nullptr || items <= 0 || currents == nullptr || a == nullptr || flux == nullptr ||
      results == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm_newton_options_t o = options ? *options : DefaultOptions();
  const femm_solve_config_t c = config ? *config : DefaultConfig();
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = femm::EnsureWork(plan->host, items, c.precision == FEMM_PRECISION_MIXED, &plan->dev,
                          c.block_size > 1);
  if (st != FEMM_OK)
    return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  femm::LaunchExtras x;
  x.variant_of_item = variant_of_item;
  if (safeguarded)
    return femm::LaunchSafeguarded(plan->dev, items, currents, o, c, warm_starts, a, flux,
                                   results, s, x);
  st = femm::LaunchNewton(plan->dev, items, currents, o, c, warm_starts, a, flux, results, s, x);
  if (st != FEMM_OK || !o.recover_nonconvergence || o.max_newton_iterations <= 0)
    return st;
  // Bounded recovery of exhausted items from the plain start
  // (SolveSafeguarded, nonlinear_solver.cu:534-659).
  x.prior = results;
  x.prior_mode = femm::kPriorRedoUnconverged;
  return femm::LaunchSafeguarded(plan->dev, items, currents, o, c, nullptr, a, flux, results, s,
                                 x);
}

}  // namespace

extern "C" {

int femm_plan_create(const femm_model_t *model, femm_plan **out)
{
  FEMM_API_BEGIN
  if (out == nullptr)
    return FEMM_INVALID_ARGUMENT;
  *out = nullptr;
  // Owned until it is handed out, so a failed host allocation below releases it.
  std::unique_ptr<femm_plan> plan(new femm_plan);
  std::string error;
  int st = femm::BuildHostPlan(model, &plan->host, &error);
  if (st != FEMM_OK)
    return st;
  ModelCopy &mc = plan->model;
  const size_t N = static_cast<size_t>(model->node_count), T = static_cast<size_t>(model->triangle_count);
  mc.node_xy.assign(model->node_xy, model->node_xy + 2 * N);
  mc.tri_nodes.assign(model->triangle_nodes, model->triangle_nodes + 3 * T);
  mc.tri_mat.assign(model->triangle_material, model->triangle_material + T);
  mc.materials.assign(model->materials, model->materials + model->material_count);
  mc.dirichlet_nodes.assign(model->dirichlet_nodes, model->dirichlet_nodes + model->dirichlet_count);
  mc.dirichlet_values.assign(model->dirichlet_values,
                             model->dirichlet_values + model->dirichlet_count);
  mc.constraints.assign(model->constraints, model->constraints + model->constraint_count);
  mc.curve_count = model->curve_count;
  if (model->curve_count > 0)
    mc.all_curve_offsets.assign(model->curve_offsets, model->curve_offsets + model->curve_count + 1);
  mc.view = *model;
  mc.view.node_xy = mc.node_xy.data();
  mc.view.triangle_nodes = mc.tri_nodes.data();
  mc.view.triangle_material = mc.tri_mat.data();
  mc.view.materials = mc.materials.data();
  mc.view.dirichlet_nodes = mc.dirichlet_nodes.data();
  mc.view.dirichlet_values = mc.dirichlet_values.data();
  mc.view.constraints = mc.constraints.data();
  mc.view.curve_count = 0;
  mc.view.curve_offsets = mc.curve_offsets.data();
  mc.view.curve_b = nullptr;
  mc.view.curve_h = nullptr;
  mc.view.air_gap_count = static_cast<int32_t>(plan->host.ages.size());
  mc.view.air_gaps = plan->host.ages.empty() ? nullptr : plan->host.ages.data();
  // The device copy is uploaded by the first call that computes, so the
  // symbolic plan is available on hosts without a GPU.
  *out = plan.release();
  return FEMM_OK;
  FEMM_API_END
}

void femm_plan_destroy(femm_plan *plan)
{
  FEMM_API_BEGIN
  if (plan == nullptr)
    return;
  if (plan->dev.base != nullptr || plan->io != nullptr)
    (void)hipDeviceSynchronize();
  femm::ReleaseDevice(&plan->dev);
  if (plan->io)
    (void)hipFree(plan->io);
  if (plan->post_base)
    (void)hipFree(plan->post_base);
  delete plan;
  FEMM_API_END_VOID
}

int femm_plan_sizes(const femm_plan *plan, int32_t sizes[6])
{
  FEMM_API_BEGIN
  if (plan == nullptr || sizes == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  sizes[0] = h.N;
  sizes[1] = h.T;
  sizes[2] = h.n;
  sizes[3] = h.nnz;
  sizes[4] = h.C;
  sizes[5] = h.Q;
  return FEMM_OK;
  FEMM_API_END
}

int femm_plan_copy_air_gap_terms(const femm_plan *plan, double *age_values, double *age_rhs)
{
  FEMM_API_BEGIN
  if (plan == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  if (age_values) {
    if (h.age_vals.empty())
      std::fill(age_values, age_values + h.nnz, 0.0);
    else
      std::copy(h.age_vals.begin(), h.age_vals.end(), age_values);
  }
  if (age_rhs) {
    if (h.age_rhs.empty())
      std::fill(age_rhs, age_rhs + h.n, 0.0);
    else
      std::copy(h.age_rhs.begin(), h.age_rhs.end(), age_rhs);
  }
  return FEMM_OK;
  FEMM_API_END
}

int femm_plan_copy_symbolic(const femm_plan *plan, int32_t *row_offsets, int32_t *column_indices,
                            int32_t *diagonal_slots, int32_t *free_nodes, int32_t *node_to_free,
                            int8_t *node_sign, double *boundary_values)
{
  FEMM_API_BEGIN
  if (plan == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  if (row_offsets)
    std::copy(h.row_offsets.begin(), h.row_offsets.end(), row_offsets);
  if (column_indices)
    std::copy(h.cols.begin(), h.cols.end(), column_indices);
  if (diagonal_slots)
    std::copy(h.diag_slot.begin(), h.diag_slot.end(), diagonal_slots);
  if (free_nodes)
    std::copy(h.free_nodes.begin(), h.free_nodes.end(), free_nodes);
  if (node_to_free)
    std::copy(h.node_to_free.begin(), h.node_to_free.end(), node_to_free);
  if (node_sign)
    std::copy(h.node_sign.begin(), h.node_sign.end(), node_sign);
  if (boundary_values)
    std::copy(h.boundary.begin(), h.boundary.end(), boundary_values);
  return FEMM_OK;
  FEMM_API_END
}

int femm_assemble_batch(femm_plan *plan, int32_t items, const double *states,
                        const double *currents, int32_t use_newton, double *values, double *rhs,
                        double *diagonal)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || states == nullptr || currents == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = femm::EnsureWork(h, items, false, &plan->dev);
  if (st != FEMM_OK)
    return st;
  const size_t ns = Round(sizeof(double) * items * static_cast<size_t>(h.N));
  const size_t nc = Round(sizeof(double) * items * static_cast<size_t>(std::max(h.C, 1)));
  const size_t nv = Round(sizeof(int32_t) * items);
  if ((st = EnsureIo(plan, ns + nc + nv)) != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_states = reinterpret_cast<double *>(io);
  double *d_cur = reinterpret_cast<double *>(io + ns);
  int32_t *d_valid = reinterpret_cast<int32_t *>(io + ns + nc);
  if (hipMemcpy(d_states, states, sizeof(double) * items * static_cast<size_t>(h.N),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_cur, currents, sizeof(double) * items * static_cast<size_t>(h.C),
                hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if ((st = femm::LaunchAssemble(plan->dev, items, d_states, d_cur, use_newton, d_valid,
                                 nullptr)) != FEMM_OK)
    return st;
  std::vector<double> sell(static_cast<size_t>(items) * h.sell_size);
  const size_t nvec = static_cast<size_t>(femm::VectorStride(h.n));   // workspace vector stride
  std::vector<double> r(static_cast<size_t>(items) * nvec);
  std::vector<int32_t> valid(items);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(sell.data(), plan->dev.work.vals, sizeof(double) * sell.size(),
                hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(r.data(), plan->dev.work.rhs, sizeof(double) * r.size(), hipMemcpyDeviceToHost) !=
          hipSuccess ||
      hipMemcpy(valid.data(), d_valid, sizeof(int32_t) * items, hipMemcpyDeviceToHost) !=
          hipSuccess)
    return FEMM_INTERNAL_ERROR;
  for (int32_t item = 0; item < items; ++item) {
    const double *src = sell.data() + static_cast<size_t>(item) * h.sell_size;
    if (values)
      for (int32_t slot = 0; slot < h.nnz; ++slot)
        values[static_cast<size_t>(item) * h.nnz + slot] = src[h.slot_to_sell[slot]];
    if (diagonal)
      for (int32_t row = 0; row < h.n; ++row)
        diagonal[static_cast<size_t>(item) * h.n + row] = src[h.diag_sell[row]];
    if (rhs)
      std::copy(r.begin() + static_cast<size_t>(item) * nvec,
                r.begin() + static_cast<size_t>(item) * nvec + h.n,
                rhs + static_cast<size_t>(item) * h.n);
  }
  for (int32_t v : valid)
    if (!v)
      return FEMM_ASSEMBLY_FAILED;
  return FEMM_OK;
  FEMM_API_END
}

int femm_pcg_batch(femm_plan *plan, int32_t items, const double *values, const double *rhs,
                   double relative_tolerance, int32_t max_iterations, double *x, double *info)
{
  FEMM_API_BEGIN
  return femm_pcg_batch_block(plan, items, values, rhs, relative_tolerance, max_iterations, 1, x,
                              info);
  FEMM_API_END
}

int femm_pcg_batch_block(femm_plan *plan, int32_t items, const double *values, const double *rhs,
                         double relative_tolerance, int32_t max_iterations, int32_t block_size,
                         double *x, double *info)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || values == nullptr || rhs == nullptr ||
      (block_size != 1 && block_size != 2 && block_size != 4 && block_size != 8))
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = femm::EnsureWork(h, items, false, &plan->dev, block_size > 1);
  if (st != FEMM_OK)
    return st;
  std::vector<double> sell(static_cast<size_t>(items) * h.sell_size, 0.0);
  for (int32_t item = 0; item < items; ++item)
    for (int32_t slot = 0; slot < h.nnz; ++slot)
      sell[static_cast<size_t>(item) * h.sell_size + h.slot_to_sell[slot]] =
          values[static_cast<size_t>(item) * h.nnz + slot];
  const size_t ni = Round(sizeof(double) * 3 * items);
  if ((st = EnsureIo(plan, ni)) != FEMM_OK)
    return st;
  double *d_info = static_cast<double *>(plan->io);
  const size_t pitch = sizeof(double) * femm::VectorStride(h.n);   // workspace vector stride
  if (hipMemcpy(plan->dev.work.vals, sell.data(), sizeof(double) * sell.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy2D(plan->dev.work.rhs, pitch, rhs, sizeof(double) * h.n, sizeof(double) * h.n,
                  items, hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if ((st = femm::LaunchPcg(plan->dev, items, relative_tolerance, max_iterations, d_info,
                            nullptr, block_size)) != FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      (x && hipMemcpy2D(x, sizeof(double) * h.n, plan->dev.work.x, pitch, sizeof(double) * h.n,
                        items, hipMemcpyDeviceToHost) != hipSuccess) ||
      (info && hipMemcpy(info, d_info, sizeof(double) * 3 * items, hipMemcpyDeviceToHost) !=
                   hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

int femm_spmv_batch_device(femm_plan *plan, int32_t items, int32_t precision, int32_t reps,
                           const void *x, void *y, void *stream)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || reps < 0 || x == nullptr || y == nullptr ||
      (precision != FEMM_PRECISION_FP64 && precision != FEMM_PRECISION_MIXED))
    return FEMM_INVALID_ARGUMENT;
  if (plan->dev.base == nullptr || plan->dev.work_base == nullptr)
    return FEMM_INVALID_ARGUMENT;   // no operator assembled on the device yet
  return femm::LaunchSpmv(plan->dev, items, precision == FEMM_PRECISION_MIXED, reps, x, y,
                          static_cast<hipStream_t>(stream));
  FEMM_API_END
}

int femm_newton_batch_device(femm_plan *plan, int32_t items, const double *currents,
                             const femm_newton_options_t *options,
                             const femm_solve_config_t *config, const double *warm_starts,
                             double *a, double *flux, femm_item_result_t *results, void *stream)
{
  FEMM_API_BEGIN
  return SolveDevice(plan, items, nullptr, currents, options, config, warm_starts, a, flux,
                     results, stream);
  FEMM_API_END
}

int femm_newton_batch_safeguarded_device(femm_plan *plan, int32_t items, const double *currents,
                                         const femm_newton_options_t *options,
                                         const femm_solve_config_t *config,
                                         const double *warm_starts, double *a, double *flux,
                                         femm_item_result_t *results, void *stream)
{
  FEMM_API_BEGIN
  return SolveDevice(plan, items, nullptr, currents, options, config, warm_starts, a, flux,
                     results, stream, true);
  FEMM_API_END
}

int femm_newton_batch_variants_device(femm_plan *plan, int32_t items,
                                      const int32_t *variant_of_item, const double *currents,
                                      const femm_newton_options_t *options,
                                      const femm_solve_config_t *config,
                                      const double *warm_starts, double *a, double *flux,
                                      femm_item_result_t *results, void *stream)
{
  FEMM_API_BEGIN
  if (plan != nullptr && items == 0)
    return FEMM_OK;
  if (plan == nullptr || variant_of_item == nullptr || plan->dev.variants.count == 0)
    return FEMM_INVALID_ARGUMENT;
  return SolveDevice(plan, items, variant_of_item, currents, options, config, warm_starts, a,
                     flux, results, stream);
  FEMM_API_END
}

int femm_plan_autotune(femm_plan *plan, int32_t items, const femm_solve_config_t *config,
                       int32_t *partitions)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || partitions == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm_solve_config_t c = config ? *config : DefaultConfig();
  int device = 0, cus = 0;
  if (!GpuPresent() || hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return FEMM_GPU_UNAVAILABLE;
  int best = 1, st = FEMM_OK;
  float best_ms = -1.0f;
  for (int p = 1; p <= femm::kMaxPartitions; p *= 2) {
    if (static_cast<long>(items) * p > cus || (p > 1 && p > plan->host.nslices))
      break;
    if (p == 1 && items * 4 < cus)
      continue;   // a lone workgroup per item cannot fill the chip
    float ms = 0.0f;
    if ((st = femm::ProbeSolve(plan, items, c, p, &ms)) != FEMM_OK)
      return st;
    if (best_ms < 0.0f || ms < best_ms) {
      best_ms = ms;
      best = p;
    }
  }
  auto &tuned = plan->dev.tuned;
  tuned.erase(std::remove_if(tuned.begin(), tuned.end(),
                             [items](const std::pair<int32_t, int32_t> &t) { return t.first == items; }),
              tuned.end());
  tuned.emplace_back(items, best);
  *partitions = best;
  return FEMM_OK;
  FEMM_API_END
}

int femm_plan_set_variants(femm_plan *plan, int32_t count, const femm_model_t *variants)
{
  FEMM_API_BEGIN
  if (plan == nullptr || count < 0 || (count > 0 && variants == nullptr))
    return FEMM_INVALID_ARGUMENT;
  int st = EnsurePlanDevice(plan);
  if (st != FEMM_OK)
    return st;
  if ((st = femm::AllocateVariants(plan->host, 0, &plan->dev)) != FEMM_OK || count == 0)
    return st;
  femm_model_t base = plan->model.view;
  base.curve_count = plan->model.curve_count;
  base.curve_offsets = plan->model.all_curve_offsets.empty() ? nullptr
                                                            : plan->model.all_curve_offsets.data();
  if ((st = femm::AllocateVariants(plan->host, count, &plan->dev)) != FEMM_OK)
    return st;
  // Material parts are built on host threads in bounded chunks and copied as they finish.
  const int workers = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
  const int chunk = 4 * workers;
  std::vector<femm::HostPlan> parts(chunk);
  std::vector<int> codes(chunk);
  for (int first = 0; first < count; first += chunk) {
    const int m = std::min(chunk, count - first);
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int k = next++; k < m; k = next++) {
        const femm_model_t *v = variants + first + k;
        codes[k] = femm::SameTopology(&base, v)
                       ? femm::BuildHostPlan(v, &parts[k], nullptr, &plan->host)
                       : FEMM_INVALID_ARGUMENT;
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::min(workers, m); ++t)
      pool.emplace_back(work);
    work();
    for (std::thread &t : pool)
      t.join();
    for (int k = 0; k < m; ++k) {
      if (codes[k] == FEMM_OK)
        codes[k] = femm::CopyVariant(plan->host, first + k, parts[k], &plan->dev);
      if (codes[k] != FEMM_OK) {
        (void)femm::AllocateVariants(plan->host, 0, &plan->dev);
        return codes[k];
      }
    }
  }
  return hipDeviceSynchronize() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
  FEMM_API_END
}

}  // extern "C"

namespace
{

// Host-pointer batched solve: stage inputs, run on the device, copy results back.
int SolveHost(femm_plan *plan, int32_t items, const int32_t *variant_of_item,
              const double *currents, const femm_newton_options_t *options,
              const femm_solve_config_t *config, const double *warm_starts, double *a,
              double *bx, double *by, double *flux, femm_item_result_t *results,
              bool safeguarded = false)
{
  if (plan != nullptr && items == 0)   // empty batch: empty result, nothing touched
    return FEMM_OK;
  if (plan == nullptr || items <= 0 || currents == nullptr || a == nullptr || flux == nullptr ||
      results == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  const size_t B = static_cast<size_t>(items);
  const size_t sa = Round(sizeof(double) * B * h.N), sc = Round(sizeof(double) * B * std::max(h.C, 1));
  const size_t sr = Round(sizeof(femm_item_result_t) * B);
  const size_t sb = (bx && by) ? Round(sizeof(double) * B * h.T) : 0;
  const size_t sw = warm_starts ? sa : 0;
  const size_t sv = variant_of_item ? Round(sizeof(int32_t) * B) : 0;
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = EnsureIo(plan, sa + 2 * sc + sr + 2 * sb + sw + sv);
  if (st != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_a = reinterpret_cast<double *>(io);
  double *d_cur = reinterpret_cast<double *>(io + sa);
  double *d_flux = reinterpret_cast<double *>(io + sa + sc);
  auto *d_res = reinterpret_cast<femm_item_result_t *>(io + sa + 2 * sc);
  double *d_bx = sb ? reinterpret_cast<double *>(io + sa + 2 * sc + sr) : nullptr;
  double *d_by = sb ? reinterpret_cast<double *>(io + sa + 2 * sc + sr + sb) : nullptr;
  double *d_warm = sw ? reinterpret_cast<double *>(io + sa + 2 * sc + sr + 2 * sb) : nullptr;
  int32_t *d_var = sv ? reinterpret_cast<int32_t *>(io + sa + 2 * sc + sr + 2 * sb + sw) : nullptr;
  if (hipMemcpy(d_cur, currents, sizeof(double) * B * h.C, hipMemcpyHostToDevice) != hipSuccess ||
      (d_warm && hipMemcpy(d_warm, warm_starts, sizeof(double) * B * h.N,
                           hipMemcpyHostToDevice) != hipSuccess) ||
      (d_var && hipMemcpy(d_var, variant_of_item, sizeof(int32_t) * B, hipMemcpyHostToDevice) !=
                    hipSuccess))
    return FEMM_INTERNAL_ERROR;
  st = SolveDevice(plan, items, d_var, d_cur, options, config, d_warm, d_a, d_flux, d_res, nullptr,
                   safeguarded);
  if (st != FEMM_OK)
    return st;
  if (sb && (st = femm::LaunchFinalize(plan->dev, items, d_a, d_bx, d_by, nullptr, nullptr)) !=
                FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(a, d_a, sizeof(double) * B * h.N, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(flux, d_flux, sizeof(double) * B * h.C, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(results, d_res, sizeof(femm_item_result_t) * B, hipMemcpyDeviceToHost) !=
          hipSuccess ||
      (sb && (hipMemcpy(bx, d_bx, sizeof(double) * B * h.T, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(by, d_by, sizeof(double) * B * h.T, hipMemcpyDeviceToHost) != hipSuccess)))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
}

}  // namespace

extern "C" {

int femm_newton_batch(femm_plan *plan, int32_t items, const double *currents,
                      const femm_newton_options_t *options, const femm_solve_config_t *config,
                      const double *warm_starts, double *a, double *bx, double *by, double *flux,
                      femm_item_result_t *results)
{
  FEMM_API_BEGIN
  return SolveHost(plan, items, nullptr, currents, options, config, warm_starts, a, bx, by, flux,
                   results);
  FEMM_API_END
}

int femm_newton_batch_safeguarded(femm_plan *plan, int32_t items, const double *currents,
                                  const femm_newton_options_t *options,
                                  const femm_solve_config_t *config, const double *warm_starts,
                                  double *a, double *bx, double *by, double *flux,
                                  femm_item_result_t *results)
{
  FEMM_API_BEGIN
  return SolveHost(plan, items, nullptr, currents, options, config, warm_starts, a, bx, by, flux,
                   results, true);
  FEMM_API_END
}

int femm_newton_batch_variants(femm_plan *plan, int32_t items, const int32_t *variant_of_item,
                               const double *currents, const femm_newton_options_t *options,
                               const femm_solve_config_t *config, const double *warm_starts,
                               double *a, double *flux, femm_item_result_t *results)
{
  FEMM_API_BEGIN
  if (plan != nullptr && items == 0)
    return FEMM_OK;
  if (plan == nullptr || variant_of_item == nullptr || plan->dev.variants.count == 0)
    return FEMM_INVALID_ARGUMENT;
  return SolveHost(plan, items, variant_of_item, currents, options, config, warm_starts, a,
                   nullptr, nullptr, flux, results);
  FEMM_API_END
}

int femm_tangent_multi_rhs(femm_plan *plan, const double *base_a, const double *base_currents,
                           int32_t directions, const double *direction_currents,
                           const femm_newton_options_t *options, double *da, double *dbx,
                           double *dby, double *dflux, femm_item_result_t *results)
{
  FEMM_API_BEGIN
  if (plan == nullptr || base_a == nullptr || base_currents == nullptr || directions <= 0 ||
      direction_currents == nullptr || options == nullptr || da == nullptr || dflux == nullptr ||
      results == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  const size_t D = static_cast<size_t>(directions), C = static_cast<size_t>(h.C);
  femm_item_result_t invalid{};
  invalid.status = FEMM_INVALID_ARGUMENT;
  bool ok = options->linear_relative_tolerance > 0.0 && options->max_linear_iterations >= 0;
  for (size_t i = 0; i < D * C; ++i)
    ok = ok && std::isfinite(direction_currents[i]);
  if (!ok) {
    // SolveTangentMultiRhs marks every direction invalid (nonlinear_solver.cu:90-104).
    for (size_t d = 0; d < D; ++d)
      results[d] = invalid;
    return FEMM_OK;
  }
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = femm::EnsureWork(h, directions, false, &plan->dev);
  if (st != FEMM_OK)
    return st;
  const size_t sa = Round(sizeof(double) * h.N), sc = Round(sizeof(double) * std::max<size_t>(C, 1));
  const size_t sd = Round(sizeof(double) * D * std::max<size_t>(C, 1));
  const size_t sda = Round(sizeof(double) * D * h.N), sb = Round(sizeof(double) * D * h.T);
  const size_t sr = Round(sizeof(femm_item_result_t) * D), sv = Round(sizeof(int32_t));
  const bool fields = dbx != nullptr && dby != nullptr;
  if ((st = EnsureIo(plan, sa + sc + sd + sda + (fields ? 2 * sb : 0) + sd + sr + sv)) != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_a = reinterpret_cast<double *>(io);
  double *d_cur = reinterpret_cast<double *>(io + sa);
  double *d_dir = reinterpret_cast<double *>(io + sa + sc);
  double *d_da = reinterpret_cast<double *>(io + sa + sc + sd);
  char *rest = io + sa + sc + sd + sda;
  double *d_bx = fields ? reinterpret_cast<double *>(rest) : nullptr;
  double *d_by = fields ? reinterpret_cast<double *>(rest + sb) : nullptr;
  rest += fields ? 2 * sb : 0;
  double *d_flux = reinterpret_cast<double *>(rest);
  auto *d_res = reinterpret_cast<femm_item_result_t *>(rest + sd);
  auto *d_valid = reinterpret_cast<int32_t *>(rest + sd + sr);
  if (hipMemcpy(d_a, base_a, sizeof(double) * h.N, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_cur, base_currents, sizeof(double) * C, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_dir, direction_currents, sizeof(double) * D * C, hipMemcpyHostToDevice) !=
          hipSuccess)
    return FEMM_INTERNAL_ERROR;
  // The Newton operator at the base state (AssembleNonlinearNewton with use_newton).
  if ((st = femm::LaunchAssemble(plan->dev, 1, d_a, d_cur, 1, d_valid, nullptr)) != FEMM_OK)
    return st;
  int32_t valid = 0;
  if (hipMemcpy(&valid, d_valid, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if (!valid) {
    femm_item_result_t failed{};
    failed.status = FEMM_ASSEMBLY_FAILED;
    for (size_t d = 0; d < D; ++d)
      results[d] = failed;
    return FEMM_OK;
  }
  if ((st = femm::LaunchTangent(plan->dev, directions, d_dir, options->linear_relative_tolerance,
                                options->max_linear_iterations, d_da, d_res, nullptr)) !=
          FEMM_OK ||
      (st = femm::LaunchFinalize(plan->dev, directions, d_da, d_bx, d_by, d_flux, nullptr)) !=
          FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(da, d_da, sizeof(double) * D * h.N, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(dflux, d_flux, sizeof(double) * D * C, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(results, d_res, sizeof(femm_item_result_t) * D, hipMemcpyDeviceToHost) !=
          hipSuccess ||
      (fields && (hipMemcpy(dbx, d_bx, sizeof(double) * D * h.T, hipMemcpyDeviceToHost) !=
                      hipSuccess ||
                  hipMemcpy(dby, d_by, sizeof(double) * D * h.T, hipMemcpyDeviceToHost) !=
                      hipSuccess)))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

int femm_finalize_batch(femm_plan *plan, int32_t items, const double *states, double *bx,
                        double *by, double *flux)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || states == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  const size_t B = static_cast<size_t>(items);
  const size_t sa = Round(sizeof(double) * B * h.N), st_ = Round(sizeof(double) * B * h.T);
  const size_t sc = Round(sizeof(double) * B * std::max(h.C, 1));
  int st = EnsurePlanDevice(plan);
  if (st == FEMM_OK)
    st = EnsureIo(plan, sa + 2 * st_ + sc);
  if (st != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_a = reinterpret_cast<double *>(io);
  double *d_bx = reinterpret_cast<double *>(io + sa);
  double *d_by = reinterpret_cast<double *>(io + sa + st_);
  double *d_flux = reinterpret_cast<double *>(io + sa + 2 * st_);
  if (hipMemcpy(d_a, states, sizeof(double) * B * h.N, hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if ((st = femm::LaunchFinalize(plan->dev, items, d_a, d_bx, d_by, d_flux, nullptr)) != FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      (bx && hipMemcpy(bx, d_bx, sizeof(double) * B * h.T, hipMemcpyDeviceToHost) != hipSuccess) ||
      (by && hipMemcpy(by, d_by, sizeof(double) * B * h.T, hipMemcpyDeviceToHost) != hipSuccess) ||
      (flux && hipMemcpy(flux, d_flux, sizeof(double) * B * h.C, hipMemcpyDeviceToHost) !=
                   hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

int femm_plan_set_postprocess(femm_plan *plan, const femm_postprocess_options_t *options)
{
  FEMM_API_BEGIN
  if (plan == nullptr || options == nullptr)
    return FEMM_INVALID_ARGUMENT;
  int st = EnsurePlanDevice(plan);
  if (st != FEMM_OK)
    return st;
  if (plan->post_base) {
    (void)hipDeviceSynchronize();
    (void)hipFree(plan->post_base);
    plan->post_base = nullptr;
  }
  plan->post_ready = false;
  // A model with air-gap elements gets the element's own postprocess, as the reference's
  // adapters choose for sliding-band artifacts (app/sample_adapter.cpp:59-61, 434-436).
  if (plan->host.ages.empty()) {
    st = femm::BuildPostPlanDevice(&plan->model.view, plan->host, plan->dev, *options,
                                   &plan->post, &plan->post_dev, &plan->post_base);
  } else {
    st = femm::BuildAgePostPlan(plan->host, plan->model.view.depth_m, *options, &plan->post);
    if (st == FEMM_OK)
      st = femm::UploadPostPlan(plan->post, plan->dev, &plan->post_dev, &plan->post_base);
  }
  if (st != FEMM_OK)
    return st;
  plan->post_ready = true;
  return FEMM_OK;
  FEMM_API_END
}

int femm_postprocess_copy_mask(const femm_plan *plan, double *mask)
{
  FEMM_API_BEGIN
  if (plan == nullptr || mask == nullptr || !plan->post_ready)
    return FEMM_INVALID_ARGUMENT;
  std::copy(plan->post.mask.begin(), plan->post.mask.end(), mask);
  return FEMM_OK;
  FEMM_API_END
}

int femm_postprocess_batch_device(femm_plan *plan, int32_t items, const double *states,
                                  double *loads, double *radial_b, void *stream)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || states == nullptr || loads == nullptr ||
      (radial_b == nullptr && plan->post_dev.nprobe > 0) || !plan->post_ready)
    return FEMM_INVALID_ARGUMENT;
  return femm::LaunchPostprocess(plan->post_dev, items, states, loads, radial_b,
                                 static_cast<hipStream_t>(stream));
  FEMM_API_END
}

int femm_postprocess_batch(femm_plan *plan, int32_t items, const double *states, double *loads,
                           double *radial_b)
{
  FEMM_API_BEGIN
  if (plan == nullptr || items <= 0 || states == nullptr || loads == nullptr || !plan->post_ready)
    return FEMM_INVALID_ARGUMENT;
  const femm::HostPlan &h = plan->host;
  const size_t B = static_cast<size_t>(items), P = static_cast<size_t>(plan->post_dev.nprobe);
  const size_t sa = Round(sizeof(double) * B * h.N), sl = Round(sizeof(double) * B * 3);
  const size_t sp = Round(sizeof(double) * B * std::max<size_t>(P, 1));
  int st = EnsureIo(plan, sa + sl + sp);
  if (st != FEMM_OK)
    return st;
  char *io = static_cast<char *>(plan->io);
  double *d_a = reinterpret_cast<double *>(io);
  double *d_l = reinterpret_cast<double *>(io + sa);
  double *d_p = reinterpret_cast<double *>(io + sa + sl);
  if (hipMemcpy(d_a, states, sizeof(double) * B * h.N, hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if ((st = femm::LaunchPostprocess(plan->post_dev, items, d_a, d_l, d_p, nullptr)) != FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(loads, d_l, sizeof(double) * B * 3, hipMemcpyDeviceToHost) != hipSuccess ||
      (P && radial_b &&
       hipMemcpy(radial_b, d_p, sizeof(double) * B * P, hipMemcpyDeviceToHost) != hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

const char *femm_status_name(int32_t status)
{
  if (status < 0 || status >= static_cast<int32_t>(sizeof(kStatusNames) / sizeof(kStatusNames[0])))
    return "UNKNOWN";
  return kStatusNames[status];
}

int femm_device_info(char *text, size_t capacity)
{
  FEMM_API_BEGIN
  if (text == nullptr || capacity == 0)
    return FEMM_INVALID_ARGUMENT;
  hipDeviceProp_t prop;
  int dev = 0;
  if (!GpuPresent() || hipGetDevice(&dev) != hipSuccess ||
      hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    std::snprintf(text, capacity, "no HIP device");
    return FEMM_GPU_UNAVAILABLE;
  }
  std::snprintf(text, capacity, "%s %s CUs=%d HBM=%.1fGiB L2=%dKiB", prop.name, prop.gcnArchName,
                prop.multiProcessorCount, prop.totalGlobalMem / 1073741824.0,
                prop.l2CacheSize / 1024);
  return FEMM_OK;
  FEMM_API_END
}

}  // extern "C"
