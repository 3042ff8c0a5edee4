// Device-side plan and launch interface of the MI355X solve path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "femm_gpu_amd.h"
#include "plan.hpp"

namespace femm
{

constexpr int kBlockThreads = 512;    // one workgroup (8 wave64) per batch item
constexpr int kElemStride = 12;       // per-triangle 3x3 Newton matrix + 3 rhs terms

// Device pointers of the immutable plan (uploaded once per geometry).
struct DevicePlan {
  int32_t N, T, n, nnz, C, Q, nslices, sell_size;
  int32_t max_slice_width;      // widest SELL slice
  int32_t planar, any_nonlinear, any_boundary_value;
  double depth;
  const int32_t *tri_nodes, *tri_mat, *tri_circuit;
  const double *tri_w, *tri_gx, *tri_gy, *tri_flux;
  const double *mat_nu0, *mat_alpha, *mat_cold;
  const int32_t *mat_curve_off, *mat_curve_cnt;
  const double *curve_b, *curve_h, *curve_s;
  const int32_t *node_to_free, *free_nodes;
  const int8_t *node_sign;
  const double *boundary;
  const int32_t *slot_off, *contrib, *row_off, *rcontrib;
  const double *pm_rhs, *circuit_rhs;
  const int32_t *slice_ptr, *slice_width, *sell_cols, *slot_to_sell, *diag_sell;
  const int32_t *row_offsets;   // CSR [n + 1]: slot range of a row partition
  // Native air-gap elements: fixed stiffness per CSR slot [nnz], null without them (their
  // fixed-node load is part of pm_rhs).
  const double *age_vals;
};

// Per-item device workspaces, item-major.  Strides in elements.
struct DeviceWork {
  int32_t capacity;
  double *elem;      // [items][T][kElemStride]
  double *vals;      // [items][sell_size]
  float *vals32;     // [items][sell_size]   (mixed precision)
  double *rhs, *inv, *x, *r, *p, *ap;   // [items][n]
  double *p_alt, *q_buf;                // [items][n] second direction and A p of team solves
  double *trial;     // [items][N] Armijo trial state (safeguarded Newton)
  float *r32, *p32, *ap32, *d32, *inv32; // [items][n]   (mixed precision)
  float *p32_alt, *q32_buf;             // [items][n]   (mixed precision, team solves)
  double *bjac;      // [items][kMaxBlock][n] block-Jacobi planes (block_size > 1 only)
  float *bjac32;     // [items][kMaxBlock][n] (mixed precision with block-Jacobi)
  // Partitioned solves (several workgroups per item): per-item reduction partials
  // [items][2][kMaxPartitions][4], barrier words [items][kBarrierWords] and abort flags
  // [items].
  double *team_partials;
  uint32_t *team_barrier;
  int32_t *team_abort;
  // Items a budget-capped launch left to a continuation: [0] = count, then the items in
  // increasing order ([1 + items]).
  int32_t *cont;
};

constexpr int kMaxPartitions = 256;   // workgroups per item in a partitioned solve
// Team barrier words per item: one arrival word per workgroup, then the release word in a
// cache line of its own.
constexpr int kBarrierRelease = kMaxPartitions;
constexpr int kBarrierWords = kMaxPartitions + 32;
constexpr int kMaxBlock = 8;          // largest block-Jacobi block (rows)

// Per-item stride of the workspace's free-row vectors: n rounded up to 64 rows, so vector
// passes may load 16-byte chunks (the rows past n stay zero).
__host__ __device__ constexpr int VectorStride(int n) { return (n + 63) / 64 * 64; }

// Per-item material variants on the plan's mesh: every array is [count][stride] with the
// plan's sizes (T triangles, `mats` materials, `knots` B-H knots, n free rows, C circuits).
struct DeviceVariants {
  int32_t count = 0, mats = 0, knots = 0;
  const int32_t *tri_mat = nullptr, *tri_circuit = nullptr;
  const double *tri_flux = nullptr;
  const double *mat_nu0 = nullptr, *mat_alpha = nullptr, *mat_cold = nullptr;
  const int32_t *mat_curve_off = nullptr, *mat_curve_cnt = nullptr;
  const double *curve_b = nullptr, *curve_h = nullptr, *curve_s = nullptr;
  const double *pm_rhs = nullptr, *circuit_rhs = nullptr;
  const int32_t *any_nonlinear = nullptr;
};

struct DeviceBuffers {
  DevicePlan plan{};
  DeviceVariants variants{};
  void *variants_base = nullptr;
  void *base = nullptr;          // one allocation for the immutable plan
  DeviceWork work{};
  void *work_base = nullptr;
  size_t work_bytes = 0;
  std::vector<std::pair<int32_t, int32_t>> tuned;   // (items, partitions) from the autotuner
};

int UploadPlan(const HostPlan &host, DeviceBuffers *dev);
// Replaces the variant set with `count` zeroed variants laid out for `topology`, then
// fills variant `index` from a material part built with BuildHostPlan(..., &topology).
int AllocateVariants(const HostPlan &topology, int32_t count, DeviceBuffers *dev);
int CopyVariant(const HostPlan &topology, int32_t index, const HostPlan &part, DeviceBuffers *dev);
// Workspace for `items` items; `mixed` adds the FP32 arrays, `blocks` the block-Jacobi planes.
// Bytes of a solve workspace for `capacity` items and its layout from `base` (every
// array 256-byte aligned); EnsureWork allocates, zeroes and carves one.
size_t WorkBytes(const HostPlan &host, int32_t capacity, bool mixed, bool blocks);
DeviceWork CarveWork(const HostPlan &host, int32_t capacity, bool mixed, bool blocks, void *base);
int EnsureWork(const HostPlan &host, int32_t items, bool mixed, DeviceBuffers *dev,
               bool blocks = false);
void ReleaseDevice(DeviceBuffers *dev);

// Options of one solve launch (kernel argument).
struct NewtonArgs {
  femm_newton_options_t options;
  femm_solve_config_t config;
  int32_t prior_mode;
  femm_newton_options_t warm_options;   // learned mode: the warm attempt's options
  int32_t resume;                       // safeguarded list launch: 1 = continue each item
                                        // from its state in A, 0 = from the plain start
};

// Which items a launch with prior results solves again (`prior` may alias `results`;
// the others keep the earlier outputs untouched, and re-solved items add their
// iteration counters to the earlier ones).
enum PriorMode : int32_t {
  kPriorNone = 0,          // no prior results: solve every item
  kPriorRedoFailed = 1,    // items whose prior status is not FEMM_OK
  kPriorRedoUnconverged = 2  // items whose prior status is NONLINEAR_SOLVE_NOT_CONVERGED
};

// One plan's batch inside a group launch (meshes of uneven size in one launch).
struct GroupEntry {
  DevicePlan plan;
  DeviceWork work;
  const double *currents;
  double *a, *flux;
  femm_item_result_t *results;
  int32_t items, partitions, first_block, reserved;
};

// The entries (also copied to `entries_device`, same order) in one cooperative launch of the
// reference driver, or with `safeguarded` of the safeguarded Newton from the plain start:
// entry e's items use blocks [first_block, first_block + items x partitions).  With
// `prior_results` only the items an earlier group launch left unconverged are solved again
// (results read and rewritten in place): the bounded recovery.
int LaunchGroup(const GroupEntry *entries_device, const std::vector<GroupEntry> &entries,
                const femm_newton_options_t &options, const femm_solve_config_t &config,
                bool safeguarded, bool prior_results, hipStream_t stream);
// Co-resident workgroup capacity of a group launch on the current device.
int GroupCapacity(bool mixed);

// Workgroups per item chosen for `items` items when the config asks for it (partitions = 0).
int AutoPartitions(const DeviceBuffers &dev, int32_t items);

// Optional per-launch inputs: earlier results (see PriorMode) and the material variant
// of each item (null = the plan's own materials).
struct LaunchExtras {
  const femm_item_result_t *prior = nullptr;
  int32_t prior_mode = kPriorNone;
  const int32_t *variant_of_item = nullptr;
};

// Block-Jacobi solve forms (device_block.hip): `form` carries the precision bit (1) and the
// block bit (2); `mode` is the solve mode (0 reference driver, 1 safeguarded, 2 learned).
int LaunchBlockSolve(int form, int mode, const DeviceBuffers &dev, int32_t items,
                     int32_t partitions, const double *currents, const NewtonArgs &args,
                     const double *warm, double *a, double *flux, femm_item_result_t *results,
                     hipStream_t stream, const LaunchExtras &x);
// The group kernel of a block-Jacobi form (reference driver or safeguarded recovery).
const void *BlockGroupKernel(int form, bool recovery);
// Tangent solves (tangent.hip): one PCG per direction on the Newton operator assembled into
// item 0 (LaunchAssemble at the base state); da[directions][N], results[directions].
int LaunchTangent(const DeviceBuffers &dev, int32_t directions, const double *direction_currents,
                  double tol, int32_t maxit, double *da, femm_item_result_t *results,
                  hipStream_t stream);
// PCG with block-Jacobi on work.vals / work.rhs (x0 = 0); info[items][3].
int LaunchBlockPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit,
                   int32_t block_size, double *info, hipStream_t stream);

// Batched Newton solve with the reference driver's policy (nonlinear_solver.cu:183-530)
// on device pointers, enqueued on `stream`; `warm` null = the plain start.
int LaunchNewton(const DeviceBuffers &dev, int32_t items, const double *currents,
                 const femm_newton_options_t &options, const femm_solve_config_t &config,
                 const double *warm, double *a, double *flux, femm_item_result_t *results,
                 hipStream_t stream, const LaunchExtras &extras = LaunchExtras{});
// Learned-start solve, two launches on `stream`: per item, the safeguarded Newton from
// `warm` with `warm_options` (its iteration budget), then, for the items that attempt did
// not accept, the reference driver from the plain start with `plain_options`.
int LaunchLearned(const DeviceBuffers &dev, int32_t items, const double *currents,
                  const femm_newton_options_t &warm_options,
                  const femm_newton_options_t &plain_options, const femm_solve_config_t &config,
                  const double *warm, double *a, double *flux, femm_item_result_t *results,
                  hipStream_t stream, const LaunchExtras &extras = LaunchExtras{});
// Safeguarded Newton (full steps with Armijo backtracking on the equation residual,
// nonlinear_solver.cu:534-659) from `warm` or the plain start.
int LaunchSafeguarded(const DeviceBuffers &dev, int32_t items, const double *currents,
                      const femm_newton_options_t &options, const femm_solve_config_t &config,
                      const double *warm, double *a, double *flux, femm_item_result_t *results,
                      hipStream_t stream, const LaunchExtras &extras = LaunchExtras{});
// Assembly only (values in SELL layout into work.vals, rhs, inv) for `items`.
int LaunchAssemble(const DeviceBuffers &dev, int32_t items, const double *states,
                   const double *currents, int32_t newton, int32_t *valid, hipStream_t stream);
// reps products y = A x per item on the workspace operator values (FP64 or the FP32 copy);
// x, y [items][n] of that type.
int LaunchSpmv(const DeviceBuffers &dev, int32_t items, bool fp32, int32_t reps, const void *x,
               void *y, hipStream_t stream);
// PCG only on work.vals / work.rhs (x0 = 0); info[items][3].
int LaunchPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit, double *info,
              hipStream_t stream, int32_t block_size = 1);
// The same PCG for item 0 over several co-resident workgroups (cooperative launch); info[3].
int LaunchTeamPcg(const DeviceBuffers &dev, double tol, int32_t maxit, double *info,
                  hipStream_t stream);
// Element B and circuit flux for states[items][N].
int LaunchFinalize(const DeviceBuffers &dev, int32_t items, const double *states, double *bx,
                   double *by, double *flux, hipStream_t stream);

}  // namespace femm
