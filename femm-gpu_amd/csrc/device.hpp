// Device-side plan and launch interface of the MI355X solve path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "femm_gpu_amd.h"
#include "plan.hpp"

namespace femm
{

constexpr int kBlockThreads = 512;    // one workgroup (8 wave64) per batch item
constexpr int kElemStride = 12;       // per-triangle 3x3 Newton matrix + 3 rhs terms

// Device pointers of the immutable plan (uploaded once per geometry).
struct DevicePlan {
  int32_t N, T, n, nnz, C, Q, nslices, sell_size;
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
};

// Per-item device workspaces, item-major.  Strides in elements.
struct DeviceWork {
  int32_t capacity;
  double *elem;      // [items][T][kElemStride]
  double *vals;      // [items][sell_size]
  float *vals32;     // [items][sell_size]   (mixed precision)
  double *rhs, *inv, *x, *r, *p, *ap;   // [items][n]
  float *r32, *p32, *ap32, *d32, *inv32; // [items][n]   (mixed precision)
};

struct DeviceBuffers {
  DevicePlan plan{};
  void *base = nullptr;          // one allocation for the immutable plan
  DeviceWork work{};
  void *work_base = nullptr;
  size_t work_bytes = 0;
};

int UploadPlan(const HostPlan &host, DeviceBuffers *dev);
int EnsureWork(const HostPlan &host, int32_t items, bool mixed, DeviceBuffers *dev);
void ReleaseDevice(DeviceBuffers *dev);

// Batched Newton solve on device pointers (enqueued on `stream`).
int LaunchNewton(const DeviceBuffers &dev, int32_t items, const double *currents,
                 const femm_newton_options_t &options, const femm_solve_config_t &config,
                 const double *warm, double *a, double *flux, femm_item_result_t *results,
                 hipStream_t stream);
// Assembly only (values in SELL layout into work.vals, rhs, inv) for `items`.
int LaunchAssemble(const DeviceBuffers &dev, int32_t items, const double *states,
                   const double *currents, int32_t newton, int32_t *valid, hipStream_t stream);
// PCG only on work.vals / work.rhs (x0 = 0); info[items][3].
int LaunchPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit, double *info,
              hipStream_t stream);
// Element B and circuit flux for states[items][N].
int LaunchFinalize(const DeviceBuffers &dev, int32_t items, const double *states, double *bx,
                   double *by, double *flux, hipStream_t stream);

}  // namespace femm
