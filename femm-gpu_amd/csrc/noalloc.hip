// Batched Newton solve whose workspace lives in a fixed device array instead of device memory
// the library allocates: for callers that may not allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).  The kernels and launches are those of
// femm_newton_batch_device; only where the workspace lives differs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "api_guard.hpp"
#include "device.hpp"
#include "femm_gpu_amd.h"
#include "handle.hpp"

namespace
{

constexpr size_t kScratchBytes = size_t(128) << 20;   // FEMM_NOALLOC_SCRATCH_BYTES
constexpr int kMaxDevices = 64;
static_assert(kScratchBytes == FEMM_NOALLOC_SCRATCH_BYTES, "header and library disagree");

// The solve workspace of every no-allocation call on one GPU; each GPU holds its own copy
// for the life of the process.  In host code this name is only a host-side stand-in: the
// device address is the pointer hipGetSymbolAddress returns on that device.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with that device
// current and kept; every later use goes through this pointer.
int ScratchAddress(void **out)
{
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess || device < 0)
    return FEMM_GPU_UNAVAILABLE;
  if (device >= kMaxDevices)
    return FEMM_INTERNAL_ERROR;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device] == nullptr) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr)
      return FEMM_INTERNAL_ERROR;
    g_scratch[device] = ptr;
  }
  *out = g_scratch[device];
  return FEMM_OK;
}

// Zeroes the part of the scratch a call uses (SELL padding values and the vectors' rows past
// the last free row must read as zero, as in a freshly zeroed plan workspace).  A kernel
// rather than hipMemsetAsync, so the call may sit in a captured graph.
__global__ __launch_bounds__(256) void ClearScratch(uint4 *words, size_t count)
{
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < count;
       k += static_cast<size_t>(gridDim.x) * blockDim.x)
    words[k] = zero;
}

}  // namespace

extern "C" {

int femm_newton_batch_device_noalloc(femm_plan *plan, int32_t items, const double *currents,
                                     const femm_newton_options_t *options,
                                     const femm_solve_config_t *config,
                                     const double *warm_starts, double *a, double *flux,
                                     femm_item_result_t *results, void *stream)
{
  FEMM_API_BEGIN
  if (plan != nullptr && items == 0)
    return FEMM_OK;
  if (plan == nullptr || items <= 0 || currents == nullptr || a == nullptr || flux == nullptr ||
      results == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm_newton_options_t o = options ? *options : femm::DefaultNewtonOptions();
  femm_solve_config_t c = config ? *config : femm::DefaultSolveConfig();
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  // Uploading the plan allocates: it must already be resident (any earlier computing call
  // on the plan uploads it).
  if (plan->dev.base == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const bool mixed = c.precision == FEMM_PRECISION_MIXED, blocks = c.block_size > 1;
  const size_t bytes = femm::WorkBytes(plan->host, items, mixed, blocks);
  if (bytes > kScratchBytes)
    return FEMM_GPU_ALLOCATION_FAILED;   // the batch's workspace does not fit the array
  void *base = nullptr;
  int st = ScratchAddress(&base);
  if (st != FEMM_OK)
    return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t words = bytes / sizeof(uint4);   // WorkBytes rounds every array to 256 bytes
  const unsigned grid = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 4096));
  ClearScratch<<<grid, 256, 0, s>>>(static_cast<uint4 *>(base), words);
  if (hipGetLastError() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  // The plan's device arrays with the workspace carved from the scratch.
  femm::DeviceBuffers view;
  view.plan = plan->dev.plan;
  view.variants = plan->dev.variants;
  view.work = femm::CarveWork(plan->host, items, mixed, blocks, base);
  view.work_base = base;
  view.work_bytes = bytes;
  if (c.partitions == 0)
    c.partitions = femm::AutoPartitions(plan->dev, items);
  femm::LaunchExtras x;
  st = femm::LaunchNewton(view, items, currents, o, c, warm_starts, a, flux, results, s, x);
  if (st != FEMM_OK || !o.recover_nonconvergence || o.max_newton_iterations <= 0)
    return st;
  // Bounded recovery of exhausted items, as femm_newton_batch_device.
  x.prior = results;
  x.prior_mode = femm::kPriorRedoUnconverged;
  return femm::LaunchSafeguarded(view, items, currents, o, c, nullptr, a, flux, results, s, x);
  FEMM_API_END
}

}  // extern "C"
