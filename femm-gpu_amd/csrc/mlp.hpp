// Device-resident MLP used by the learned Newton warm start.
//
// Network: inputs -> hidden (SiLU) -> hidden (SiLU) -> outputs, y = x W + b per layer.
// Products run as BF16 MFMA tiles with FP32 accumulation; FP32 master weights are
// updated by Adam and re-rounded to BF16 after every step.  Every matrix dimension
// is padded to a multiple of 16 (padded weights are held at zero by an update mask).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace femm
{

using bf16 = uint16_t;

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet) and the exact widening.
__host__ __device__ inline bf16 ToBf16(float f)
{
  const uint32_t bits = __builtin_bit_cast(uint32_t, f);
  if ((bits & 0x7f800000u) == 0x7f800000u)
    return static_cast<bf16>((bits >> 16) | ((bits & 0xffffu) ? 0x40u : 0u));
  return static_cast<bf16>((bits + 0x7fffu + ((bits >> 16) & 1u)) >> 16);
}

__host__ __device__ inline float FromBf16(bf16 h)
{
  return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

inline int Pad16(int v) { return (v + 15) / 16 * 16; }

// Owning device allocation (hipMalloc + zero fill).
template <typename T>
class DeviceArray
{
 public:
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { Reset(); }
  bool Allocate(size_t count)
  {
    Reset();
    void *mem = nullptr;
    const size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc(&mem, bytes) != hipSuccess)
      return false;
    if (hipMemset(mem, 0, bytes) != hipSuccess) {
      (void)hipFree(mem);
      return false;
    }
    ptr_ = static_cast<T *>(mem);
    count_ = count;
    return true;
  }
  void Reset()
  {
    if (ptr_)
      (void)hipFree(ptr_);
    ptr_ = nullptr;
    count_ = 0;
  }
  T *get() const { return ptr_; }
  size_t size() const { return count_; }

 private:
  T *ptr_ = nullptr;
  size_t count_ = 0;
};

struct MlpLayout {
  int in = 0, hid = 0, out = 0, in_p = 0, out_p = 0;
  size_t w1 = 0, b1 = 0, w2 = 0, b2 = 0, w3 = 0, b3 = 0, count = 0;
  static MlpLayout Make(int inputs, int hidden, int outputs);
};

// Activations for up to `capacity` rows (a multiple of 16).
struct MlpActivations {
  int capacity = 0;
  DeviceArray<bf16> x16, h1, h2, dy16, dz2_16, dz1_16;
  DeviceArray<float> z1, z2, y, dy, dh2, dz2, dh1, dz1, target, loss_rows;
  bool Ensure(const MlpLayout &L, int rows);
};

struct Mlp {
  MlpLayout L;
  DeviceArray<float> w, g, m, v;
  DeviceArray<bf16> w16;
  DeviceArray<uint8_t> mask;
  DeviceArray<double> in_mean, in_inv_std;   // feature standardisation (FP64 features)
  MlpActivations act;
  int64_t step = 0;

  // Fixed-seed He-uniform initialisation; zero Adam state and standardisation = identity.
  int Init(int inputs, int hidden, int outputs, uint64_t seed);
  int SetParameters(const float *host);
  int GetParameters(float *host) const;
  int SetStandardisation(const double *mean, const double *stddev);
  // Standardise FP64 features x[rows_index[r] or r][in] into act.x16 (rows padded to 16).
  void Pack(const double *x, const int32_t *rows_index, int rows, hipStream_t s);
  // Forward for `rows` packed rows: act.y [pad16(rows)][out_p] FP32.
  void Forward(int rows, hipStream_t s);
  // Gradients of 1/(rows*out) sum (y - target)^2 into g; act.target [rows][out] FP32.
  void Backward(int rows, hipStream_t s);
  void AdamStep(float lr, hipStream_t s);
};

}  // namespace femm
