// BF16 MFMA kernels of the warm-start MLP (forward, backward, Adam) and the
// femm_mlp_* C entry points used to train and check the network on its own.
//
// GEMM tiling: one wave64 owns one 16x16 output tile and walks K in steps of 32 with
// gfx950's v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16] (i = 0..7) and receives D[4(l/16)+i][l%16] (i = 0..3).  An operand
// stored K-contiguous (A as M x K, or B given as its transpose, N x K) arrives as one
// 16-byte load per lane; the others as eight 2-byte loads.  A remaining 16-wide K step
// (K is padded to 16) uses v_mfma_f32_16x16x16_bf16 with four values per lane.  Four
// waves per workgroup cover a 32x32 block; every output element is accumulated by one
// lane in a fixed K order, so results are bitwise repeatable.

#include "api_guard.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <new>
#include <random>

#include "femm_gpu_amd.h"

namespace femm
{
namespace
{

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kTile = 16;
constexpr int kThreads = 256;

enum Activation { kLinear = 0, kSilu = 1 };

__device__ inline float Sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// Z = op(A) op(B) + bias (Z optional), H = act(Z) as BF16 (optional).
// op(A) is M x K, op(B) is K x N; M, N, K are multiples of 16.
template <bool kTransA, bool kTransB>
__global__ __launch_bounds__(kThreads) void GemmBf16(const bf16 *__restrict__ A,
                                                     const bf16 *__restrict__ B,
                                                     const float *__restrict__ bias,
                                                     float *__restrict__ Z, bf16 *__restrict__ H,
                                                     int M, int N, int K, int act)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tm = (blockIdx.y * 2 + (wave >> 1)) * kTile;
  const int tn = (blockIdx.x * 2 + (wave & 1)) * kTile;
  if (tm >= M || tn >= N)
    return;
  const int row = lane & 15, quad = lane >> 4;
  float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
  int k0 = 0;
  for (; k0 + 2 * kTile <= K; k0 += 2 * kTile) {
    const int kq = k0 + 8 * quad;
    short8_t a, b;
    if constexpr (kTransA) {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        a[i] = static_cast<short>(A[static_cast<size_t>(kq + i) * M + tm + row]);
    } else {
      a = *reinterpret_cast<const short8_t *>(A + static_cast<size_t>(tm + row) * K + kq);
    }
    if constexpr (kTransB) {
      b = *reinterpret_cast<const short8_t *>(B + static_cast<size_t>(tn + row) * K + kq);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        b[i] = static_cast<short>(B[static_cast<size_t>(kq + i) * N + tn + row]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b), acc, 0, 0, 0);
  }
  if (k0 < K) {   // K % 32 == 16
    short4_t a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + 4 * quad + i;
      a[i] = static_cast<short>(kTransA ? A[static_cast<size_t>(k) * M + tm + row]
                                        : A[static_cast<size_t>(tm + row) * K + k]);
      b[i] = static_cast<short>(kTransB ? B[static_cast<size_t>(tn + row) * K + k]
                                        : B[static_cast<size_t>(k) * N + tn + row]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = tm + 4 * quad + i, n = tn + row;
    const float z = acc[i] + (bias ? bias[n] : 0.0f);
    const size_t at = static_cast<size_t>(m) * N + n;
    if (Z)
      Z[at] = z;
    if (H)
      H[at] = ToBf16(act == kSilu ? z * Sigmoid(z) : z);
  }
}

// dZ = dH * silu'(Z); FP32 kept for the bias sums, BF16 for the next products.
__global__ void SiluBackward(const float *__restrict__ dH, const float *__restrict__ Z,
                             float *__restrict__ dZ, bf16 *__restrict__ dZ16, size_t count)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const float s = Sigmoid(Z[i]);
  const float g = dH[i] * (s * (1.0f + Z[i] * (1.0f - s)));
  dZ[i] = g;
  dZ16[i] = ToBf16(g);
}

// Mean-squared error over `rows` x `outputs`: dY = 2 (Y - T) / (rows outputs); rows at
// or beyond `rows` and padded columns get zero gradient.  Per-row squared error sums
// into loss_rows.
__global__ void MseGradient(const float *__restrict__ Y, const float *__restrict__ T,
                            float *__restrict__ dY, bf16 *__restrict__ dY16,
                            float *__restrict__ loss_rows, int rows, int padded_rows, int outputs,
                            int ld)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= padded_rows)
    return;
  const float scale = 2.0f / (static_cast<float>(rows) * static_cast<float>(outputs));
  float loss = 0.0f;
  for (int c = 0; c < ld; ++c) {
    const size_t at = static_cast<size_t>(r) * ld + c;
    const float d =
        (r < rows && c < outputs) ? Y[at] - T[static_cast<size_t>(r) * outputs + c] : 0.0f;
    loss += d * d;
    dY[at] = scale * d;
    dY16[at] = ToBf16(scale * d);
  }
  if (r < rows)
    loss_rows[r] = loss;
}

// db[c] = sum_r dZ[r][c]: a workgroup takes 64 columns; its kSumGroups row groups sum the
// rows r = g, g + kSumGroups, ... in order (coalesced 64-column rows), and the group sums
// are added in group order, so the result is repeatable.
constexpr int kSumGroups = kThreads / 64;

__global__ __launch_bounds__(kThreads) void ColumnSum(const float *__restrict__ dZ,
                                                      float *__restrict__ db, int rows, int cols)
{
  __shared__ float part[kSumGroups][64];
  const int g = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int c = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (c < cols)
    for (int r = g; r < rows; r += kSumGroups)
      s += dZ[static_cast<size_t>(r) * cols + c];
  part[g][lane] = s;
  __syncthreads();
  if (g == 0 && c < cols) {
    float t = part[0][lane];
    for (int h = 1; h < kSumGroups; ++h)
      t += part[h][lane];
    db[c] = t;
  }
}

unsigned ColumnBlocks(int cols) { return static_cast<unsigned>((cols + 63) / 64); }

__global__ void AdamKernel(float *__restrict__ w, bf16 *__restrict__ w16,
                           const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, const uint8_t *__restrict__ mask, size_t count,
                           float lr, float b1, float b2, float eps, float c1, float c2)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count || !mask[i])
    return;
  const float mi = b1 * m[i] + (1.0f - b1) * g[i];
  const float vi = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  w16[i] = ToBf16(wi);
}

// x16[r][c] = bf16((x[src][c] - mean[c]) * inv_std[c]) for c < in, 0 for padding.
__global__ void PackKernel(const double *__restrict__ x, const double *__restrict__ mean,
                           const double *__restrict__ inv_std, const int32_t *__restrict__ index,
                           bf16 *__restrict__ out, int rows, int padded_rows, int in, int in_p)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= padded_rows)
    return;
  const int src = r < rows ? (index ? index[r] : r) : -1;
  for (int c = 0; c < in_p; ++c) {
    float v = 0.0f;
    if (src >= 0 && c < in)
      v = static_cast<float>((x[static_cast<size_t>(src) * in + c] - mean[c]) * inv_std[c]);
    out[static_cast<size_t>(r) * in_p + c] = ToBf16(v);
  }
}

__global__ void GatherTargets(const float *__restrict__ src, const int32_t *__restrict__ index,
                              float *__restrict__ dst, int rows, int cols)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(rows) * cols)
    return;
  const size_t r = i / cols, c = i % cols;
  dst[i] = src[static_cast<size_t>(index[r]) * cols + c];
}

unsigned Blocks(size_t count) { return static_cast<unsigned>((count + kThreads - 1) / kThreads); }

template <bool TA, bool TB>
void Gemm(const bf16 *A, const bf16 *B, const float *bias, float *Z, bf16 *H, int M, int N, int K,
          int act, hipStream_t s)
{
  hipLaunchKernelGGL((GemmBf16<TA, TB>), dim3((N + 31) / 32, (M + 31) / 32), dim3(kThreads), 0, s,
                     A, B, bias, Z, H, M, N, K, act);
}

}  // namespace

// out[0] = scale * sum of v[0..count), fixed-shape tree in one 256-thread block.
__global__ __launch_bounds__(256) void SumToScalar(const float *__restrict__ v, int count,
                                                   float scale, float *__restrict__ out)
{
  __shared__ float part[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < count; i += 256)
    s += v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[0] = part[0] * scale;
}

MlpLayout MlpLayout::Make(int inputs, int hidden, int outputs)
{
  MlpLayout L;
  L.in = inputs;
  L.hid = hidden;
  L.out = outputs;
  L.in_p = Pad16(inputs);
  L.out_p = Pad16(outputs);
  L.w1 = 0;
  L.b1 = L.w1 + static_cast<size_t>(L.in_p) * hidden;
  L.w2 = L.b1 + hidden;
  L.b2 = L.w2 + static_cast<size_t>(hidden) * hidden;
  L.w3 = L.b2 + hidden;
  L.b3 = L.w3 + static_cast<size_t>(hidden) * L.out_p;
  L.count = L.b3 + L.out_p;
  return L;
}

bool MlpActivations::Ensure(const MlpLayout &L, int rows)
{
  const int R = Pad16(std::max(rows, 1));
  if (R <= capacity)
    return true;
  const size_t r = static_cast<size_t>(R);
  const bool ok = x16.Allocate(r * L.in_p) && h1.Allocate(r * L.hid) && h2.Allocate(r * L.hid) &&
                  dy16.Allocate(r * L.out_p) && dz2_16.Allocate(r * L.hid) &&
                  dz1_16.Allocate(r * L.hid) && z1.Allocate(r * L.hid) && z2.Allocate(r * L.hid) &&
                  y.Allocate(r * L.out_p) && dy.Allocate(r * L.out_p) && dh2.Allocate(r * L.hid) &&
                  dz2.Allocate(r * L.hid) && dh1.Allocate(r * L.hid) && dz1.Allocate(r * L.hid) &&
                  target.Allocate(r * L.out) && loss_rows.Allocate(r);
  capacity = ok ? R : 0;
  return ok;
}

int Mlp::Init(int inputs, int hidden, int outputs, uint64_t seed)
{
  if (inputs <= 0 || hidden <= 0 || hidden % 16 || outputs <= 0 || inputs > 4096 ||
      hidden > 4096 || outputs > 4096)
    return FEMM_INVALID_ARGUMENT;
  L = MlpLayout::Make(inputs, hidden, outputs);
  std::vector<float> host(L.count, 0.0f);
  std::vector<uint8_t> live(L.count, 0);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, int rows, int cols, int real_rows, int real_cols) {
    const float bound = std::sqrt(6.0f / static_cast<float>(real_rows));
    std::uniform_real_distribution<float> d(-bound, bound);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) {
        const size_t at = off + static_cast<size_t>(r) * cols + c;
        if (r < real_rows && c < real_cols) {
          host[at] = d(rng);
          live[at] = 1;
        }
      }
  };
  fill(L.w1, L.in_p, L.hid, L.in, L.hid);
  fill(L.w2, L.hid, L.hid, L.hid, L.hid);
  fill(L.w3, L.hid, L.out_p, L.hid, L.out);
  for (int c = 0; c < L.hid; ++c)
    live[L.b1 + c] = live[L.b2 + c] = 1;
  for (int c = 0; c < L.out; ++c)
    live[L.b3 + c] = 1;
  std::vector<double> zero(L.in, 0.0), one(L.in, 1.0);
  if (!w.Allocate(L.count) || !g.Allocate(L.count) || !m.Allocate(L.count) ||
      !v.Allocate(L.count) || !w16.Allocate(L.count) || !mask.Allocate(L.count) ||
      !in_mean.Allocate(L.in) || !in_inv_std.Allocate(L.in))
    return FEMM_GPU_ALLOCATION_FAILED;
  if (hipMemcpy(mask.get(), live.data(), L.count, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(in_mean.get(), zero.data(), sizeof(double) * L.in, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(in_inv_std.get(), one.data(), sizeof(double) * L.in, hipMemcpyHostToDevice) !=
          hipSuccess)
    return FEMM_INTERNAL_ERROR;
  return SetParameters(host.data());
}

int Mlp::SetParameters(const float *host)
{
  std::vector<bf16> low(L.count);
  for (size_t i = 0; i < L.count; ++i)
    low[i] = ToBf16(host[i]);
  if (hipMemcpy(w.get(), host, sizeof(float) * L.count, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(w16.get(), low.data(), sizeof(bf16) * L.count, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(m.get(), 0, sizeof(float) * L.count) != hipSuccess ||
      hipMemset(v.get(), 0, sizeof(float) * L.count) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  step = 0;
  return FEMM_OK;
}

int Mlp::GetParameters(float *host) const
{
  return hipMemcpy(host, w.get(), sizeof(float) * L.count, hipMemcpyDeviceToHost) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
}

int Mlp::SetStandardisation(const double *mean, const double *stddev)
{
  std::vector<double> inv(L.in);
  for (int i = 0; i < L.in; ++i) {
    if (!std::isfinite(mean[i]) || !(stddev[i] > 0.0) || !std::isfinite(stddev[i]))
      return FEMM_INVALID_ARGUMENT;
    inv[i] = 1.0 / stddev[i];
  }
  if (hipMemcpy(in_mean.get(), mean, sizeof(double) * L.in, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(in_inv_std.get(), inv.data(), sizeof(double) * L.in, hipMemcpyHostToDevice) !=
          hipSuccess)
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
}

void Mlp::Pack(const double *x, const int32_t *rows_index, int rows, hipStream_t s)
{
  const int R = Pad16(rows);
  hipLaunchKernelGGL(PackKernel, dim3(Blocks(R)), dim3(kThreads), 0, s, x, in_mean.get(),
                     in_inv_std.get(), rows_index, act.x16.get(), rows, R, L.in, L.in_p);
}

void Mlp::Forward(int rows, hipStream_t s)
{
  const int R = Pad16(rows);
  Gemm<false, false>(act.x16.get(), w16.get() + L.w1, w.get() + L.b1, act.z1.get(), act.h1.get(),
                     R, L.hid, L.in_p, kSilu, s);
  Gemm<false, false>(act.h1.get(), w16.get() + L.w2, w.get() + L.b2, act.z2.get(), act.h2.get(),
                     R, L.hid, L.hid, kSilu, s);
  Gemm<false, false>(act.h2.get(), w16.get() + L.w3, w.get() + L.b3, act.y.get(), nullptr, R,
                     L.out_p, L.hid, kLinear, s);
}

void Mlp::Backward(int rows, hipStream_t s)
{
  const int R = Pad16(rows);
  const size_t cnt = static_cast<size_t>(R) * L.hid;
  hipLaunchKernelGGL(MseGradient, dim3(Blocks(R)), dim3(kThreads), 0, s, act.y.get(),
                     act.target.get(), act.dy.get(), act.dy16.get(), act.loss_rows.get(), rows, R,
                     L.out, L.out_p);
  // Output layer: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T.
  Gemm<true, false>(act.h2.get(), act.dy16.get(), nullptr, g.get() + L.w3, nullptr, L.hid,
                    L.out_p, R, kLinear, s);
  hipLaunchKernelGGL(ColumnSum, dim3(ColumnBlocks(L.out_p)), dim3(kThreads), 0, s, act.dy.get(),
                     g.get() + L.b3, R, L.out_p);
  Gemm<false, true>(act.dy16.get(), w16.get() + L.w3, nullptr, act.dh2.get(), nullptr, R, L.hid,
                    L.out_p, kLinear, s);
  hipLaunchKernelGGL(SiluBackward, dim3(Blocks(cnt)), dim3(kThreads), 0, s, act.dh2.get(),
                     act.z2.get(), act.dz2.get(), act.dz2_16.get(), cnt);
  // Hidden layer 2.
  Gemm<true, false>(act.h1.get(), act.dz2_16.get(), nullptr, g.get() + L.w2, nullptr, L.hid, L.hid,
                    R, kLinear, s);
  hipLaunchKernelGGL(ColumnSum, dim3(ColumnBlocks(L.hid)), dim3(kThreads), 0, s, act.dz2.get(),
                     g.get() + L.b2, R, L.hid);
  Gemm<false, true>(act.dz2_16.get(), w16.get() + L.w2, nullptr, act.dh1.get(), nullptr, R, L.hid,
                    L.hid, kLinear, s);
  hipLaunchKernelGGL(SiluBackward, dim3(Blocks(cnt)), dim3(kThreads), 0, s, act.dh1.get(),
                     act.z1.get(), act.dz1.get(), act.dz1_16.get(), cnt);
  // Hidden layer 1.
  Gemm<true, false>(act.x16.get(), act.dz1_16.get(), nullptr, g.get() + L.w1, nullptr, L.in_p,
                    L.hid, R, kLinear, s);
  hipLaunchKernelGGL(ColumnSum, dim3(ColumnBlocks(L.hid)), dim3(kThreads), 0, s, act.dz1.get(),
                     g.get() + L.b1, R, L.hid);
}

void Mlp::AdamStep(float lr, hipStream_t s)
{
  ++step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, static_cast<float>(step));
  const float c2 = 1.0f - std::pow(b2, static_cast<float>(step));
  hipLaunchKernelGGL(AdamKernel, dim3(Blocks(L.count)), dim3(kThreads), 0, s, w.get(), w16.get(),
                     g.get(), m.get(), v.get(), mask.get(), L.count, lr, b1, b2, 1e-8f, c1, c2);
}

int TrainMlp(Mlp *net, int samples, const double *x, const float *t, int steps, int batch,
             float lr, uint64_t seed, float *loss_history_dev, hipStream_t s)
{
  batch = std::min(batch, samples);
  if (!net->act.Ensure(net->L, batch))
    return FEMM_GPU_ALLOCATION_FAILED;
  DeviceArray<int32_t> order;
  const size_t total = static_cast<size_t>(steps) * batch;
  if (!order.Allocate(std::max<size_t>(total, 1)))
    return FEMM_GPU_ALLOCATION_FAILED;
  // Epoch-wise shuffles drawn on the host from the seed, uploaded once.
  std::vector<int32_t> perm(samples), host(total);
  for (int i = 0; i < samples; ++i)
    perm[i] = i;
  std::mt19937_64 rng(seed);
  size_t cursor = samples;
  for (size_t k = 0; k < total; ++k) {
    if (cursor == static_cast<size_t>(samples)) {
      std::shuffle(perm.begin(), perm.end(), rng);
      cursor = 0;
    }
    host[k] = perm[cursor++];
  }
  if (total && hipMemcpy(order.get(), host.data(), sizeof(int32_t) * total,
                         hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  const float pi = 3.14159265358979f;
  for (int k = 0; k < steps; ++k) {
    const int32_t *rows = order.get() + static_cast<size_t>(k) * batch;
    net->Pack(x, rows, batch, s);
    hipLaunchKernelGGL(GatherTargets, dim3(Blocks(static_cast<size_t>(batch) * net->L.out)),
                       dim3(kThreads), 0, s, t, rows, net->act.target.get(), batch, net->L.out);
    net->Forward(batch, s);
    net->Backward(batch, s);
    const float frac = steps > 1 ? static_cast<float>(k) / static_cast<float>(steps - 1) : 0.0f;
    net->AdamStep(lr * (0.01f + 0.99f * 0.5f * (1.0f + std::cos(pi * frac))), s);
    if (loss_history_dev)
      hipLaunchKernelGGL(SumToScalar, dim3(1), dim3(256), 0, s, net->act.loss_rows.get(), batch,
                         1.0f / (static_cast<float>(batch) * net->L.out), loss_history_dev + k);
  }
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm

struct femm_mlp {
  femm::Mlp net;
};

namespace
{

bool HaveDevice()
{
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

// Host FP64 features (and FP32 targets) staged on the device for one call.
int Stage(const double *x, size_t nx, const float *t, size_t nt, femm::DeviceArray<double> *dx,
          femm::DeviceArray<float> *dt)
{
  if (!dx->Allocate(nx) || (dt && !dt->Allocate(nt)))
    return FEMM_GPU_ALLOCATION_FAILED;
  if (hipMemcpy(dx->get(), x, sizeof(double) * nx, hipMemcpyHostToDevice) != hipSuccess ||
      (dt && hipMemcpy(dt->get(), t, sizeof(float) * nt, hipMemcpyHostToDevice) != hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
}

}  // namespace

extern "C" {

int femm_mlp_create(int32_t inputs, int32_t hidden, int32_t outputs, uint64_t seed, femm_mlp **out)
{
  FEMM_API_BEGIN
  if (out == nullptr)
    return FEMM_INVALID_ARGUMENT;
  *out = nullptr;
  if (!HaveDevice())
    return FEMM_GPU_UNAVAILABLE;
  auto *p = new (std::nothrow) femm_mlp;
  if (p == nullptr)
    return FEMM_INTERNAL_ERROR;
  const int st = p->net.Init(inputs, hidden, outputs, seed);
  if (st != FEMM_OK) {
    delete p;
    return st;
  }
  *out = p;
  return FEMM_OK;
  FEMM_API_END
}

void femm_mlp_destroy(femm_mlp *p)
{
  FEMM_API_BEGIN
  if (p == nullptr)
    return;
  (void)hipDeviceSynchronize();
  delete p;
  FEMM_API_END_VOID
}

int femm_mlp_layout(const femm_mlp *p, int64_t layout[10])
{
  FEMM_API_BEGIN
  if (p == nullptr || layout == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const femm::MlpLayout &L = p->net.L;
  const int64_t v[10] = {L.in,  L.hid, L.out, L.in_p, L.out_p, static_cast<int64_t>(L.count),
                         static_cast<int64_t>(L.b1), static_cast<int64_t>(L.w2),
                         static_cast<int64_t>(L.w3), static_cast<int64_t>(L.b3)};
  std::copy(v, v + 10, layout);
  return FEMM_OK;
  FEMM_API_END
}

int femm_mlp_get_parameters(const femm_mlp *p, float *params)
{
  FEMM_API_BEGIN
  if (p == nullptr || params == nullptr)
    return FEMM_INVALID_ARGUMENT;
  return p->net.GetParameters(params);
  FEMM_API_END
}

int femm_mlp_set_parameters(femm_mlp *p, const float *params)
{
  FEMM_API_BEGIN
  if (p == nullptr || params == nullptr)
    return FEMM_INVALID_ARGUMENT;
  for (size_t i = 0; i < p->net.L.count; ++i)
    if (!std::isfinite(params[i]))
      return FEMM_INVALID_ARGUMENT;
  return p->net.SetParameters(params);
  FEMM_API_END
}

int femm_mlp_set_standardization(femm_mlp *p, const double *mean, const double *stddev)
{
  FEMM_API_BEGIN
  if (p == nullptr || mean == nullptr || stddev == nullptr)
    return FEMM_INVALID_ARGUMENT;
  return p->net.SetStandardisation(mean, stddev);
  FEMM_API_END
}

int femm_mlp_forward(femm_mlp *p, int32_t rows, const double *x, float *y)
{
  FEMM_API_BEGIN
  if (p == nullptr || rows <= 0 || x == nullptr || y == nullptr)
    return FEMM_INVALID_ARGUMENT;
  femm::Mlp &net = p->net;
  if (!net.act.Ensure(net.L, rows))
    return FEMM_GPU_ALLOCATION_FAILED;
  femm::DeviceArray<double> dx;
  int st = Stage(x, static_cast<size_t>(rows) * net.L.in, nullptr, 0, &dx, nullptr);
  if (st != FEMM_OK)
    return st;
  net.Pack(dx.get(), nullptr, rows, nullptr);
  net.Forward(rows, nullptr);
  std::vector<float> full(static_cast<size_t>(femm::Pad16(rows)) * net.L.out_p);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(full.data(), net.act.y.get(), sizeof(float) * full.size(),
                hipMemcpyDeviceToHost) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < net.L.out; ++c)
      y[static_cast<size_t>(r) * net.L.out + c] = full[static_cast<size_t>(r) * net.L.out_p + c];
  return FEMM_OK;
  FEMM_API_END
}

int femm_mlp_gradients(femm_mlp *p, int32_t rows, const double *x, const float *target,
                       float *loss, float *gradients)
{
  FEMM_API_BEGIN
  if (p == nullptr || rows <= 0 || x == nullptr || target == nullptr)
    return FEMM_INVALID_ARGUMENT;
  femm::Mlp &net = p->net;
  if (!net.act.Ensure(net.L, rows))
    return FEMM_GPU_ALLOCATION_FAILED;
  femm::DeviceArray<double> dx;
  int st = Stage(x, static_cast<size_t>(rows) * net.L.in, nullptr, 0, &dx, nullptr);
  if (st != FEMM_OK)
    return st;
  if (hipMemcpy(net.act.target.get(), target, sizeof(float) * rows * net.L.out,
                hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  femm::DeviceArray<float> total;
  if (!total.Allocate(1))
    return FEMM_GPU_ALLOCATION_FAILED;
  net.Pack(dx.get(), nullptr, rows, nullptr);
  net.Forward(rows, nullptr);
  net.Backward(rows, nullptr);
  hipLaunchKernelGGL(femm::SumToScalar, dim3(1), dim3(256), 0, nullptr, net.act.loss_rows.get(),
                     rows, 1.0f / (static_cast<float>(rows) * net.L.out), total.get());
  if (hipDeviceSynchronize() != hipSuccess ||
      (loss && hipMemcpy(loss, total.get(), sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) ||
      (gradients && hipMemcpy(gradients, net.g.get(), sizeof(float) * net.L.count,
                              hipMemcpyDeviceToHost) != hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

int femm_mlp_train(femm_mlp *p, int32_t samples, const double *x, const float *target,
                   int32_t steps, int32_t batch, float learning_rate, uint64_t seed,
                   float *loss_history)
{
  FEMM_API_BEGIN
  if (p == nullptr || samples <= 0 || x == nullptr || target == nullptr || steps < 0 ||
      batch <= 0 || !(learning_rate > 0.0f) || !std::isfinite(learning_rate))
    return FEMM_INVALID_ARGUMENT;
  femm::Mlp &net = p->net;
  femm::DeviceArray<double> dx;
  femm::DeviceArray<float> dt, hist;
  int st = Stage(x, static_cast<size_t>(samples) * net.L.in, target,
                 static_cast<size_t>(samples) * net.L.out, &dx, &dt);
  if (st != FEMM_OK)
    return st;
  if (!hist.Allocate(std::max(steps, 1)))
    return FEMM_GPU_ALLOCATION_FAILED;
  st = femm::TrainMlp(&net, samples, dx.get(), dt.get(), steps, batch, learning_rate, seed,
                      hist.get(), nullptr);
  if (st != FEMM_OK)
    return st;
  if (hipDeviceSynchronize() != hipSuccess ||
      (loss_history && steps > 0 &&
       hipMemcpy(loss_history, hist.get(), sizeof(float) * steps, hipMemcpyDeviceToHost) !=
           hipSuccess))
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
  FEMM_API_END
}

}  // extern "C"
