// Device code of the batched Newton solve shared by the solve translation units:
// reductions, B-H evaluation, assembly, SELL SpMV, block-Jacobi, FP64 and mixed PCG,
// the Newton drivers and the solve kernels, plus the launch templates.  Each solve form
// (kFormMixed | kFormBlock) is instantiated in one translation unit only
// (device.hip: point Jacobi; device_block.hip: block-Jacobi), so the compiler's
// inlining and register allocation of one form do not depend on the others.
//
// Work decomposition: one 512-thread workgroup owns one batch item for the
// whole Newton solve (or `partitions` workgroups own it together, with team
// reductions in workgroup order).  Every reduction is a fixed-shape wave64 shuffle
// tree followed by a fixed-order sum of the 8 wave partials, so results are
// bitwise reproducible from run to run and independent of batch composition.
//
// Memory layout (HBM, item-major): per item the nodal state A[N], the element
// Newton blocks elem[T][12], the operator values in SELL-64 order, and the PCG
// vectors x, r, p, Ap, D^-1 [n].  The SELL column indices, contributor lists and
// element geometry are shared by all items and stay resident in L2 / MALL.

#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "post.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>
#include <vector>

namespace femm
{
namespace
{

constexpr int kWaves = kBlockThreads / 64;

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    if ((expr) != hipSuccess)                                                                      \
      return FEMM_INTERNAL_ERROR;                                                                  \
  } while (0)

// ----------------------------------------------------------------------------- reductions

// The same value in every lane, made wave-uniform (held in scalar registers).
__device__ inline double Uniform(double x)
{
  const uint64_t bits = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(bits));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(bits >> 32));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

struct BlockReducer {
  double *lds;   // [2][kWaves][4]
  int buf;

  template <int K>
  __device__ __attribute__((always_inline)) void Sum(double (&v)[K])
  {
    static_assert(K <= 4, "at most four simultaneous sums");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double x = v[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        x += __shfl_down(x, off, 64);
      v[k] = x;
    }
    double *b = lds + buf * (kWaves * 4);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        b[wave * 4 + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = b[k];
#pragma unroll
      for (int w = 1; w < kWaves; ++w)
        s += b[w * 4 + k];
      v[k] = Uniform(s);
    }
    buf ^= 1;
  }

  __device__ __attribute__((always_inline)) double Max(double v)
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      v = fmax(v, __shfl_down(v, off, 64));
    double *b = lds + buf * (kWaves * 4);
    if (lane == 0)
      b[wave * 4] = v;
    __syncthreads();
    double s = b[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w)
      s = fmax(s, b[w * 4]);
    buf ^= 1;
    return Uniform(s);
  }

  __device__ __attribute__((always_inline)) int AnyTrue(bool local)
  {
    double v = local ? 1.0 : 0.0;
    return Max(v) > 0.0;
  }
};

// Index ranges one workgroup covers for one item: free rows (slice-aligned), SELL
// slices, CSR slots of those rows, nodes and triangles.
struct Span {
  int row0, row1, slice0, slice1, slot0, slot1, node0, node1, tri0, tri1;
};

// The whole item (one workgroup per item).
__device__ inline Span WholeItem(const DevicePlan &P)
{
  return Span{0, P.n, 0, P.nslices, 0, P.nnz, 0, P.N, 0, P.T};
}

// The workgroups that solve one item together, as seen by one of them: its ranges,
// team-wide reductions (identical in every lane and every workgroup) and a team-wide
// barrier.  A single workgroup owns the whole item here.
struct BlockTeam {
  BlockReducer R;
  Span S;
  template <int K>
  __device__ __attribute__((always_inline)) void Sum(double (&v)[K])
  {
    R.Sum(v);
  }
  __device__ __attribute__((always_inline)) double Max(double v) { return R.Max(v); }
  __device__ __attribute__((always_inline)) int AnyTrue(bool local) { return R.AnyTrue(local); }
  __device__ __attribute__((always_inline)) void Sync() { __syncthreads(); }
};


// Several workgroups solving one item together (a partitioned solve).  Workgroup g of G
// owns a slice-aligned share of the free rows (with their CSR slots) and equal shares of
// the nodes and triangles.  A reduction turns each workgroup's block sum into a partial in
// global memory; after a team barrier every workgroup adds the G partials in workgroup
// order, so the result is identical in every workgroup and repeatable for a fixed G.  The
// barrier is flat: workgroup g publishes its writes (agent-scope release) and stores the
// barrier's number in its own arrival word; workgroup 0 watches all G arrival words with
// one wave (one load per 64 workgroups), then acquires and stores the number in the release
// word, which the others watch before they acquire.  No read-modify-write is contended and
// each workgroup fences twice per barrier.  All workgroups of a launch are co-resident
// (cooperative launch); a wait that still exceeds kMaxBarrierPolls (seconds) marks the
// team aborted, and every later barrier of the team returns at its first poll, so no
// workgroup can wait forever.
constexpr long kMaxBarrierPolls = 1L << 27;

// Wave 0's part of team barrier number `target` for workgroup g of G; one out-of-line copy
// keeps the wait loops from being expanded at every barrier of the solver.
__device__ __noinline__ void TeamBarrier(uint32_t *bar, int32_t *abort, int G, int g,
                                        uint32_t target)
{
  static_assert(kMaxPartitions <= 4 * 64, "four arrival words per lane");
  const int lane = static_cast<int>(threadIdx.x);
  // The workgroup's writes (completed to the L2 by the block barrier before), then arrival.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0)
    __hip_atomic_store(&bar[g], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // Each poll issues all of its loads (the watched words and the abort flag) before it looks
  // at any of them: one memory round trip per poll.  An aborted team stops at the first poll.
  bool aborted = false;
  if (g == 0) {
    for (long polls = 0;; ++polls) {
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int h = lane + 64 * j;
        w[j] = h < G ? __hip_atomic_load(&bar[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : target;
      }
      const int32_t stop = __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bool here = true;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        here &= static_cast<int32_t>(w[j] - target) >= 0;
      if (__all(here))
        break;
      if (stop != 0 || polls > kMaxBarrierPolls) {
        aborted = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // every member's writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // passed on with the release word
    if (!aborted && lane == 0)
      __hip_atomic_store(&bar[kBarrierRelease], target, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else {
    for (long polls = 0;; ++polls) {
      const uint32_t r =
          __hip_atomic_load(&bar[kBarrierRelease], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int32_t stop = __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (static_cast<int32_t>(r - target) >= 0)
        break;
      if (stop != 0 || polls > kMaxBarrierPolls) {
        aborted = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the other workgroups' writes
  }
  if (aborted && lane == 0)
    __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct GridTeam {
  BlockReducer R;
  Span S;
  double *partials;   // [2][G][4]
  uint32_t *bar;      // [kBarrierWords]: arrival word per workgroup, release word
  int32_t *abort;
  int G, g, buf;
  uint32_t passed = 0;   // barriers passed so far (the same count in every member)

  // A team of one workgroup (G == 1) reduces and synchronises like BlockTeam.
  __device__ __attribute__((always_inline)) void Sync()
  {
    __syncthreads();
    if (G == 1)
      return;
    ++passed;
    if (threadIdx.x < 64)
      TeamBarrier(bar, abort, G, g, passed);
    __syncthreads();
  }

  template <int K>
  __device__ __attribute__((always_inline)) void Sum(double (&v)[K])
  {
    R.Sum(v);
    if (G == 1)
      return;
    double *mine = partials + (static_cast<size_t>(buf) * G + g) * 4;
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        mine[k] = v[k];
    }
    Sync();
    static_assert(kMaxPartitions <= kBlockThreads, "one partial per thread");
    // Every member sums the G partials in the same fixed order: thread h loads partial h
    // (all loads issued together, one memory round trip instead of G dependent loads) and
    // the block reduction combines them in its fixed order.
    const double *all = partials + static_cast<size_t>(buf) * G * 4;
    const int h = threadIdx.x;
    double x[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
      x[k] = h < G ? all[h * 4 + k] : 0.0;
    R.Sum(x);
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = x[k];
    buf ^= 1;
  }

  __device__ __attribute__((always_inline)) double Max(double v)
  {
    v = R.Max(v);
    if (G == 1)
      return v;
    double *mine = partials + (static_cast<size_t>(buf) * G + g) * 4;
    if (threadIdx.x == 0)
      mine[0] = v;
    Sync();
    const double *all = partials + static_cast<size_t>(buf) * G * 4;   // as in Sum
    const int h = threadIdx.x;
    const double m = R.Max(h < G ? all[h * 4] : all[0]);
    buf ^= 1;
    return m;
  }

  __device__ __attribute__((always_inline)) int AnyTrue(bool local)
  {
    return Max(local ? 1.0 : 0.0) > 0.0;
  }
};

__device__ inline int UniformInt(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroups in a team (a BlockTeam is one workgroup).
__device__ inline int TeamSize(const BlockTeam &) { return 1; }
__device__ inline int TeamSize(const GridTeam &R) { return R.G; }

// Workgroup g's share of an item split over G workgroups (wave-uniform values).
__device__ inline Span Partition(const DevicePlan &P, int G, int g)
{
  Span s;
  s.slice0 = static_cast<int>(static_cast<long>(P.nslices) * g / G);
  s.slice1 = static_cast<int>(static_cast<long>(P.nslices) * (g + 1) / G);
  s.row0 = min(s.slice0 * kSliceRows, P.n);
  s.row1 = min(s.slice1 * kSliceRows, P.n);
  s.slot0 = P.row_offsets[s.row0];
  s.slot1 = P.row_offsets[s.row1];
  s.node0 = static_cast<int>(static_cast<long>(P.N) * g / G);
  s.node1 = static_cast<int>(static_cast<long>(P.N) * (g + 1) / G);
  s.tri0 = static_cast<int>(static_cast<long>(P.T) * g / G);
  s.tri1 = static_cast<int>(static_cast<long>(P.T) * (g + 1) / G);
  return Span{UniformInt(s.row0),  UniformInt(s.row1),  UniformInt(s.slice0), UniformInt(s.slice1),
              UniformInt(s.slot0), UniformInt(s.slot1), UniformInt(s.node0),  UniformInt(s.node1),
              UniformInt(s.tri0),  UniformInt(s.tri1)};
}

// ----------------------------------------------------------------------------- B-H curve

// Hermite B-H evaluation of the reference (physics/nonlinear_model.cpp:115-148,
// physics/nonlinear_assembly.cu:12-54): linear scan for the segment, linear
// extrapolation beyond the last knot, |B| = 0 gives nu = slope[0], nu' = 0.
__device__ bool EvalBh(const DevicePlan &P, int off, int cnt, double field, double *nu,
                       double *dnu)
{
  const double b = fabs(field);
  if (b == 0.0) {
    *nu = P.curve_s[off];
    *dnu = 0.0;
    return isfinite(*nu);
  }
  const double *cb = P.curve_b + off, *ch = P.curve_h + off, *cs = P.curve_s + off;
  const int last = cnt - 1;
  double h = 0.0, dh = 0.0;
  if (b > cb[last]) {
    h = ch[last] + cs[last] * (b - cb[last]);
    dh = cs[last];
  } else {
    bool found = false;
    for (int i = 0; i < last; ++i) {
      if (b >= cb[i] && b <= cb[i + 1]) {
        const double L = cb[i + 1] - cb[i];
        const double z = (b - cb[i]) / L, z2 = z * z;
        h = (1.0 - 3.0 * z2 + 2.0 * z2 * z) * ch[i] + z * (1.0 - 2.0 * z + z2) * L * cs[i] +
            z2 * (3.0 - 2.0 * z) * ch[i + 1] + z2 * (z - 1.0) * L * cs[i + 1];
        dh = 6.0 * z * (z - 1.0) * ch[i] / L + (1.0 - 4.0 * z + 3.0 * z2) * cs[i] +
             6.0 * z * (1.0 - z) * ch[i + 1] / L + z * (3.0 * z - 2.0) * cs[i + 1];
        found = true;
        break;
      }
    }
    if (!found)
      return false;
  }
  *nu = h / b;
  *dnu = 0.5 * (dh / (b * b) - h / (b * b * b));
  return isfinite(*nu) && isfinite(*dnu);
}

// ----------------------------------------------------------------------------- assembly

// Element pass: B at every quadrature point from the nodal state, nu and
// dnu/dB^2, then the 3x3 Newton block K + C and the per-node rhs terms
// C a_old - (K + C) A_boundary (physics/nonlinear_assembly.cu:56-98 and the
// term-by-term formulas of :100-186).  Returns true if every element is valid.
__device__ __attribute__((always_inline)) bool ElementPass(const DevicePlan &P, const Span &S, const double *a, double *elem,
                            bool newton)
{
  bool ok = true;
  for (int t = S.tri0 + threadIdx.x; t < S.tri1; t += kBlockThreads) {
    const int n0 = P.tri_nodes[3 * t], n1 = P.tri_nodes[3 * t + 1], n2 = P.tri_nodes[3 * t + 2];
    const double av[3] = {a[n0], a[n1], a[n2]};
    double bv[3] = {0.0, 0.0, 0.0};
    if (P.any_boundary_value) {
      bv[0] = P.boundary[n0];
      bv[1] = P.boundary[n1];
      bv[2] = P.boundary[n2];
    }
    const int mat = P.tri_mat[t];
    const int coff = P.mat_curve_off[mat], ccnt = P.mat_curve_cnt[mat];
    const double nu0 = P.mat_nu0[mat], alpha = P.mat_alpha[mat];
    double K[9], e[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 9; ++k)
      K[k] = 0.0;
    for (int q = 0; q < P.Q; ++q) {
      const size_t g = (static_cast<size_t>(t) * P.Q + q) * 3;
      const double w = P.tri_w[static_cast<size_t>(t) * P.Q + q];
      const double gx[3] = {P.tri_gx[g], P.tri_gx[g + 1], P.tri_gx[g + 2]};
      const double gy[3] = {P.tri_gy[g], P.tri_gy[g + 1], P.tri_gy[g + 2]};
      const double bx = av[0] * gx[0] + av[1] * gx[1] + av[2] * gx[2];
      const double by = av[0] * gy[0] + av[1] * gy[1] + av[2] * gy[2];
      const double mag = hypot(bx, by);
      double nu = 0.0, dnu = 0.0;
      bool valid = isfinite(mag);
      if (valid && ccnt > 0) {
        if (!newton) {
          nu = P.mat_cold[mat];
          dnu = 0.0;
        } else {
          valid = EvalBh(P, coff, ccnt, mag, &nu, &dnu);
        }
      } else if (valid) {
        nu = nu0 * (1.0 + alpha * mag * mag);
        dnu = newton ? nu0 * alpha : 0.0;
      }
      if (!valid || !(nu > 0.0) || !isfinite(dnu))
        ok = false;
      double bdot[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        bdot[i] = bx * gx[i] + by * gy[i];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double k = w * nu * (gx[i] * gx[j] + gy[i] * gy[j]);
          const double c = newton ? 2.0 * w * dnu * bdot[i] * bdot[j] : 0.0;
          K[3 * i + j] += k + c;
          if (newton)
            e[i] += c * av[j];
          if (bv[j] != 0.0)
            e[i] -= (k + c) * bv[j];
        }
      }
    }
    double *out = elem + static_cast<size_t>(t) * kElemStride;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      out[k] = K[k];
    out[9] = e[0];
    out[10] = e[1];
    out[11] = e[2];
  }
  return ok;
}

// Slot-owned ordered gather into SELL positions (physics/nonlinear_assembly.cu:100-140):
// contributor order is the host element/i/j emission order; one lane owns a slot.
template <typename Real32>
__device__ __attribute__((always_inline)) void GatherValues(const DevicePlan &P, const Span &S, const double *elem, double *vals,
                             Real32 *vals32)
{
  for (int slot = S.slot0 + threadIdx.x; slot < S.slot1; slot += kBlockThreads) {
    double sum = 0.0;
    const int end = P.slot_off[slot + 1];
    for (int k = P.slot_off[slot]; k < end; ++k) {
      const int c = P.contrib[k];
      const double v = elem[static_cast<size_t>(c >> 5) * kElemStride + ((c >> 3) & 3) * 3 +
                            ((c >> 1) & 3)];
      sum += (c & 1) ? -v : v;
    }
    // The fixed air-gap element stiffness comes after every triangle term, as the host
    // assembler emits it (nonlinear_assembly.cu:134-137).
    if (P.age_vals != nullptr)
      sum += P.age_vals[slot];
    const int pos = P.slot_to_sell[slot];
    vals[pos] = sum;
    if (vals32)
      vals32[pos] = static_cast<float>(sum);
  }
}

// Row-owned rhs (physics/nonlinear_assembly.cu:142-186): PM (+ air-gap element) load and
// circuit sources, then the ordered element terms.  Also the Jacobi inverse (:188-206).
__device__ __attribute__((always_inline)) bool GatherRhs(const DevicePlan &P, const Span &S, const double *elem,
                          const double *vals, const double *cur, double *rhs, double *inv,
                          float *inv32)
{
  bool ok = true;
  for (int row = S.row0 + threadIdx.x; row < S.row1; row += kBlockThreads) {
    double s = P.pm_rhs[row];   // PM load + fixed air-gap element load, summed on the host
    for (int c = 0; c < P.C; ++c)
      s += cur[c] * P.circuit_rhs[static_cast<size_t>(c) * P.n + row];
    const int end = P.row_off[row + 1];
    for (int k = P.row_off[row]; k < end; ++k) {
      const int c = P.rcontrib[k];
      const double v = elem[static_cast<size_t>(c >> 3) * kElemStride + 9 + ((c >> 1) & 3)];
      s += (c & 1) ? -v : v;
    }
    rhs[row] = s;
    const double d = vals[P.diag_sell[row]];
    const bool good = isfinite(d) && d > 0.0;
    inv[row] = good ? 1.0 / d : 0.0;
    if (inv32)
      inv32[row] = good ? static_cast<float>(1.0 / d) : 0.0f;
    ok = ok && good;
  }
  return ok;
}

// ----------------------------------------------------------------------------- SpMV

// y = A v over SELL-64 (lane = row of the slice); returns the local sum of
// w[row] * y[row] when w is given.  A workgroup owns one item, so the memory
// parallelism of one CU must come from within each wave.  Slices of width
// <= kUnrolledWidth are software-pipelined: the gathers of slice s are issued
// first, then the column and value loads of the wave's next slice, so waiting for
// the gathers (vector loads complete in order) leaves the next slice's loads in
// flight; every slice costs one memory round trip instead of two.  The sum order
// over a row's entries is the CSR order either way.
constexpr int kUnrolledWidth = 10;
constexpr int kNarrowWidth = 9;   // the pipelined products' unroll when no slice is wider

template <typename T, int U = kUnrolledWidth>
struct SliceLoads {
  int c[U];
  T a[U];
  int width;
};

// Row of lane `lane` of slice s, -1 for padding lanes past the last row.
__device__ inline int SellRow(const DevicePlan &P, int s, int lane)
{
  const int row = s * kSliceRows + lane;
  return row < P.n ? row : -1;
}

// Branch-free loads of one slice: entries past the slice width repeat the last column
// (a valid address in the same cache line) with a zero value, so the whole slice is one
// straight-line run of loads and later adds only +-0 terms.
template <typename T, int U>
__device__ inline void LoadSlice(const DevicePlan &P, const T *__restrict__ vals, int s, int lane,
                                 SliceLoads<T, U> &L)
{
  const int base = P.slice_ptr[s] + lane;
  L.width = P.slice_width[s];
  const int last = L.width - 1;   // every slice holds a diagonal, so width >= 1
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int pos = base + max(min(k, last), 0) * kSliceRows;
    L.c[k] = P.sell_cols[pos];
    const T value = vals[pos];
    L.a[k] = k <= last ? value : T(0);
  }
}

// Generic (unpipelined) product of one slice wider than kUnrolledWidth.
template <typename T, typename V>
__device__ inline T WideSlice(const DevicePlan &P, const T *__restrict__ vals,
                              const V *__restrict__ v, int s, int lane)
{
  const int base = P.slice_ptr[s] + lane, width = P.slice_width[s];
  T sum = 0;
  for (int k = 0; k < width; ++k) {
    const int pos = base + k * kSliceRows;
    sum += vals[pos] * static_cast<T>(v[P.sell_cols[pos]]);
  }
  return sum;
}

// Two slices per wave per step (slices s and s + kWaves): twice the gathers in flight per
// round trip for the FP32 iterations, whose loads take half the registers of FP64 ones.
// U entries per slice are loaded whatever its width (the last column repeated): a plan
// whose slices are all at most kNarrowWidth wide takes the narrower unroll (Spmv2Fit,
// SpmvFit), one load of each kind fewer per slice.
template <typename T, typename V, int U = kUnrolledWidth>
__device__ __attribute__((always_inline)) double Spmv2(const DevicePlan &P, const Span &S,
                                                        const T *__restrict__ vals,
                        const V *__restrict__ v, V *__restrict__ y, const V *__restrict__ w)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double local = 0.0;
  const int first = S.slice0 + wave;
  if (first >= S.slice1)
    return local;
  auto clamp = [&](int s) { return s < S.slice1 ? s : first; };
  SliceLoads<T, U> c0, c1;
  LoadSlice(P, vals, first, lane, c0);
  LoadSlice(P, vals, clamp(first + kWaves), lane, c1);
  for (int s = first; s < S.slice1; s += 2 * kWaves) {
    const int s1 = s + kWaves;
    const int row0 = SellRow(P, s, lane), row1 = s1 < S.slice1 ? SellRow(P, s1, lane) : -1;
    T sum0 = 0, sum1 = 0;
    if (c0.width <= U && c1.width <= U) {
      V g0[U], g1[U];
#pragma unroll
      for (int k = 0; k < U; ++k)
        g0[k] = v[c0.c[k]];
#pragma unroll
      for (int k = 0; k < U; ++k)
        g1[k] = v[c1.c[k]];
      SliceLoads<T, U> n0, n1;
      // Past the end, the first slice is loaded again (in bounds, never used).
      LoadSlice(P, vals, clamp(s + 2 * kWaves), lane, n0);
      LoadSlice(P, vals, clamp(s + 3 * kWaves), lane, n1);
#pragma unroll
      for (int k = 0; k < U; ++k)
        sum0 += c0.a[k] * static_cast<T>(g0[k]);
#pragma unroll
      for (int k = 0; k < U; ++k)
        sum1 += c1.a[k] * static_cast<T>(g1[k]);
      c0 = n0;
      c1 = n1;
    } else {
      sum0 = WideSlice(P, vals, v, s, lane);
      if (s1 < S.slice1)
        sum1 = WideSlice(P, vals, v, s1, lane);
      LoadSlice(P, vals, clamp(s + 2 * kWaves), lane, c0);
      LoadSlice(P, vals, clamp(s + 3 * kWaves), lane, c1);
    }
    if (row0 >= 0) {
      y[row0] = static_cast<V>(sum0);
      if (w)
        local += static_cast<double>(w[row0]) * static_cast<double>(sum0);
    }
    if (row1 >= 0) {
      y[row1] = static_cast<V>(sum1);
      if (w)
        local += static_cast<double>(w[row1]) * static_cast<double>(sum1);
    }
  }
  return local;
}

template <typename T, typename V>
__device__ __attribute__((always_inline)) double Spmv2Fit(const DevicePlan &P, const Span &S,
                                                           const T *__restrict__ vals,
                           const V *__restrict__ v, V *__restrict__ y, const V *__restrict__ w)
{
  if (P.max_slice_width <= kNarrowWidth)
    return Spmv2<T, V, kNarrowWidth>(P, S, vals, v, y, w);
  return Spmv2<T, V>(P, S, vals, v, y, w);
}

// kPipeline = false: the plain two-round-trip loop, for rarely used products (FP64
// residuals inside the mixed form, equation residuals) where registers matter more.
// U: entries loaded per slice by the pipelined loop (SpmvFit picks it as Spmv2Fit does).
template <typename T, typename V, bool kPipeline = true, int U = kUnrolledWidth>
__device__ __attribute__((always_inline)) double Spmv(const DevicePlan &P, const Span &S,
                                                       const T *__restrict__ vals,
                       const V *__restrict__ v, V *__restrict__ y, const V *__restrict__ w)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double local = 0.0;
  if constexpr (!kPipeline) {
    for (int s = S.slice0 + wave; s < S.slice1; s += kWaves) {
      const int base = P.slice_ptr[s] + lane, width = P.slice_width[s];
      T sum = 0;
      if (width <= kUnrolledWidth) {
        int c[kUnrolledWidth];
        T a[kUnrolledWidth];
#pragma unroll
        for (int k = 0; k < kUnrolledWidth; ++k)
          if (k < width) {
            c[k] = P.sell_cols[base + k * kSliceRows];
            a[k] = vals[base + k * kSliceRows];
          }
        V g[kUnrolledWidth];
#pragma unroll
        for (int k = 0; k < kUnrolledWidth; ++k)
          if (k < width)
            g[k] = v[c[k]];
#pragma unroll
        for (int k = 0; k < kUnrolledWidth; ++k)
          if (k < width)
            sum += a[k] * static_cast<T>(g[k]);
      } else {
        for (int k = 0; k < width; ++k) {
          const int pos = base + k * kSliceRows;
          sum += vals[pos] * static_cast<T>(v[P.sell_cols[pos]]);
        }
      }
      const int row = SellRow(P, s, lane);
      if (row >= 0) {
        y[row] = static_cast<V>(sum);
        if (w)
          local += static_cast<double>(w[row]) * static_cast<double>(sum);
      }
    }
    return local;
  }
  SliceLoads<T, U> cur;
  cur.width = 0;
  if (S.slice0 + wave < S.slice1)
    LoadSlice(P, vals, S.slice0 + wave, lane, cur);
  for (int s = S.slice0 + wave; s < S.slice1; s += kWaves) {
    T sum = 0;
    const int next = s + kWaves;
    const int row = SellRow(P, s, lane);
    if (cur.width <= U) {
      V g[U];
#pragma unroll
      for (int k = 0; k < U; ++k)
        g[k] = v[cur.c[k]];
      SliceLoads<T, U> nxt;
      // The last slice prefetches itself again (in bounds, never used).
      LoadSlice(P, vals, next < S.slice1 ? next : s, lane, nxt);
#pragma unroll
      for (int k = 0; k < U; ++k)
        sum += cur.a[k] * static_cast<T>(g[k]);
      cur = nxt;
    } else {
      const int base = P.slice_ptr[s] + lane;
      for (int k = 0; k < cur.width; ++k) {
        const int pos = base + k * kSliceRows;
        sum += vals[pos] * static_cast<T>(v[P.sell_cols[pos]]);
      }
      cur.width = 0;
      if (next < S.slice1)
        LoadSlice(P, vals, next, lane, cur);
    }
    if (row >= 0) {
      y[row] = static_cast<V>(sum);
      if (w)
        local += static_cast<double>(w[row]) * static_cast<double>(sum);
    }
  }
  return local;
}

template <typename T, typename V>
__device__ __attribute__((always_inline)) double SpmvFit(const DevicePlan &P, const Span &S,
                                                          const T *__restrict__ vals,
                          const V *__restrict__ v, V *__restrict__ y, const V *__restrict__ w)
{
  if (P.max_slice_width <= kNarrowWidth)
    return Spmv<T, V, true, kNarrowWidth>(P, S, vals, v, y, w);
  return Spmv<T, V, true>(P, S, vals, v, y, w);
}

// Team form of the product that opens a PCG iteration: the search direction
// p = z + beta p_old is formed on the fly from the gathered z and p_old (the same two
// operations as the direction pass, so p and y are bitwise those of the unfused
// iteration) and stored for the workgroup's own rows, y = A p, and the local sum of
// p[row] * y[row] is returned.  The direction pass and the team barrier after it drop out
// of the iteration; p_old and p_new must be different vectors (other workgroups gather
// p_old while this one writes p_new).
template <typename T>
__device__ __attribute__((always_inline)) double SpmvDirection(
    const DevicePlan &P, const Span &S, const T *__restrict__ vals, const T *__restrict__ z,
    const T *__restrict__ pold, T beta, T *__restrict__ pnew, T *__restrict__ y)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double local = 0.0;
  SliceLoads<T> cur;
  cur.width = 0;
  if (S.slice0 + wave < S.slice1)
    LoadSlice(P, vals, S.slice0 + wave, lane, cur);
  for (int s = S.slice0 + wave; s < S.slice1; s += kWaves) {
    T sum = 0;
    const int next = s + kWaves;
    const int row = SellRow(P, s, lane);
    if (cur.width <= kUnrolledWidth) {
      T g[kUnrolledWidth];
#pragma unroll
      for (int k = 0; k < kUnrolledWidth; ++k)
        g[k] = z[cur.c[k]] + beta * pold[cur.c[k]];
      SliceLoads<T> nxt;
      // The last slice prefetches itself again (in bounds, never used).
      LoadSlice(P, vals, next < S.slice1 ? next : s, lane, nxt);
#pragma unroll
      for (int k = 0; k < kUnrolledWidth; ++k)
        sum += cur.a[k] * g[k];
      cur = nxt;
    } else {
      const int base = P.slice_ptr[s] + lane;
      for (int k = 0; k < cur.width; ++k) {
        const int pos = base + k * kSliceRows;
        const int c = P.sell_cols[pos];
        sum += vals[pos] * (z[c] + beta * pold[c]);
      }
      cur.width = 0;
      if (next < S.slice1)
        LoadSlice(P, vals, next, lane, cur);
    }
    if (row >= 0) {
      const T pr = z[row] + beta * pold[row];
      pnew[row] = pr;
      y[row] = sum;
      local += static_cast<double>(pr) * static_cast<double>(sum);
    }
  }
  return local;
}

// The hot PCG vector passes move 16 bytes per lane per load (two doubles or four floats,
// consecutive rows): the per-CU rate of vector memory instructions, not HBM, bounds
// one-workgroup-per-item passes with 4-byte loads.  Per-item vectors are padded to
// VectorStride rows (zero past n), so every 16-byte chunk is aligned and chunks past
// the last row only ever hold zeros.
constexpr int kChunkUnroll = 2;   // chunks per thread in flight per step

template <typename T>
struct Chunk {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

template <typename T>
__device__ inline Chunk<T> LoadChunk(const T *p)
{
  Chunk<T> c;
  if constexpr (sizeof(T) == 8) {
    const double2 t = *reinterpret_cast<const double2 *>(p);
    c.v[0] = t.x;
    c.v[1] = t.y;
  } else {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    c.v[0] = t.x;
    c.v[1] = t.y;
    c.v[2] = t.z;
    c.v[3] = t.w;
  }
  return c;
}

template <typename T>
__device__ inline void StoreChunk(T *p, const Chunk<T> &c)
{
  if constexpr (sizeof(T) == 8) {
    *reinterpret_cast<double2 *>(p) = make_double2(c.v[0], c.v[1]);
  } else {
    *reinterpret_cast<float4 *>(p) = make_float4(c.v[0], c.v[1], c.v[2], c.v[3]);
  }
}

// ----------------------------------------------------------------------------- block-Jacobi

// Block-Jacobi preconditioner (PcgPreconditioner::kBlockJacobi, solvers/options.h:8-10):
// blocks of bs consecutive free rows (bs = 2, 4 or 8 divides the 64 rows of a wave, so a
// block never leaves a wave or a partition's slice range).  planes[j][i] multiplies the
// residual of row j of i's block: z_i = sum_j planes[j][i] r_(first + j), in j order.
template <typename T>
struct BlockPrec {
  const T *planes;   // [kMaxBlock][VectorStride(n)] of this item, or nullptr
  int bs;            // 1: point Jacobi by the inverse diagonal
  int n;             // plane stride (VectorStride of the free-row count)
};

// All LDS accesses of this wave issued so far are complete before any later one
// (lanes of one wave exchanging values through LDS).
__device__ inline void WaveLdsSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// z_i for row i; every lane of the wave calls it with its row (valid = row inside the
// span; invalid lanes pass r = 0).  Rows are lane-contiguous: row i sits in lane i % 64.
template <typename T>
__device__ inline T BlockApply(const BlockPrec<T> &B, int i, bool valid, T ri)
{
  const int lane = threadIdx.x & 63;
  const int lead = lane & ~(B.bs - 1);
  T pj[kMaxBlock];
#pragma unroll
  for (int j = 0; j < kMaxBlock; ++j)
    pj[j] = (j < B.bs && valid) ? B.planes[static_cast<size_t>(j) * B.n + i] : T(0);
  T z = T(0);
#pragma unroll
  for (int j = 0; j < kMaxBlock; ++j)
    if (j < B.bs) {
      const T rj = __shfl(ri, lead + j, 64);
      z += pj[j] * rj;
    }
  return z;
}

// Blocks of 4 rows held whole by one thread: the thread's 4 consecutive rows arrive as
// 16-byte chunks (one float4, or two double2), the block's 4 planes likewise (plane j's
// chunk is column j of the block inverse), so z = M^-1 r needs no exchange between lanes.
// Rows past the span's last row are zero padding in every vector and plane.
template <typename T>
struct Rows4 {
  T v[4];
};

template <typename T>
__device__ inline Rows4<T> LoadRows4(const T *p)
{
  Rows4<T> q;
  if constexpr (sizeof(T) == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    q.v[0] = t.x;
    q.v[1] = t.y;
    q.v[2] = t.z;
    q.v[3] = t.w;
  } else {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(p + 2);
    q.v[0] = a.x;
    q.v[1] = a.y;
    q.v[2] = b.x;
    q.v[3] = b.y;
  }
  return q;
}

template <typename T>
__device__ inline void StoreRows4(T *p, const Rows4<T> &q)
{
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
  } else {
    *reinterpret_cast<double2 *>(p) = make_double2(q.v[0], q.v[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(q.v[2], q.v[3]);
  }
}

// z for one block of 4 rows from its planes: z_k = sum_j Minv[k][j] r_j (j in order).
template <typename T>
__device__ inline Rows4<T> Block4Apply(const BlockPrec<T> &B, int row, const Rows4<T> &r)
{
  Rows4<T> pl[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    pl[j] = LoadRows4(B.planes + static_cast<size_t>(j) * B.n + row);
  Rows4<T> z;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    T s = T(0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      s += pl[j].v[k] * r.v[j];
    z.v[k] = s;
  }
  return z;
}

// BuildBlockInverse (solvers/block_jacobi_preconditioner.cu:73-168) for the rows of span S:
// each block's principal submatrix, scaled by sqrt(diag) on both sides, must have finite
// reciprocal diagonals >= DBL_MIN and be symmetric to 1e-12 relative to max(1, |a|, |b|);
// the block's first lane Cholesky-factors it in LDS, then lane k solves column k of the
// inverse by forward / back substitution and unscales it.  An unusable block becomes the
// point Jacobi rows (the reference's NaN-marked fallback, pcg_device.cuh:56-72).  Lane k
// writes plane k of its block's rows (column k of the inverse).
// The plan arrays and the row range come as scalars: this out-of-line build must not take
// the address of the caller's plan or team, which would keep those in scratch memory for
// the whole solve.
__device__ void BuildBlockJacobi(int n_rows, const int32_t *slice_ptr, const int32_t *slice_width,
                                 const int32_t *sell_cols, const int32_t *diag_sell, int row0,
                                 int row1, const double *vals, const double *inv, int bs,
                                 double *planes, float *planes32)
{
  __shared__ double blk[kWaves * (64 * kMaxBlock + 128)];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *F_all = blk + wave * (64 * kMaxBlock + 128);
  double *sc = F_all + 64 * kMaxBlock;         // sqrt(diag) by lane
  double *flag = sc + 64;                      // per block: 1 = unusable
  const int bl = lane / bs, k = lane - bl * bs, lead = bl * bs;
  double *F = F_all + bl * bs * bs;
  const size_t n = VectorStride(n_rows);   // plane stride
  for (int base = row0 + wave * 64; base < row1; base += kBlockThreads) {
    const int i = base + lane, first = base + lead;
    const bool valid = i < row1;
    const int width = max(0, min(bs, row1 - first));
    const double d = valid ? vals[diag_sell[i]] : 1.0;
    sc[lane] = valid ? sqrt(d) : 1.0;
    for (int c = 0; c < bs; ++c)
      F[k * bs + c] = 0.0;
    WaveLdsSync();
    bool bad_row = false;
    if (valid) {
      const double reciprocal = 1.0 / d;
      bad_row = !isfinite(reciprocal) || reciprocal < DBL_MIN;
      const int slice = i / kSliceRows;
      const int pos0 = slice_ptr[slice] + lane, entries = slice_width[slice];
      for (int e = 0; e < entries; ++e) {
        const int pos = pos0 + e * kSliceRows;
        const int c = sell_cols[pos] - first;
        if (c >= 0 && c < width)
          F[k * bs + c] += vals[pos] / sc[lane] / sc[lead + c];
      }
    }
    const uint64_t bad_rows = __ballot(bad_row);
    WaveLdsSync();
    if (k == 0 && width > 0) {
      bool bad = ((bad_rows >> lead) & ((1ull << bs) - 1ull)) != 0;
      for (int r = 0; r < width; ++r)
        for (int c = 0; c <= r; ++c) {
          const double u = F[r * bs + c], v = F[c * bs + r];
          if (!isfinite(u) || !isfinite(v) || fabs(u - v) > 1e-12 * fmax(1.0, fmax(fabs(u), fabs(v))))
            bad = true;
        }
      for (int c = 0; c < width && !bad; ++c) {
        double pivot = F[c * bs + c];
        for (int q = 0; q < c; ++q)
          pivot -= F[c * bs + q] * F[c * bs + q];
        if (!(pivot > 0.0) || !isfinite(pivot)) {
          bad = true;
          break;
        }
        F[c * bs + c] = sqrt(pivot);
        for (int r = c + 1; r < width; ++r) {
          double v = F[r * bs + c];
          for (int q = 0; q < c; ++q)
            v -= F[r * bs + q] * F[c * bs + q];
          F[r * bs + c] = v / F[c * bs + c];
        }
      }
      flag[bl] = bad ? 1.0 : 0.0;
    }
    WaveLdsSync();
    bool bad = width > 0 && flag[bl] != 0.0;
    double x[kMaxBlock];
    bool broken = false;
    if (!bad && k < width) {
#pragma unroll
      for (int r = 0; r < kMaxBlock; ++r)
        if (r < width) {
          double v = r == k ? 1.0 : 0.0;
#pragma unroll
          for (int q = 0; q < r; ++q)
            v -= F[r * bs + q] * x[q];
          x[r] = v / F[r * bs + r];
        }
#pragma unroll
      for (int r = kMaxBlock - 1; r >= 0; --r)
        if (r < width) {
          double v = x[r];
#pragma unroll
          for (int q = r + 1; q < kMaxBlock; ++q)
            if (q < width)
              v -= F[q * bs + r] * x[q];
          x[r] = v / F[r * bs + r];
        }
#pragma unroll
      for (int r = 0; r < kMaxBlock; ++r)
        if (r < width) {
          x[r] = x[r] / sc[lead + r] / sc[lane];
          broken = broken || !isfinite(x[r]);
        }
    }
    const uint64_t broken_lanes = __ballot(broken);
    bad = bad || ((broken_lanes >> lead) & ((1ull << bs) - 1ull)) != 0;
    // Lane k writes plane k at its block's rows; rows past the span get nothing.
#pragma unroll
    for (int r = 0; r < kMaxBlock; ++r)
      if (r < width) {
        double v;
        if (bad)
          v = r == k ? inv[first + r] : 0.0;
        else
          v = k < width ? x[r] : 0.0;
        planes[static_cast<size_t>(k) * n + first + r] = v;
        if (planes32)
          planes32[static_cast<size_t>(k) * n + first + r] = static_cast<float>(v);
      }
    WaveLdsSync();   // the block storage is rewritten by the next chunk
  }
}

// ----------------------------------------------------------------------------- FP64 PCG

struct LinearInfo {
  int status;
  int iterations;
  double residual;
};

// Jacobi PCG of the reference (solvers/pcg_kernels.cu:10-344) for one item:
// rhs scaled by max|b|, fused update of x, r and z = D^-1 r, acceptance only on
// the recomputed true residual with restart otherwise.  With warm != 0 the
// iteration starts from the x already in `x` (unscaled) instead of zero; the
// acceptance test ||b - A x|| <= tol ||b|| is unchanged.
// kBlock: 0 = point Jacobi; 1 = block-Jacobi of B.bs rows (4 in-thread, others by lane
// exchange); 2 = block-Jacobi of 4 rows only (the lane-exchange code is not compiled).
// stop_on_stall (a tolerance below what FP64 rounding lets the true residual reach, as a
// tight acceptance re-solve may ask): a restart whose true residual is not below half the
// previous one ends the solve with the iteration cap's status instead of restarting until
// maxit; the reference's solves keep the default, restarting until maxit.
// A team of workgroups (Team other than BlockTeam) forms each search direction inside
// the next product (SpmvDirection) instead of a pass of its own: p alternates between
// p and p_alt, A p goes to q_buf and z = M^-1 r to ap, one team barrier fewer per
// iteration with the same arithmetic.  One workgroup per item keeps the separate direction pass (its
// barrier is a block barrier) and needs neither p_alt nor q_buf.
template <bool kPipeline = true, int kBlock = 0, class Team = BlockTeam>
__device__ __attribute__((always_inline)) LinearInfo Pcg64(const DevicePlan &P, const double *vals, const double *inv,
                            const BlockPrec<double> &B, const double *rhs, double *x, double *r,
                            double *p, double *ap, double tol, int maxit, bool warm, Team &R,
                            double *p_alt = nullptr, double *q_buf = nullptr,
                            bool stop_on_stall = false)
{
  constexpr bool kFuse = !std::is_same_v<Team, BlockTeam>;
  constexpr bool block = kBlock != 0;
  constexpr bool only4 = kBlock == 2;
  const bool fast4 = only4 || B.bs == 4;
  const int lane = threadIdx.x & 63;
  const int wave_row0 = R.S.row0 + (threadIdx.x & ~63);   // block mode: wave-uniform row loops
  LinearInfo info{FEMM_INTERNAL_ERROR, 0, INFINITY};
  const int n = P.n;
  if (maxit < 0 || !(tol > 0.0) || !(tol < 1.0) || !isfinite(tol)) {
    info.status = FEMM_INVALID_ARGUMENT;
    return info;
  }
  // First-invalid-row precedence of the reference's validation scan.
  double bad = 0.0, bmax = 0.0;
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
    const double d = vals[P.diag_sell[i]];
    if (!isfinite(rhs[i]) || !isfinite(d) || !(d > 0.0))
      bad = fmax(bad, static_cast<double>(n - i));
    bmax = fmax(bmax, fabs(rhs[i]));
  }
  const double first = R.Max(bad);
  if (first != 0.0) {
    const int row = n - static_cast<int>(first);
    const double d = vals[P.diag_sell[row]];
    info.status = (!isfinite(rhs[row]) || !isfinite(d)) ? FEMM_NUMERICAL_NONFINITE
                                                        : FEMM_LINEAR_SOLVE_BREAKDOWN;
    return info;
  }
  const double scale = R.Max(bmax);
  if (scale == 0.0) {
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
      x[i] = 0.0;
    R.Sync();
    info.status = FEMM_OK;
    info.residual = 0.0;
    return info;
  }
  if (warm) {
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
      x[i] = x[i] / scale;
    R.Sync();
    Spmv<double, double, false>(P, R.S, vals, x, ap, static_cast<const double *>(nullptr));
    R.Sync();
  }
  double s0[3] = {0.0, 0.0, 0.0};
  if (block && fast4) {
    const int q1 = (R.S.row1 + 3) / 4;
    for (int q = R.S.row0 / 4 + threadIdx.x; q < q1; q += kBlockThreads) {
      const int row = 4 * q;
      const Rows4<double> bv = LoadRows4(rhs + row);
      const Rows4<double> av = warm ? LoadRows4(ap + row) : Rows4<double>{};
      Rows4<double> rv;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double b = bv.v[e] / scale;
        rv.v[e] = warm ? b - av.v[e] : b;
        s0[0] += rv.v[e] * rv.v[e];
        s0[2] += b * b;
      }
      const Rows4<double> zv = Block4Apply(B, row, rv);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        s0[1] += rv.v[e] * zv.v[e];
      if (!warm)
        StoreRows4(x + row, Rows4<double>{});
      StoreRows4(r + row, rv);
      StoreRows4(p + row, zv);
    }
  } else if constexpr (block && !only4) {
    for (int base = wave_row0; base < R.S.row1; base += kBlockThreads) {
      const int i = base + lane;
      const bool valid = i < R.S.row1;
      double b = 0.0, ri = 0.0;
      if (valid) {
        b = rhs[i] / scale;
        ri = warm ? b - ap[i] : b;
        if (!warm)
          x[i] = 0.0;
        r[i] = ri;
      }
      const double zi = BlockApply(B, i, valid, ri);
      if (valid) {
        p[i] = zi;
        s0[0] += ri * ri;
        s0[1] += ri * zi;
        s0[2] += b * b;
      }
    }
  } else {
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
      const double b = rhs[i] / scale;
      const double ri = warm ? b - ap[i] : b;
      if (!warm)
        x[i] = 0.0;
      const double zi = ri * inv[i];
      r[i] = ri;
      p[i] = zi;
      s0[0] += ri * ri;
      s0[1] += ri * zi;
      s0[2] += b * b;
    }
  }
  R.Sum(s0);
  if (!isfinite(s0[0]) || !isfinite(s0[1])) {
    info.status = FEMM_NUMERICAL_NONFINITE;
    return info;
  }
  info.residual = sqrt(s0[0]);
  double rho = s0[1];
  const double threshold = tol * sqrt(s0[2]);
  if (warm && info.residual <= threshold) {
    // The start already satisfies the acceptance test.
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
      x[i] *= scale;
    R.Sync();
    info.status = FEMM_OK;
    info.residual *= scale;
    return info;
  }
  double *pc = p;                  // current search direction
  const bool fuse = kFuse && TeamSize(R) > 1;   // a one-workgroup team keeps the pass
  double *pn = fuse ? p_alt : p;   // the next one (fused form)
  double *qv = fuse ? q_buf : ap;  // A p (the unfused form keeps z over it after the update)
  bool fresh = true;               // pc holds the direction itself (start or restart)
  double beta = 0.0;
  double last_true = INFINITY;     // true residual at the last restart
  for (int it = 0; it < maxit; ++it) {
    double pap[1];
    if (fuse && !fresh) {
      pap[0] = SpmvDirection<double>(P, R.S, vals, ap, pc, beta, pn, qv);
      double *t = pc;
      pc = pn;
      pn = t;
    } else {
      if constexpr (kPipeline)
        pap[0] = SpmvFit<double, double>(P, R.S, vals, pc, qv, pc);
      else
        pap[0] = Spmv<double, double, false>(P, R.S, vals, pc, qv, pc);
    }
    fresh = false;
    R.Sum(pap);
    if (!isfinite(pap[0]) || !isfinite(rho)) {
      info.status = FEMM_NUMERICAL_NONFINITE;
      return info;
    }
    if (!(pap[0] > 0.0) || !(rho > 0.0)) {
      info.status = FEMM_LINEAR_SOLVE_BREAKDOWN;
      return info;
    }
    const double alpha = rho / pap[0];
    double s1[2] = {0.0, 0.0};
    if constexpr (block) {
      // x, r updated; z = M^-1 r written to ap (over A p in the unfused form).
      if (fast4) {
        const int q1 = (R.S.row1 + 3) / 4;
        for (int q = R.S.row0 / 4 + threadIdx.x; q < q1; q += kBlockThreads) {
          const int row = 4 * q;
          Rows4<double> xv = LoadRows4(x + row), pv = LoadRows4(pc + row),
                        rv = LoadRows4(r + row), av = LoadRows4(qv + row);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            xv.v[e] = xv.v[e] + alpha * pv.v[e];
            rv.v[e] = rv.v[e] - alpha * av.v[e];
            s1[0] += rv.v[e] * rv.v[e];
          }
          const Rows4<double> zv = Block4Apply(B, row, rv);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            s1[1] += rv.v[e] * zv.v[e];
          StoreRows4(x + row, xv);
          StoreRows4(r + row, rv);
          StoreRows4(ap + row, zv);
        }
      } else if constexpr (!only4) {
        for (int base = wave_row0; base < R.S.row1; base += kBlockThreads) {
          const int i = base + lane;
          const bool valid = i < R.S.row1;
          double ri = 0.0;
          if (valid) {
            x[i] = x[i] + alpha * pc[i];
            ri = r[i] - alpha * qv[i];
            r[i] = ri;
            s1[0] += ri * ri;
          }
          const double zi = BlockApply(B, i, valid, ri);
          if (valid) {
            ap[i] = zi;
            s1[1] += ri * zi;
          }
        }
      }
    } else {
      constexpr int V = Chunk<double>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c0 = R.S.row0 / V + threadIdx.x; c0 < c1; c0 += kBlockThreads * kChunkUnroll) {
        Chunk<double> xv[kChunkUnroll], pv[kChunkUnroll], rv[kChunkUnroll], av[kChunkUnroll],
            iv[kChunkUnroll];
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
            xv[u] = LoadChunk(x + c * V);
            pv[u] = LoadChunk(pc + c * V);
            rv[u] = LoadChunk(r + c * V);
            av[u] = LoadChunk(qv + c * V);
            iv[u] = LoadChunk(inv + c * V);
          }
        }
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
              xv[u].v[e] = xv[u].v[e] + alpha * pv[u].v[e];
              const double ri = rv[u].v[e] - alpha * av[u].v[e];
              rv[u].v[e] = ri;
              s1[0] += ri * ri;
              const double zi = ri * iv[u].v[e];
              s1[1] += ri * zi;
              iv[u].v[e] = zi;
            }
            StoreChunk(x + c * V, xv[u]);
            StoreChunk(r + c * V, rv[u]);
            if (fuse)
              StoreChunk(ap + c * V, iv[u]);   // z for the next product
          }
        }
      }
    }
    R.Sum(s1);
    info.iterations = it + 1;
    if (!isfinite(s1[0])) {
      info.status = FEMM_NUMERICAL_NONFINITE;
      return info;
    }
    info.residual = sqrt(s1[0]);
    if (info.residual <= threshold) {
      // True residual on the recomputed b/scale - A x (pcg_kernels.cu:192-270).
      Spmv<double, double, false>(P, R.S, vals, x, ap, static_cast<const double *>(nullptr));
      R.Sync();
      double tr[1] = {0.0};
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
        const double ri = rhs[i] / scale - ap[i];
        r[i] = ri;
        tr[0] += ri * ri;
      }
      R.Sum(tr);
      if (!isfinite(tr[0])) {
        info.status = FEMM_NUMERICAL_NONFINITE;
        return info;
      }
      info.residual = sqrt(tr[0]);
      if (stop_on_stall && !(info.residual <= threshold) && !(info.residual < 0.5 * last_true))
        break;   // the restarts no longer lower the true residual: report the cap's status
      last_true = info.residual;
      if (info.residual <= threshold) {
        bool finite = true;
        for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
          x[i] *= scale;
          finite = finite && isfinite(x[i]);
        }
        const bool nonfinite = R.AnyTrue(!finite);
        info.residual *= scale;
        info.status = (nonfinite || !isfinite(info.residual)) ? FEMM_NUMERICAL_NONFINITE : FEMM_OK;
        return info;
      }
      double rs[1] = {0.0};
      if (block && fast4) {
        const int q1 = (R.S.row1 + 3) / 4;
        for (int q = R.S.row0 / 4 + threadIdx.x; q < q1; q += kBlockThreads) {
          const int row = 4 * q;
          const Rows4<double> rv = LoadRows4(r + row);
          const Rows4<double> zv = Block4Apply(B, row, rv);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            rs[0] += rv.v[e] * zv.v[e];
          StoreRows4(pc + row, zv);
        }
      } else if constexpr (block && !only4) {
        for (int base = wave_row0; base < R.S.row1; base += kBlockThreads) {
          const int i = base + lane;
          const bool valid = i < R.S.row1;
          const double ri = valid ? r[i] : 0.0;
          const double zi = BlockApply(B, i, valid, ri);
          if (valid) {
            pc[i] = zi;
            rs[0] += ri * zi;
          }
        }
      } else {
        for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
          const double zi = r[i] * inv[i];
          pc[i] = zi;
          rs[0] += r[i] * zi;
        }
      }
      R.Sum(rs);
      if (!isfinite(rs[0]) || !(rs[0] > 0.0)) {
        info.status = FEMM_LINEAR_SOLVE_BREAKDOWN;
        return info;
      }
      rho = rs[0];
      fresh = true;   // the restart direction is z itself
      continue;
    }
    if (!isfinite(s1[1]) || !(s1[1] > 0.0)) {
      info.status = FEMM_LINEAR_SOLVE_BREAKDOWN;
      return info;
    }
    beta = s1[1] / rho;
    rho = s1[1];
    if (fuse)
      continue;   // the next product forms p = z + beta p (z is published by the sum above)
    if constexpr (block) {
      constexpr int V = Chunk<double>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c = R.S.row0 / V + threadIdx.x; c < c1; c += kBlockThreads) {
        const Chunk<double> zv = LoadChunk(ap + c * V);
        Chunk<double> pv = LoadChunk(p + c * V);
#pragma unroll
        for (int e = 0; e < V; ++e)
          pv.v[e] = zv.v[e] + beta * pv.v[e];
        StoreChunk(p + c * V, pv);
      }
    } else {
      constexpr int V = Chunk<double>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c0 = R.S.row0 / V + threadIdx.x; c0 < c1; c0 += kBlockThreads * kChunkUnroll) {
        Chunk<double> rv[kChunkUnroll], iv[kChunkUnroll], pv[kChunkUnroll];
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
            rv[u] = LoadChunk(r + c * V);
            iv[u] = LoadChunk(inv + c * V);
            pv[u] = LoadChunk(p + c * V);
          }
        }
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
#pragma unroll
            for (int e = 0; e < V; ++e)
              pv[u].v[e] = rv[u].v[e] * iv[u].v[e] + beta * pv[u].v[e];
            StoreChunk(p + c * V, pv[u]);
          }
        }
      }
    }
    R.Sync();
  }
  // Iteration cap: report the true residual (pcg_kernels.cu:302-343).
  Spmv<double, double, false>(P, R.S, vals, x, ap, static_cast<const double *>(nullptr));
  R.Sync();
  double fr[1] = {0.0};
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
    const double ri = rhs[i] / scale - ap[i];
    fr[0] += ri * ri;
  }
  R.Sum(fr);
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
    x[i] *= scale;
  R.Sync();
  info.residual = sqrt(fr[0]) * scale;
  info.status = isfinite(info.residual) ? FEMM_LINEAR_SOLVE_NOT_CONVERGED : FEMM_NUMERICAL_NONFINITE;
  return info;
}

// ----------------------------------------------------------------------------- mixed precision

struct RefinedInfo {
  LinearInfo linear;
  int refinements;   // FP64 true-residual evaluations
  int fp64_iterations;   // of linear.iterations, those of the FP64 finish
};

// r = b - A64 x in FP64; returns ||r||_2 and max|r| through the reducer.
template <class Team>
__device__ __attribute__((always_inline)) void TrueResidual(const DevicePlan &P, const double *vals, const double *rhs,
                             const double *x, double *r, double *ap, double *norm, double *amax,
                             Team &R)
{
  Spmv<double, double, false>(P, R.S, vals, x, ap, static_cast<const double *>(nullptr));
  R.Sync();
  double rr[1] = {0.0};
  double rmax = 0.0;
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
    const double ri = rhs[i] - ap[i];
    r[i] = ri;
    rr[0] += ri * ri;
    rmax = fmax(rmax, fabs(ri));
  }
  R.Sum(rr);
  *amax = R.Max(rmax);
  *norm = sqrt(rr[0]);
}

// Mixed-precision Jacobi PCG with reliable updates (Sleijpen and van der Vorst,
// "Reliable updated residuals in hybrid Bi-CG methods", Computing 56, 1996): the
// iteration runs in FP32 (values, SpMV, preconditioner, vectors; dot products
// accumulated in FP64) on a correction d relative to the FP64 iterate x.  Whenever
// the FP32 recurrence residual has dropped by `delta` since the last FP64 residual,
// or claims convergence, x += sigma d and the FP64 true residual b - A64 x replaces
// the recurrence residual (FP32 vectors are kept in units of sigma = max|r| at the last
// replacement, so they never underflow); the search direction is kept, so the Krylov
// space is not restarted.  Acceptance is the reference's test on the FP64 true residual,
// ||b - A x|| <= tol ||b||.  A breakdown in FP32, a replacement that fails to halve the
// true residual, or running out of replacements finishes the solve in FP64 PCG
// from the current x.  Called with maxit = 0 and fp64_maxit > 0, that FP64 PCG (at most
// fp64_maxit iterations) is the whole solve.
template <bool kBlock, class Team>
__device__ __attribute__((always_inline)) RefinedInfo ReliableSolve(const DevicePlan &P, const double *vals, const double *inv,
                                     const float *vals32, const float *inv32,
                                     const BlockPrec<double> &B64, const BlockPrec<float> &B32,
                                     const double *rhs, double *x, double *r, double *p,
                                     double *ap, float *d32, float *r32, float *p32, float *ap32,
                                     double tol, int maxit, bool warm, int max_updates,
                                     double delta, Team &R, int fp64_maxit = 0,
                                     float *p32_alt = nullptr, float *q32_buf = nullptr,
                                     double *p_alt = nullptr, double *q_buf = nullptr,
                                     bool stop_on_stall = false)
{
  RefinedInfo out{{FEMM_INTERNAL_ERROR, 0, INFINITY}, 0, 0};
  constexpr bool block = kBlock;
  constexpr bool kFuse = !std::is_same_v<Team, BlockTeam>;   // as in Pcg64
  if (maxit < 0 || !(tol > 0.0) || !(tol < 1.0) || !isfinite(tol)) {
    out.linear.status = FEMM_INVALID_ARGUMENT;
    return out;
  }
  double b2[1] = {0.0};
  bool finite = true;
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
    b2[0] += rhs[i] * rhs[i];
    finite = finite && isfinite(rhs[i]);
    if (!warm)
      x[i] = 0.0;
  }
  R.Sum(b2);
  if (R.AnyTrue(!finite)) {
    out.linear.status = FEMM_NUMERICAL_NONFINITE;
    return out;
  }
  const double bnorm = sqrt(b2[0]);
  if (bnorm == 0.0) {
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
      x[i] = 0.0;
    R.Sync();
    out.linear = {FEMM_OK, 0, 0.0};
    return out;
  }
  const double threshold = tol * bnorm;
  if (!(delta > 0.0) || !(delta < 1.0))
    delta = 0.1;
  double rnorm, rmax;
  TrueResidual(P, vals, rhs, x, r, ap, &rnorm, &rmax, R);
  out.refinements = 1;
  out.linear.residual = rnorm;
  if (!isfinite(rnorm)) {
    out.linear.status = FEMM_NUMERICAL_NONFINITE;
    return out;
  }
  if (rnorm <= threshold) {
    out.linear.status = FEMM_OK;
    return out;
  }
  double sigma = rmax;
  double rho, rr;
  {
    double s0[2] = {0.0, 0.0};
    if constexpr (block) {
      const int q1 = (R.S.row1 + 3) / 4;
      for (int q = R.S.row0 / 4 + threadIdx.x; q < q1; q += kBlockThreads) {
        const int row = 4 * q;
        const Rows4<double> r64 = LoadRows4(r + row);
        Rows4<float> rv, dv;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          rv.v[e] = static_cast<float>(r64.v[e] / sigma);
          dv.v[e] = 0.0f;
          s0[0] += static_cast<double>(rv.v[e]) * rv.v[e];
        }
        const Rows4<float> zv = Block4Apply(B32, row, rv);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          s0[1] += static_cast<double>(rv.v[e]) * zv.v[e];
        StoreRows4(r32 + row, rv);
        StoreRows4(p32 + row, zv);
        StoreRows4(d32 + row, dv);
      }
    } else {
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
        const float ri = static_cast<float>(r[i] / sigma);
        const float zi = ri * inv32[i];
        r32[i] = ri;
        p32[i] = zi;
        d32[i] = 0.0f;
        s0[0] += static_cast<double>(ri) * ri;
        s0[1] += static_cast<double>(ri) * zi;
      }
    }
    R.Sum(s0);
    rr = s0[0];
    rho = s0[1];
  }
  double rr_ref = rr;           // ||r32||^2 at the last replacement
  double last_true = rnorm;     // FP64 true residual at the last replacement
  bool fallback = false;
  int it = 0;
  float *pc = p32;                      // current search direction
  constexpr bool fuse = kFuse;          // (a runtime choice here costs registers: spills)
  float *pn = fuse ? p32_alt : p32;     // the next one (fused form)
  float *qv = fuse ? q32_buf : ap32;    // A p; z = M^-1 r goes to ap32
  bool fresh = true;
  float beta = 0.0f;
  while (it < maxit) {
    // Two slices per step for a workgroup that owns the item (the same sums in the same
    // order as one slice per step); team members form the direction inside the product.
    double pap[1];
    if constexpr (kFuse) {
      if (fresh) {
        pap[0] = Spmv<float, float, true>(P, R.S, vals32, pc, qv, pc);
      } else {
        pap[0] = SpmvDirection<float>(P, R.S, vals32, ap32, pc, beta, pn, qv);
        float *t = pc;
        pc = pn;
        pn = t;
      }
      fresh = false;
    } else {
      pap[0] = Spmv2Fit(P, R.S, vals32, pc, qv, pc);
    }
    R.Sum(pap);
    if (!(pap[0] > 0.0) || !(rho > 0.0) || !isfinite(pap[0]) || !isfinite(rho)) {
      fallback = true;
      break;
    }
    const float alpha = static_cast<float>(rho / pap[0]);
    double s1[2] = {0.0, 0.0};
    if constexpr (block) {
      // d, r updated; z = M^-1 r written to ap32 (over A p in the unfused form).
      const int q1 = (R.S.row1 + 3) / 4;
      for (int q = R.S.row0 / 4 + threadIdx.x; q < q1; q += kBlockThreads) {
        const int row = 4 * q;
        Rows4<float> dv = LoadRows4(d32 + row), pv = LoadRows4(pc + row),
                     rv = LoadRows4(r32 + row), av = LoadRows4(qv + row);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          dv.v[e] = dv.v[e] + alpha * pv.v[e];
          rv.v[e] = rv.v[e] - alpha * av.v[e];
          s1[0] += static_cast<double>(rv.v[e]) * rv.v[e];
        }
        const Rows4<float> zv = Block4Apply(B32, row, rv);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          s1[1] += static_cast<double>(rv.v[e]) * zv.v[e];
        StoreRows4(d32 + row, dv);
        StoreRows4(r32 + row, rv);
        StoreRows4(ap32 + row, zv);
      }
    } else {
      constexpr int V = Chunk<float>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c0 = R.S.row0 / V + threadIdx.x; c0 < c1; c0 += kBlockThreads * kChunkUnroll) {
        Chunk<float> dv[kChunkUnroll], pv[kChunkUnroll], rv[kChunkUnroll], av[kChunkUnroll],
            iv[kChunkUnroll];
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
            dv[u] = LoadChunk(d32 + c * V);
            pv[u] = LoadChunk(pc + c * V);
            rv[u] = LoadChunk(r32 + c * V);
            av[u] = LoadChunk(qv + c * V);
            iv[u] = LoadChunk(inv32 + c * V);
          }
        }
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
              dv[u].v[e] = dv[u].v[e] + alpha * pv[u].v[e];
              const float ri = rv[u].v[e] - alpha * av[u].v[e];
              rv[u].v[e] = ri;
              s1[0] += static_cast<double>(ri) * ri;
              const float zi = ri * iv[u].v[e];
              s1[1] += static_cast<double>(ri) * zi;
              iv[u].v[e] = zi;
            }
            StoreChunk(d32 + c * V, dv[u]);
            StoreChunk(r32 + c * V, rv[u]);
            if constexpr (fuse)
              StoreChunk(ap32 + c * V, iv[u]);   // z for the next product
          }
        }
      }
    }
    R.Sum(s1);
    ++it;
    if (!isfinite(s1[0]) || !isfinite(s1[1])) {
      fallback = true;
      break;
    }
    double rho_next = s1[1];
    const bool claims = s1[0] * sigma * sigma <= threshold * threshold;
    if (claims || s1[0] <= delta * delta * rr_ref) {
      // Reliable update: fold the correction into x and replace the residual.
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
        x[i] += sigma * static_cast<double>(d32[i]);
        d32[i] = 0.0f;
      }
      R.Sync();
      TrueResidual(P, vals, rhs, x, r, ap, &rnorm, &rmax, R);
      ++out.refinements;
      out.linear.residual = rnorm;
      if (!isfinite(rnorm)) {
        out.linear.iterations = it;
        out.linear.status = FEMM_NUMERICAL_NONFINITE;
        return out;
      }
      if (rnorm <= threshold) {
        out.linear.iterations = it;
        out.linear.status = FEMM_OK;
        return out;
      }
      if (rnorm > 0.5 * last_true || out.refinements > max_updates) {
        fallback = true;
        break;
      }
      last_true = rnorm;
      // Rescale to the new residual unit; the direction carries over.
      const double q = sigma / rmax;
      sigma = rmax;
      double s2[2] = {0.0, 0.0};
      if constexpr (block) {
        const int q1 = (R.S.row1 + 3) / 4;
        for (int c = R.S.row0 / 4 + threadIdx.x; c < q1; c += kBlockThreads) {
          const int row = 4 * c;
          const Rows4<double> r64 = LoadRows4(r + row);
          Rows4<float> rv, pv = LoadRows4(pc + row);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            rv.v[e] = static_cast<float>(r64.v[e] / sigma);
            pv.v[e] = static_cast<float>(q * static_cast<double>(pv.v[e]));
            s2[0] += static_cast<double>(rv.v[e]) * rv.v[e];
          }
          const Rows4<float> zv = Block4Apply(B32, row, rv);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            s2[1] += static_cast<double>(rv.v[e]) * zv.v[e];
          StoreRows4(r32 + row, rv);
          StoreRows4(pc + row, pv);
          StoreRows4(ap32 + row, zv);
        }
      } else {
        for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
          const float ri = static_cast<float>(r[i] / sigma);
          r32[i] = ri;
          pc[i] = static_cast<float>(q * static_cast<double>(pc[i]));
          s2[0] += static_cast<double>(ri) * ri;
          const float zi = ri * inv32[i];
          s2[1] += static_cast<double>(ri) * zi;
          if constexpr (fuse)
            ap32[i] = zi;
        }
      }
      R.Sum(s2);
      rho *= q * q;
      rr_ref = s2[0];
      rho_next = s2[1];
    }
    if (!(rho_next > 0.0)) {
      fallback = true;
      break;
    }
    beta = static_cast<float>(rho_next / rho);
    if constexpr (fuse) {
      rho = rho_next;
      continue;   // the next product forms p = z + beta p (z is published by the sums above)
    }
    if constexpr (block) {
      constexpr int V = Chunk<float>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c = R.S.row0 / V + threadIdx.x; c < c1; c += kBlockThreads) {
        const Chunk<float> zv = LoadChunk(ap32 + c * V);
        Chunk<float> pv = LoadChunk(pc + c * V);
#pragma unroll
        for (int e = 0; e < V; ++e)
          pv.v[e] = zv.v[e] + beta * pv.v[e];
        StoreChunk(pc + c * V, pv);
      }
    } else {
      constexpr int V = Chunk<float>::N;
      const int c1 = (R.S.row1 + V - 1) / V;
      for (int c0 = R.S.row0 / V + threadIdx.x; c0 < c1; c0 += kBlockThreads * kChunkUnroll) {
        Chunk<float> rv[kChunkUnroll], iv[kChunkUnroll], pv[kChunkUnroll];
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
            rv[u] = LoadChunk(r32 + c * V);
            iv[u] = LoadChunk(inv32 + c * V);
            pv[u] = LoadChunk(pc + c * V);
          }
        }
#pragma unroll
        for (int u = 0; u < kChunkUnroll; ++u) {
          const int c = c0 + u * kBlockThreads;
          if (c < c1) {
#pragma unroll
            for (int e = 0; e < V; ++e)
              pv[u].v[e] = rv[u].v[e] * iv[u].v[e] + beta * pv[u].v[e];
            StoreChunk(pc + c * V, pv[u]);
          }
        }
      }
    }
    rho = rho_next;
    R.Sync();
  }
  // Fold any pending correction, then finish in FP64 from x (or report the cap).
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
    x[i] += sigma * static_cast<double>(d32[i]);
  R.Sync();
  out.linear.iterations = it;
  // FP64 finish from the current x (a breakdown or stall of the FP32 iteration, or only the
  // final true-residual check).  With block-Jacobi the mixed form supports 4-row blocks.
  const LinearInfo li = Pcg64<true, kBlock ? 2 : 0>(P, vals, inv, B64, rhs, x, r, p, ap, tol,
                                                     fallback ? max(0, maxit - it) : fp64_maxit,
                                                     true, R, p_alt, q_buf, stop_on_stall);
  out.linear.iterations += li.iterations;
  out.fp64_iterations = li.iterations;
  out.linear.residual = li.residual;
  out.linear.status = li.status;
  if (li.status == FEMM_OK || fallback)
    return out;
  out.linear.status = FEMM_LINEAR_SOLVE_NOT_CONVERGED;
  return out;
}

// ----------------------------------------------------------------------------- Newton

struct ItemPointers {
  double *elem, *vals, *rhs, *inv, *x, *r, *p, *ap, *trial;
  double *p_alt, *q_buf;   // team solves: second search direction, A p
  float *vals32, *r32, *p32, *ap32, *d32, *inv32;
  float *p32_alt, *q32_buf;
  double *bjac;      // [kMaxBlock][n] block-Jacobi planes (nullptr: not allocated)
  float *bjac32;
};

__device__ __attribute__((always_inline)) ItemPointers ItemWork(const DevicePlan &P, const DeviceWork &W, int item, bool mixed)
{
  ItemPointers ip;
  const size_t n = VectorStride(P.n), s = P.sell_size;
  ip.elem = W.elem + static_cast<size_t>(item) * P.T * kElemStride;
  ip.vals = W.vals + item * s;
  ip.rhs = W.rhs + item * n;
  ip.inv = W.inv + item * n;
  ip.x = W.x + item * n;
  ip.r = W.r + item * n;
  ip.p = W.p + item * n;
  ip.ap = W.ap + item * n;
  ip.p_alt = W.p_alt + item * n;
  ip.q_buf = W.q_buf + item * n;
  ip.trial = W.trial + static_cast<size_t>(item) * P.N;
  ip.vals32 = mixed ? W.vals32 + item * s : nullptr;
  ip.r32 = mixed ? W.r32 + item * n : nullptr;
  ip.p32 = mixed ? W.p32 + item * n : nullptr;
  ip.ap32 = mixed ? W.ap32 + item * n : nullptr;
  ip.d32 = mixed ? W.d32 + item * n : nullptr;
  ip.inv32 = mixed ? W.inv32 + item * n : nullptr;
  ip.p32_alt = mixed ? W.p32_alt + item * n : nullptr;
  ip.q32_buf = mixed ? W.q32_buf + item * n : nullptr;
  ip.bjac = W.bjac ? W.bjac + item * kMaxBlock * n : nullptr;
  ip.bjac32 = (mixed && W.bjac32) ? W.bjac32 + item * kMaxBlock * n : nullptr;
  return ip;
}

// Full device assembly for one item; returns false if any element or diagonal is invalid.
template <class Team>
__device__ __attribute__((always_inline)) bool AssembleItem(const DevicePlan &P, const ItemPointers &w, const double *a,
                             const double *cur, bool newton, Team &R)
{
  const bool e_ok = ElementPass(P, R.S, a, w.elem, newton);
  R.Sync();
  GatherValues(P, R.S, w.elem, w.vals, w.vals32);
  __syncthreads();   // GatherRhs reads the diagonal of this workgroup's own rows
  const bool r_ok = GatherRhs(P, R.S, w.elem, w.vals, cur, w.rhs, w.inv, w.inv32);
  return !R.AnyTrue(!(e_ok && r_ok));
}

// ||K(A) A - b(A)|| over the free rows, with A taken at each free root
// (solvers/nonlinear_solver.cu:549-558).  Requires a fresh assembly at `a`.
// `xf` receives A at the free roots (n values).
template <class Team>
__device__ __attribute__((always_inline)) double EquationNorm(const DevicePlan &P, const ItemPointers &w, const double *a,
                               Team &R, double *xf)
{
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
    xf[i] = a[P.free_nodes[i]];
  R.Sync();
  Spmv<double, double, false>(P, R.S, w.vals, xf, w.ap, static_cast<const double *>(nullptr));
  R.Sync();
  double s[1] = {0.0};
  for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
    const double v = w.ap[i] - w.rhs[i];
    s[0] += v * v;
  }
  R.Sum(s);
  return sqrt(s[0]);
}

// Points the item's plan at its material variant; false for an index outside the set.
__device__ __attribute__((always_inline)) bool ApplyVariant(DevicePlan &P, const DeviceVariants &V, const int32_t *variant_of_item,
                             int item)
{
  if (variant_of_item == nullptr)
    return true;
  const int v = variant_of_item[item];
  if (v < 0 || v >= V.count)
    return false;
  const size_t T = P.T, M = V.mats, K = V.knots, n = P.n;
  P.tri_mat = V.tri_mat + v * T;
  P.tri_circuit = V.tri_circuit + v * T;
  P.tri_flux = V.tri_flux + v * 3 * T;
  P.mat_nu0 = V.mat_nu0 + v * M;
  P.mat_alpha = V.mat_alpha + v * M;
  P.mat_cold = V.mat_cold + v * M;
  P.mat_curve_off = V.mat_curve_off + v * M;
  P.mat_curve_cnt = V.mat_curve_cnt + v * M;
  P.curve_b = V.curve_b + v * K;
  P.curve_h = V.curve_h + v * K;
  P.curve_s = V.curve_s + v * K;
  P.pm_rhs = V.pm_rhs + v * n;
  P.circuit_rhs = V.circuit_rhs + v * static_cast<size_t>(max(P.C, 1)) * n;
  P.any_nonlinear = V.any_nonlinear[v];
  return true;
}

// Item `item`'s row of a [items][N] warm-start array (null: the plain start).
__device__ inline const double *WarmItem(const DevicePlan &P, const double *warm, int item)
{
  return warm ? warm + static_cast<size_t>(item) * P.N : nullptr;
}

// True when a launch with prior results leaves this item as the earlier attempt left it.
__device__ __attribute__((always_inline)) bool KeepPrior(const femm_item_result_t &earlier,
                                                         int32_t mode)
{
  if (mode == kPriorRedoFailed)
    return earlier.status == FEMM_OK;
  if (mode == kPriorRedoUnconverged)
    return earlier.status != FEMM_NONLINEAR_SOLVE_NOT_CONVERGED;
  return false;
}

// Linear tolerance of a safeguarded Newton step: the options' exact tolerance, or with
// cfg.inexact_forcing = c > 0 an inexact-Newton forcing term c x the relative equation
// residual (Eisenstat & Walker, SIAM J. Sci. Comput. 17, 1996) within [exact, 1e-3].
__device__ __attribute__((always_inline)) double ForcingTolerance(const femm_newton_options_t &o,
                                                                   const femm_solve_config_t &cfg,
                                                                   double rel_residual)
{
  const double exact = o.linear_relative_tolerance;
  if (!(cfg.inexact_forcing > 0.0))
    return exact;
  return fmax(exact, fmin(1e-3, cfg.inexact_forcing * rel_residual));
}

// Safeguarded acceptance: the linearization that meets both gates is solved to
// cfg.acceptance_linear_tolerance when that is tighter than the options' exact tolerance,
// and the item is then returned at that candidate.  An item whose tight re-solve did not
// converge (`relaxed`: the FP64 residual floor of a large model can lie above it) falls
// back to the options' exact tolerance, as without the setting.
__device__ __attribute__((always_inline)) bool TightAcceptance(const femm_newton_options_t &o,
                                                                const femm_solve_config_t &cfg,
                                                                bool relaxed = false)
{
  return !relaxed && cfg.acceptance_linear_tolerance > 0.0 &&
         cfg.acceptance_linear_tolerance < o.linear_relative_tolerance;
}

__device__ __attribute__((always_inline)) double AcceptTolerance(const femm_newton_options_t &o,
                                                                  const femm_solve_config_t &cfg,
                                                                  bool relaxed = false)
{
  return TightAcceptance(o, cfg, relaxed) ? cfg.acceptance_linear_tolerance
                                          : o.linear_relative_tolerance;
}

// Inexact-Newton inner tolerance (physics/nonlinear_model.cpp:166-174).
__device__ double LinearTolerance(const femm_newton_options_t &o, bool adaptive, double prev)
{
  const double exact = o.linear_relative_tolerance;
  if (!adaptive || !isfinite(prev) || prev <= 100.0 * o.relative_tolerance || !(exact > 0.0) ||
      !isfinite(exact))
    return exact;
  return fmax(exact, 1e-8 * fmin(1.0, prev));
}

template <class Team>
__device__ __attribute__((always_inline)) void WriteFlux(const DevicePlan &P, const double *a, double *flux, Team &R,
                          bool lead = true)
{
  for (int c = 0; c < P.C; ++c) {
    double s[1] = {0.0};
    for (int t = R.S.tri0 + threadIdx.x; t < R.S.tri1; t += kBlockThreads) {
      if (P.tri_circuit[t] != c)
        continue;
      const double *fw = P.tri_flux + 3 * static_cast<size_t>(t);
      s[0] += fw[0] * a[P.tri_nodes[3 * t]] + fw[1] * a[P.tri_nodes[3 * t + 1]] +
              fw[2] * a[P.tri_nodes[3 * t + 2]];
    }
    R.Sum(s);
    if (lead && threadIdx.x == 0)
      flux[c] = P.planar ? s[0] * P.depth : s[0];
  }
}

// Result of an item before its solve starts.
__device__ femm_item_result_t FreshResult()
{
  femm_item_result_t res{};
  res.status = FEMM_INVALID_ARGUMENT;
  res.relative_step = INFINITY;
  res.equation_residual = INFINITY;
  res.last_linear_tolerance = NAN;
  res.linear_residual = INFINITY;
  return res;
}

__device__ bool OptionsValid(const femm_newton_options_t &o)
{
  return o.relative_tolerance > 0.0 && isfinite(o.relative_tolerance) &&
         o.max_newton_iterations >= 0;
}

// One linear solve of the current assembly into w.x (x0 = w.x when warm).
// Solve forms (template argument of the solve kernels): FP64 or mixed precision, point or
// block-Jacobi preconditioner.  With `fp64` a mixed form solves in FP64 (the FP64 finish
// of ReliableSolve run from w.x for the whole solve): the safeguarded policy's acceptance
// re-solve, whose solution is the returned state.  FP32 iterations with FP64 residual
// replacement reach the same ||b - A x|| as FP64 PCG but leave more error in the soft
// (low-reluctivity) modes that this norm hardly sees; the FP64 re-solve from the mixed
// solution removes it.  With `stop_on_stall` (the tight acceptance re-solve) the solve also
// stops once its restarts no longer lower the true residual: a large model's FP64 floor can
// lie above the tight tolerance, and the policy then falls back to the exact tolerance.
constexpr int kFormMixed = 1, kFormBlock = 2;

template <int kForm, class Team>
__device__ __attribute__((always_inline)) LinearInfo LinearSolve(const DevicePlan &P, const ItemPointers &w,
                                  const femm_solve_config_t &cfg, double tol, int maxit,
                                  femm_item_result_t &res, Team &R, bool fp64 = false,
                                  bool stop_on_stall = false)
{
  constexpr bool kMixed = (kForm & kFormMixed) != 0, kBlock = (kForm & kFormBlock) != 0;
  LinearInfo li;
  // Block-Jacobi of the current operator (rebuilt before every solve, as the reference's
  // BlockJacobiPlan::Prepare, block_jacobi_preconditioner.cu:188-258).
  const int bs = kBlock ? cfg.block_size : 1;
  if constexpr (kBlock) {
    BuildBlockJacobi(P.n, P.slice_ptr, P.slice_width, P.sell_cols, P.diag_sell, R.S.row0,
                     R.S.row1, w.vals, w.inv, bs, w.bjac, kMixed ? w.bjac32 : nullptr);
    R.Sync();
  }
  const BlockPrec<double> B64{w.bjac, bs, VectorStride(P.n)};
  if constexpr (kMixed) {
    const BlockPrec<float> B32{w.bjac32, bs, VectorStride(P.n)};
    const RefinedInfo ri = ReliableSolve<kBlock>(P, w.vals, w.inv, w.vals32, w.inv32, B64, B32, w.rhs, w.x,
                                         w.r, w.p, w.ap, w.d32, w.r32, w.p32, w.ap32, tol,
                                         fp64 ? 0 : maxit, cfg.pcg_warm_start != 0,
                                         cfg.max_refinements, cfg.inner_relative_tolerance, R,
                                         fp64 ? maxit : 0, w.p32_alt, w.q32_buf, w.p_alt,
                                         w.q_buf, stop_on_stall);
    li = ri.linear;
    res.refinements += ri.refinements;
    res.fp64_iterations += ri.fp64_iterations;
  } else {
    li = Pcg64<true, kBlock ? 1 : 0>(P, w.vals, w.inv, B64, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit,
                             cfg.pcg_warm_start != 0, R, w.p_alt, w.q_buf, stop_on_stall);
    res.fp64_iterations += li.iterations;
  }
  R.Sync();
  res.linear_iterations += li.iterations;
  res.last_linear_tolerance = tol;
  res.linear_residual = li.residual;
  return li;
}

// Undamped Newton step ||A_cand - A|| / ||A_cand|| over all nodes, A_cand the free-row
// solution in w.x expanded with signs and the prescribed values (nonlinear_solver.cu:326-335).
template <class Team>
__device__ __attribute__((always_inline)) double StepOf(const DevicePlan &P,
                                                         const ItemPointers &w, const double *a,
                                                         Team &R)
{
  double s[2] = {0.0, 0.0};
  for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads) {
    const int f = P.node_to_free[v];
    const double cand = f >= 0 ? static_cast<double>(P.node_sign[v]) * w.x[f] : P.boundary[v];
    const double d = cand - a[v];
    s[0] += d * d;
    s[1] += cand * cand;
  }
  R.Sum(s);
  return sqrt(s[0]) / fmax(sqrt(s[1]), DBL_MIN);
}

// One item's nonlinear solve; leaves the state in `a`.
//
// safeguarded = false: the reference Newton driver (NonlinearP1Solver::Solve / SolveBatch,
// nonlinear_solver.cu:183-530, 667-769): plain start (cold secant at iteration 0) or
// `warm` (Newton terms from iteration 0), per-item inexact-Newton tolerance, fkn relaxation
// after iteration 5, NewtonStepConverged, and with require_equation_residual the
// equation-residual gate at the returned state.
//
// safeguarded = true: SolveSafeguarded (nonlinear_solver.cu:534-659).  Each iteration
// solves for the Newton candidate at the current state with the exact linear tolerance;
// the item is accepted when the candidate's undamped step is below 100x
// relative_tolerance and the equation residual at the current state is within
// relative_tolerance (the returned state is the current one).  Otherwise the first alpha
// in 1, 1/2, ..., 2^-19 whose trial state either reduces the equation residual by the
// Armijo factor (1 - 1e-4 alpha) or already passes the residual gate becomes the new
// state; if none does, the item stops unconverged.  From `warm` (the reference's warm-start
// hook, nonlinear_solver.cu:667-698) the iteration starts there.
//
// Both policies scale the equation residual by its norm at the Dirichlet-expanded plain
// start (nonlinear_solver.cu:565-576), independent of any warm start.  The two policies
// share one body, the assembly and the linear solve.
template <int kForm, class Team>
__device__ __attribute__((always_inline)) femm_item_result_t ItemSolve(const DevicePlan &P, const ItemPointers &w,
                                        const double *cur, double *a, bool inputs_ok,
                                        const femm_newton_options_t &o,
                                        const femm_solve_config_t &cfg, const double *warm,
                                        bool safeguarded, Team &R)
{
  femm_item_result_t res = FreshResult();
  inputs_ok = inputs_ok && OptionsValid(o);
  const bool has_warm = warm != nullptr;
  const bool adaptive = o.adaptive_linear_tolerance && P.any_nonlinear;
  const bool gate_eq = o.require_equation_residual != 0;

  // Plain start: prescribed values on every fixed node (signed roots expanded), zero
  // elsewhere; the residual scale is taken here, then the warm start replaces the free
  // nodes.
  for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads)
    a[v] = P.node_to_free[v] < 0 ? P.boundary[v] : 0.0;
  R.Sync();
  int status = inputs_ok ? FEMM_NONLINEAR_SOLVE_NOT_CONVERGED : FEMM_INVALID_ARGUMENT;
  double scale = 1.0, norm = INFINITY;
  if (inputs_ok && (safeguarded || cfg.verify_equation_residual || gate_eq)) {
    if (AssembleItem(P, w, a, cur, true, R)) {
      R.Sync();
      norm = EquationNorm(P, w, a, R, w.x);
      scale = fmax(norm, DBL_MIN);
    } else if (safeguarded) {
      status = FEMM_ASSEMBLY_FAILED;
    }
  }
  if (has_warm) {
    for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads)
      if (P.node_to_free[v] >= 0)
        a[v] = warm[v];
    R.Sync();
    if (safeguarded && status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED) {
      if (AssembleItem(P, w, a, cur, true, R)) {
        R.Sync();
        norm = EquationNorm(P, w, a, R, w.x);
      } else {
        status = FEMM_ASSEMBLY_FAILED;
      }
    }
  }

  double omega = 1.0, prev = INFINITY;
  bool pending = false;   // step gate passed; equation residual not yet checked
  bool exact_next = false; // safeguarded: re-solve the current linearization exactly
  bool relaxed = false;    // safeguarded: the tight acceptance re-solve did not converge
  bool eq_current = false;
  for (int it = 0; status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED; ++it) {
    if (safeguarded) {
      // The assembly at the current state is current (start or accepted trial).  A pending
      // re-solve finishes the step that met the gates, whatever the budget.
      if (it >= o.max_newton_iterations && !exact_next)
        break;
      if (!isfinite(norm)) {
        status = FEMM_NUMERICAL_NONFINITE;
        break;
      }
    } else {
      if (!pending && it >= o.max_newton_iterations)
        break;
      if (!AssembleItem(P, w, a, cur, has_warm || it > 0, R)) {
        status = FEMM_ASSEMBLY_FAILED;
        break;
      }
      R.Sync();
      if (pending) {
        // The Newton assembly at the returned state also yields its equation
        // residual (nonlinear_solver.cu:549-558); accept only if it passes too.
        res.equation_residual = EquationNorm(P, w, a, R, w.x) / scale;
        eq_current = true;
        pending = false;
        if (res.equation_residual <= o.relative_tolerance) {
          status = FEMM_OK;
          break;
        }
        if (it >= o.max_newton_iterations)
          break;
      }
    }
    const double tol = !safeguarded ? LinearTolerance(o, adaptive, prev)
                       : exact_next ? AcceptTolerance(o, cfg, relaxed)
                                    : ForcingTolerance(o, cfg, norm / scale);
    // A re-solve of the same linearization starts from its last solution.
    const bool resolve = exact_next;
    exact_next = false;
    if (cfg.pcg_warm_start && !resolve) {
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
        w.x[i] = a[P.free_nodes[i]];
      R.Sync();
    }
    // The tight acceptance re-solve gives up on a stall (then the exact tolerance is solved
    // as the reference would, restarting to the cap).
    const LinearInfo li = LinearSolve<kForm>(P, w, cfg, tol, o.max_linear_iterations, res, R,
                                             resolve, resolve && TightAcceptance(o, cfg, relaxed));
    if (li.status != FEMM_OK) {
      if (resolve && li.status == FEMM_LINEAR_SOLVE_NOT_CONVERGED &&
          TightAcceptance(o, cfg, relaxed)) {
        relaxed = true;      // solve this linearization again to the exact tolerance
        exact_next = true;
        continue;
      }
      status = li.status;
      break;
    }
    // Undamped step ||A_cand - A|| / ||A_cand|| over all nodes (nonlinear_solver.cu:326-335).
    double step = StepOf(P, w, a, R);
    res.newton_iterations = it + 1;
    res.relative_step = step;
    if (!isfinite(step)) {
      status = FEMM_NUMERICAL_NONFINITE;
      break;
    }
    if (safeguarded) {
      const bool gates =
          step < 100.0 * o.relative_tolerance && norm / scale <= o.relative_tolerance;
      if (gates && tol > AcceptTolerance(o, cfg, relaxed)) {
        // Both gates met on an inexact solve: solve this linearization again to the
        // acceptance tolerance (the assembly is still the current state's) before accepting.
        exact_next = true;
        continue;
      }
      if (gates && !(cfg.inexact_forcing > 0.0) && !TightAcceptance(o, cfg, relaxed)) {
        status = FEMM_OK;
        break;
      }
      // Armijo backtracking.  With inexact forcing or a tight acceptance tolerance, a step
      // that met both gates is tried at alpha = 1 only: the item is accepted at that
      // candidate if its own equation residual passes, else at the current state.
      bool accepted = false;
      double alpha = 1.0;
      for (int bt = 0; bt < 20; ++bt, alpha *= 0.5) {
        for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads) {
          const int f = P.node_to_free[v];
          const double cand =
              f >= 0 ? static_cast<double>(P.node_sign[v]) * w.x[f] : P.boundary[v];
          w.trial[v] = a[v] + alpha * (cand - a[v]);
        }
        R.Sync();
        ++res.trial_assemblies;
        if (!AssembleItem(P, w, w.trial, cur, true, R)) {
          if (!gates)
            status = FEMM_ASSEMBLY_FAILED;
          break;
        }
        R.Sync();
        const double tn = EquationNorm(P, w, w.trial, R, w.p);
        const bool take = gates ? isfinite(tn) && tn / scale <= o.relative_tolerance
                                : isfinite(tn) && (tn <= (1.0 - 1e-4 * alpha) * norm ||
                                                   tn / scale <= o.relative_tolerance);
        if (take) {
          for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads)
            a[v] = w.trial[v];
          R.Sync();
          norm = tn;
          accepted = true;
        }
        if (take || gates)
          break;
      }
      if (gates) {
        status = FEMM_OK;
        break;
      }
      if (!accepted)
        break;
      continue;
    }
    // fkn-style relaxation after iteration 5 (nonlinear_solver.cu:344-351).
    if (it > 5) {
      if (step > prev && omega > 0.125)
        omega *= 0.5;
      else
        omega += 0.1 * (1.0 - omega);
    }
    for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads) {
      const int f = P.node_to_free[v];
      const double cand = f >= 0 ? static_cast<double>(P.node_sign[v]) * w.x[f] : P.boundary[v];
      a[v] += omega * (cand - a[v]);
    }
    R.Sync();
    eq_current = false;
    // NewtonStepConverged (nonlinear_model.cpp:186-193).
    if (it > 0 && step < 100.0 * o.relative_tolerance && tol <= o.linear_relative_tolerance) {
      if (!gate_eq) {
        status = FEMM_OK;
        break;
      }
      pending = true;
    }
    prev = step;
  }
  if (safeguarded) {
    if (inputs_ok)
      res.equation_residual = norm / scale;
  } else if (inputs_ok && cfg.verify_equation_residual && !eq_current &&
             (status == FEMM_OK || status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED)) {
    if (AssembleItem(P, w, a, cur, true, R)) {
      R.Sync();
      res.equation_residual = EquationNorm(P, w, a, R, w.x) / scale;
    } else {
      res.equation_residual = INFINITY;
    }
  }
  res.status = status;
  return res;
}

enum SolveMode : int32_t {
  kModeNewton = 0,       // the reference driver
  kModeSafeguarded = 1,  // SolveSafeguarded
  kModeLearned = 2       // launch pair: safeguarded from `warm` within a budget, then the
                         // driver from the plain start for the items that attempt left
};

// The nonlinear solve of `kMode` for one item by `R`'s team, results of workgroup 0
// written out.  With prior results, an item the earlier launch kept (KeepPrior) is left
// untouched and the others add their counters to the earlier ones.
template <int kForm, int kMode, class Team>
__device__ __attribute__((always_inline)) void SolveItem(
    const DevicePlan &P, const DeviceWork &W, int item, bool lead, bool inputs_ok,
    const double *currents, const NewtonArgs &args, const double *warm_item, double *A,
    double *flux_out, femm_item_result_t *results, const femm_item_result_t *prior, Team &R)
{
  femm_item_result_t earlier{};
  if (prior != nullptr)
    earlier = prior[item];
  const ItemPointers w = ItemWork(P, W, item, (kForm & kFormMixed) != 0);
  double *a = A + static_cast<size_t>(item) * P.N;
  const double *cur = currents + static_cast<size_t>(item) * P.C;
  for (int c = 0; c < P.C; ++c)
    inputs_ok = inputs_ok && isfinite(cur[c]);
  static_assert(kMode == kModeNewton || kMode == kModeSafeguarded,
                "the learned mode is a pair of launches, not a kernel");
  femm_item_result_t res = ItemSolve<kForm>(P, w, cur, a, inputs_ok, args.options, args.config,
                                            warm_item, kMode == kModeSafeguarded, R);
  WriteFlux(P, a, flux_out + static_cast<size_t>(item) * P.C, R, lead);
  if (prior != nullptr) {
    res.newton_iterations += earlier.newton_iterations;
    res.linear_iterations += earlier.linear_iterations;
    res.refinements += earlier.refinements;
    res.fp64_iterations += earlier.fp64_iterations;
    res.trial_assemblies += earlier.trial_assemblies;
  }
  if (lead && threadIdx.x == 0)
    results[item] = res;
}

// One workgroup per batch item.
template <int kForm, int kMode>
__global__ __launch_bounds__(kBlockThreads) void SolveKernel(
    DevicePlan base, DeviceWork W, const double *currents, NewtonArgs args, const double *warm,
    double *A, double *flux_out, femm_item_result_t *results, const femm_item_result_t *prior,
    DeviceVariants VS, const int32_t *variant_of_item)
{
  __shared__ double lds[2 * kWaves * 4];
  const int item = blockIdx.x;
  if (prior != nullptr && KeepPrior(prior[item], args.prior_mode))
    return;
  DevicePlan P = base;
  const bool inputs_ok = ApplyVariant(P, VS, variant_of_item, item);
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  SolveItem<kForm, kMode>(P, W, item, true, inputs_ok, currents, args, WarmItem(P, warm, item), A,
                           flux_out, results, prior, R);
}

// `partitions` workgroups per batch item (grid = items x partitions, cooperative launch).
// An aborted team (a barrier wait that ran out) reports INTERNAL_ERROR for its item.
template <int kForm, int kMode>
__global__ __launch_bounds__(kBlockThreads) void TeamSolveKernel(
    DevicePlan base, DeviceWork W, const double *currents, NewtonArgs args, const double *warm,
    double *A, double *flux_out, femm_item_result_t *results, const femm_item_result_t *prior,
    DeviceVariants VS, const int32_t *variant_of_item, int partitions)
{
  __shared__ double lds[2 * kWaves * 4];
  const int item = UniformInt(static_cast<int>(blockIdx.x) / partitions);
  const int g = UniformInt(static_cast<int>(blockIdx.x) % partitions);
  if (prior != nullptr && KeepPrior(prior[item], args.prior_mode))
    return;
  DevicePlan P = base;
  const bool inputs_ok = ApplyVariant(P, VS, variant_of_item, item);
  GridTeam R{BlockReducer{lds, 0},
             Partition(P, partitions, g),
             W.team_partials + static_cast<size_t>(item) * 2 * kMaxPartitions * 4,
             W.team_barrier + static_cast<size_t>(item) * kBarrierWords,
             W.team_abort + item,
             partitions,
             g,
             0};
  SolveItem<kForm, kMode>(P, W, item, g == 0, inputs_ok, currents, args, WarmItem(P, warm, item),
                           A, flux_out, results, prior, R);
  R.Sync();
  if (g == 0 && threadIdx.x == 0 &&
      __hip_atomic_load(R.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    results[item].status = FEMM_INTERNAL_ERROR;
}

// Lists, in increasing order, the items a later team launch takes: with cap > 0 those a
// safeguarded launch capped at `cap` Newton iterations left unconverged at the cap, with
// cap = 0 those whose status is not FEMM_OK.  cont[0] = count, cont[1 + k] = k-th item.
// One wave; a ballot per 64 items keeps the order.
__global__ __launch_bounds__(64) void ListItemsKernel(const femm_item_result_t *results,
                                                      int items, int cap, int32_t *cont)
{
  int count = 0;
  for (int base = 0; base < items; base += 64) {
    const int i = base + static_cast<int>(threadIdx.x);
    bool go = false;
    if (i < items)
      go = cap > 0 ? results[i].status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED &&
                         results[i].newton_iterations >= cap
                   : results[i].status != FEMM_OK;
    const unsigned long long mask = __ballot(go);
    if (go)
      cont[1 + count + __popcll(mask & ((1ull << threadIdx.x) - 1ull))] = i;
    count += __popcll(mask);
  }
  if (threadIdx.x == 0)
    cont[0] = count;
}

// Workgroups per listed item when `grid` workgroups share `count` items: as many as divide
// evenly, at most kMaxPartitions and at most one per kTeamRowSlices row slices (a team's
// barriers cost more than they save below about 3.5k rows per workgroup).
constexpr int kTeamRowSlices = 56;
__host__ __device__ inline int ListTeams(int grid, int count, int nslices)
{
  int p = grid / (count > 0 ? count : 1);
  p = p < kMaxPartitions ? p : kMaxPartitions;
  const int by_rows = nslices / kTeamRowSlices;
  p = p < by_rows ? p : by_rows;
  return p > 1 ? p : 1;
}

// Solves the items ListItemsKernel listed, the launch's workgroups shared out evenly
// (ListTeams per item; with more items than teams, each team takes items t, t + teams, ...),
// counters added to the earlier launch's.  kModeSafeguarded with args.resume continues each
// item from the state the capped launch left in A with the options' budget (the rest of the
// capped attempt's): the state is copied to the item's trial buffer first, because the solve
// rewrites A with the plain start (for the residual scale) before it takes the warm start,
// and it uses the trial buffer only after that.  Without args.resume, and in kModeNewton,
// each item is solved again from the plain start.
template <int kForm, int kMode>
__global__ __launch_bounds__(kBlockThreads) void TeamListKernel(
    DevicePlan base, DeviceWork W, const double *currents, NewtonArgs args, double *A,
    double *flux_out, femm_item_result_t *results, DeviceVariants VS,
    const int32_t *variant_of_item)
{
  __shared__ double lds[2 * kWaves * 4];
  const int count = UniformInt(W.cont[0]);
  const int p = UniformInt(ListTeams(static_cast<int>(gridDim.x), count, base.nslices));
  const int teams = static_cast<int>(gridDim.x) / p;
  const int g = UniformInt(static_cast<int>(blockIdx.x) % p);
  const int team = UniformInt(static_cast<int>(blockIdx.x) / p);
  if (team >= teams)
    return;   // workgroups past the last whole team
  for (int t = team; t < count; t += teams) {
    const int item = UniformInt(W.cont[1 + t]);
    DevicePlan P = base;
    const bool inputs_ok = ApplyVariant(P, VS, variant_of_item, item);
    GridTeam R{BlockReducer{lds, 0},
               Partition(P, p, g),
               W.team_partials + static_cast<size_t>(item) * 2 * kMaxPartitions * 4,
               W.team_barrier + static_cast<size_t>(item) * kBarrierWords,
               W.team_abort + item,
               p,
               g,
               0};
    const double *start = nullptr;
    if (kMode == kModeSafeguarded && args.resume) {
      const double *a = A + static_cast<size_t>(item) * P.N;
      double *keep = W.trial + static_cast<size_t>(item) * P.N;
      for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads)
        keep[v] = a[v];
      start = keep;
    }
    SolveItem<kForm, kMode>(P, W, item, g == 0, inputs_ok, currents, args, start, A, flux_out,
                            results, results, R);
    R.Sync();
    if (g == 0 && threadIdx.x == 0 &&
        __hip_atomic_load(R.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
      results[item].status = FEMM_INTERNAL_ERROR;
  }
}

// Several plans (meshes of uneven size) in one cooperative launch: entry e covers
// blocks [first_block, first_block + items * partitions) and solves its items with its
// own plan and workspace, `partitions` workgroups per item (one workgroup: no team
// barrier or partial exchange at all).  With `prior`, each entry's results are the
// earlier launch's (read and rewritten in place, as KeepPrior decides).
template <int kForm, int kMode>
__global__ __launch_bounds__(kBlockThreads) void GroupSolveKernel(const GroupEntry *entries,
                                                                  int count, NewtonArgs args,
                                                                  int prior)
{
  __shared__ double lds[2 * kWaves * 4];
  int e = 0;
  while (e + 1 < count && static_cast<int>(blockIdx.x) >= entries[e + 1].first_block)
    ++e;
  e = UniformInt(e);
  const GroupEntry &E = entries[e];
  const int local = static_cast<int>(blockIdx.x) - E.first_block;
  const int G = UniformInt(E.partitions);
  const int item = UniformInt(local / G), g = UniformInt(local % G);
  const femm_item_result_t *earlier = prior ? E.results : nullptr;
  if (earlier != nullptr && KeepPrior(earlier[item], args.prior_mode))
    return;
  const DevicePlan P = E.plan;
  const DeviceWork W = E.work;
  // One team type for every entry (a one-workgroup team reduces and synchronises as a
  // block): one inlined copy of the solve keeps the kernel's registers in bounds.
  GridTeam R{BlockReducer{lds, 0},
             Partition(P, G, g),
             W.team_partials + static_cast<size_t>(item) * 2 * kMaxPartitions * 4,
             W.team_barrier + static_cast<size_t>(item) * kBarrierWords,
             W.team_abort + item,
             G,
             g,
             0};
  SolveItem<kForm, kMode>(P, W, item, g == 0, true, E.currents, args, nullptr, E.a, E.flux,
                           E.results, earlier, R);
  R.Sync();
  if (g == 0 && threadIdx.x == 0 &&
      __hip_atomic_load(R.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    E.results[item].status = FEMM_INTERNAL_ERROR;
}

// Fresh team barrier words (kBarrierWords per item) and abort flags before a partitioned
// launch; a kernel rather than a memset so the pair can sit in a captured graph.
__global__ void ClearTeamKernel(uint32_t *barrier, int32_t *abort, int items)
{
  const long words = static_cast<long>(items) * kBarrierWords;
  for (long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; k < words;
       k += static_cast<long>(gridDim.x) * blockDim.x) {
    barrier[k] = 0u;
    if (k % kBarrierWords == 0)
      abort[k / kBarrierWords] = 0;
  }
}

// The same for every partitioned entry of a group (block e clears entry e).
__global__ void ClearGroupKernel(const GroupEntry *entries, int count)
{
  const GroupEntry &E = entries[blockIdx.x];
  if (E.partitions > 1) {
    const long words = static_cast<long>(E.items) * kBarrierWords;
    for (long k = threadIdx.x; k < words; k += blockDim.x)
      E.work.team_barrier[k] = 0u;
    for (int i = threadIdx.x; i < E.items; i += blockDim.x)
      E.work.team_abort[i] = 0;
  }
}

// ----------------------------------------------------------------------------- launches

// Largest grid a cooperative launch of `kernel` may use on the current device.
int CoResidentBlocks(const void *kernel)
{
  int device = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlockThreads, 0) !=
          hipSuccess)
    return 0;
  return cus * per_cu;
}

template <int kForm, int kMode>
int LaunchTeams(const DeviceBuffers &dev, int32_t items, int32_t partitions,
                const double *currents, const NewtonArgs &args, const double *warm, double *a,
                double *flux, femm_item_result_t *results, hipStream_t stream,
                const LaunchExtras &x)
{
  const void *kernel = reinterpret_cast<const void *>(&TeamSolveKernel<kForm, kMode>);
  const long grid = static_cast<long>(items) * partitions;
  if (partitions > kMaxPartitions || grid > CoResidentBlocks(kernel))
    return FEMM_INVALID_ARGUMENT;
  // Fresh barrier words and abort flags for the launched items.
  ClearTeamKernel<<<(items + 255) / 256, 256, 0, stream>>>(dev.work.team_barrier,
                                                           dev.work.team_abort, items);
  if (hipGetLastError() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  DevicePlan plan = dev.plan;
  DeviceWork work = dev.work;
  NewtonArgs launch_args = args;
  DeviceVariants variants = dev.variants;
  const femm_item_result_t *prior = x.prior;
  const int32_t *variant_of_item = x.variant_of_item;
  int parts = partitions;
  void *params[] = {&plan, &work, &currents, &launch_args, &warm, &a, &flux, &results,
                    &prior, &variants, &variant_of_item, &parts};
  return hipLaunchCooperativeKernel(kernel, dim3(static_cast<unsigned>(grid)), dim3(kBlockThreads),
                                    params, 0, stream) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
}

// Lists the items (ListItemsKernel with `cap`) and solves them in teams over every
// co-resident workgroup (TeamListKernel<kForm, kMode>, cooperative launch) with `args`.
template <int kForm, int kMode>
int LaunchList(const DeviceBuffers &dev, int32_t items, int32_t cap, const double *currents,
               const NewtonArgs &args, double *a, double *flux, femm_item_result_t *results,
               hipStream_t stream, const LaunchExtras &x)
{
  const void *kernel = reinterpret_cast<const void *>(&TeamListKernel<kForm, kMode>);
  const long grid = std::min<long>(CoResidentBlocks(kernel),
                                   static_cast<long>(items) * kMaxPartitions);
  if (grid <= 0)
    return FEMM_INTERNAL_ERROR;
  ListItemsKernel<<<1, 64, 0, stream>>>(results, items, cap, dev.work.cont);
  ClearTeamKernel<<<(items + 255) / 256, 256, 0, stream>>>(dev.work.team_barrier,
                                                           dev.work.team_abort, items);
  if (hipGetLastError() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  DevicePlan plan = dev.plan;
  DeviceWork work = dev.work;
  NewtonArgs launch_args = args;
  DeviceVariants variants = dev.variants;
  const int32_t *variant_of_item = x.variant_of_item;
  void *params[] = {&plan, &work, &currents, &launch_args, &a, &flux, &results, &variants,
                    &variant_of_item};
  return hipLaunchCooperativeKernel(kernel, dim3(static_cast<unsigned>(grid)), dim3(kBlockThreads),
                                    params, 0, stream) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
}

// Runs f(std::integral_constant<int, form>{}) for a runtime solve form (kFormMixed |
// kFormBlock bits).
template <class F>
int WithForm(int form, F &&f)
{
  switch (form & 3) {
  case 0:
    return f(std::integral_constant<int, 0>{});
  case 1:
    return f(std::integral_constant<int, 1>{});
  case 2:
    return f(std::integral_constant<int, 2>{});
  default:
    return f(std::integral_constant<int, 3>{});
  }
}


// One solve form's launch: `partitions` > 1 runs the team kernel cooperatively, otherwise
// one workgroup per item.  The learned mode is launches of the other modes: the
// safeguarded attempt from `warm` with the warm options over every item, then the same
// safeguarded policy from the plain start over the items that attempt did not accept,
// their counters added to the attempt's.  (One kernel holding both attempts per item needs more registers than a
// 512-thread workgroup has per lane and spills inside the PCG loop.)  With
// one workgroup per item, the plain-start redo runs the rejected items only, in teams that
// share the whole GPU (TeamListKernel), so a few fallbacks do not run on a few compute units
// while the rest idle.  With config.split_newton_iterations = k (below the warm budget),
// the attempt itself is two launches: at most k iterations per item, then TeamListKernel
// resumes the items still iterating with the rest of the budget.  Which items a team
// launch takes, and their team size, depend only on the items' own iterations, so repeated
// runs stay bitwise identical.
template <int kForm, int kMode>
int LaunchFormSolve(const DeviceBuffers &dev, int32_t items, int32_t partitions,
                    const double *currents, const NewtonArgs &args, const double *warm, double *a,
                    double *flux, femm_item_result_t *results, hipStream_t stream,
                    const LaunchExtras &x)
{
  if constexpr (kMode == kModeLearned) {
    NewtonArgs warm_args = args;
    warm_args.options = args.warm_options;
    const int budget = args.warm_options.max_newton_iterations;
    const int split = args.config.split_newton_iterations;
    const bool resume = split > 0 && split < budget && partitions <= 1;
    if (resume)
      warm_args.options.max_newton_iterations = split;
    int st = LaunchFormSolve<kForm, kModeSafeguarded>(dev, items, partitions, currents,
                                                      warm_args, warm, a, flux, results, stream,
                                                      x);
    if (st != FEMM_OK)
      return st;
    if (resume) {
      NewtonArgs rest = warm_args;
      rest.options.max_newton_iterations = budget - split;
      rest.resume = 1;
      st = LaunchList<kForm, kModeSafeguarded>(dev, items, split, currents, rest, a, flux,
                                               results, stream, x);
      if (st != FEMM_OK)
        return st;
    }
    // The items the attempt did not accept: the same safeguarded policy (the reference's
    // SolveSafeguarded, its own recovery) from the plain start with the full budget.
    NewtonArgs plain_args = args;
    plain_args.resume = 0;
    if (partitions <= 1)
      return LaunchList<kForm, kModeSafeguarded>(dev, items, 0, currents, plain_args, a, flux,
                                                 results, stream, x);
    LaunchExtras redo = x;
    redo.prior = results;
    plain_args.prior_mode = kPriorRedoFailed;
    return LaunchFormSolve<kForm, kModeSafeguarded>(dev, items, partitions, currents, plain_args,
                                                    nullptr, a, flux, results, stream, redo);
  } else {
    if (partitions > 1)
      return LaunchTeams<kForm, kMode>(dev, items, partitions, currents, args, warm, a, flux,
                                       results, stream, x);
    if (partitions < 0)
      return FEMM_INVALID_ARGUMENT;
    hipLaunchKernelGGL((SolveKernel<kForm, kMode>), dim3(items), dim3(kBlockThreads), 0, stream,
                       dev.plan, dev.work, currents, args, warm, a, flux, results, x.prior,
                       dev.variants, x.variant_of_item);
    return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
  }
}

// Runs one of the three solve modes of form kForm.
template <int kForm>
int LaunchFormSolveMode(int mode, const DeviceBuffers &dev, int32_t items, int32_t partitions,
                        const double *currents, const NewtonArgs &args, const double *warm,
                        double *a, double *flux, femm_item_result_t *results, hipStream_t stream,
                        const LaunchExtras &x)
{
  switch (mode) {
  case kModeNewton:
    return LaunchFormSolve<kForm, kModeNewton>(dev, items, partitions, currents, args, warm, a,
                                               flux, results, stream, x);
  case kModeSafeguarded:
    return LaunchFormSolve<kForm, kModeSafeguarded>(dev, items, partitions, currents, args, warm,
                                                    a, flux, results, stream, x);
  default:
    return LaunchFormSolve<kForm, kModeLearned>(dev, items, partitions, currents, args, warm, a,
                                                flux, results, stream, x);
  }
}

// The group kernel of form kForm (reference driver or its safeguarded recovery).
template <int kForm>
const void *FormGroupKernel(bool recovery)
{
  return recovery ? reinterpret_cast<const void *>(&GroupSolveKernel<kForm, kModeSafeguarded>)
                  : reinterpret_cast<const void *>(&GroupSolveKernel<kForm, kModeNewton>);
}

}  // namespace
}  // namespace femm
