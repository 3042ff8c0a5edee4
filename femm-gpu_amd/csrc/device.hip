// HIP kernels of the MI355X nonlinear magnetostatic solve path (gfx950, wave64).
//
// Work decomposition: one 512-thread workgroup owns one batch item for the
// whole Newton solve.  Every reduction is a fixed-shape wave64 shuffle tree
// followed by a fixed-order sum of the 8 wave partials, so results are
// bitwise reproducible from run to run and independent of batch composition.
// There are no atomics and no inter-workgroup synchronisation; the hardware
// dispatcher load-balances items across the 256 CUs as workgroups retire.
//
// Memory layout (HBM, item-major): per item the nodal state A[N], the element
// Newton blocks elem[T][12], the operator values in SELL-64 order, and the PCG
// vectors x, r, p, Ap, D^-1 [n].  The SELL column indices, contributor lists and
// element geometry are shared by all items and stay resident in L2 / MALL.

#include "device.hpp"
#include "post.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
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
  __device__ void Sum(double (&v)[K])
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

  __device__ double Max(double v)
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

  __device__ int AnyTrue(bool local)
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
  __device__ void Sum(double (&v)[K])
  {
    R.Sum(v);
  }
  __device__ double Max(double v) { return R.Max(v); }
  __device__ int AnyTrue(bool local) { return R.AnyTrue(local); }
  __device__ void Sync() { __syncthreads(); }
};


// Several workgroups solving one item together (a partitioned solve).  Workgroup g of G
// owns a slice-aligned share of the free rows (with their CSR slots) and equal shares of
// the nodes and triangles.  A reduction turns each workgroup's block sum into a partial in
// global memory; after a team barrier every workgroup adds the G partials in workgroup
// order, so the result is identical in every workgroup and repeatable for a fixed G.  The
// barrier is sense-reversing on two words per item (arrivals, generation) with agent-scope
// release / acquire, so each workgroup's global writes are visible to the others after
// it.  All workgroups of a launch are co-resident (cooperative launch); a wait that still
// exceeds kMaxBarrierPolls (seconds) marks the team aborted, and every later barrier of
// the team returns at once, so no workgroup can wait forever.
constexpr long kMaxBarrierPolls = 1L << 27;

// Thread 0's part of the team barrier; one out-of-line copy keeps the spin loop from
// being expanded at every barrier of the solver.
__device__ __noinline__ void TeamArrive(uint32_t *bar, int32_t *abort, int G)
{
  __threadfence();   // the workgroup's writes (ordered before by the block barrier)
  if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    return;
  const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t arrived =
      __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (arrived == static_cast<uint32_t>(G - 1)) {
    __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    // Relaxed polls (no cache invalidation per poll); the fence below acquires once.
    for (long polls = 0;
         __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen; ++polls) {
      if (polls > kMaxBarrierPolls ||
          __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __threadfence();   // acquire: the other workgroups' writes, for the loads after the barrier
}

struct GridTeam {
  BlockReducer R;
  Span S;
  double *partials;   // [2][G][4]
  uint32_t *bar;      // [2]: arrivals, generation
  int32_t *abort;
  int G, g, buf;

  __device__ void Sync()
  {
    __syncthreads();
    if (threadIdx.x == 0)
      TeamArrive(bar, abort, G);
    __syncthreads();
  }

  template <int K>
  __device__ void Sum(double (&v)[K])
  {
    R.Sum(v);
    double *mine = partials + (static_cast<size_t>(buf) * G + g) * 4;
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        mine[k] = v[k];
    }
    Sync();
    const double *all = partials + static_cast<size_t>(buf) * G * 4;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int h = 0; h < G; ++h)
        s += all[h * 4 + k];
      v[k] = Uniform(s);
    }
    buf ^= 1;
  }

  __device__ double Max(double v)
  {
    v = R.Max(v);
    double *mine = partials + (static_cast<size_t>(buf) * G + g) * 4;
    if (threadIdx.x == 0)
      mine[0] = v;
    Sync();
    const double *all = partials + static_cast<size_t>(buf) * G * 4;
    double m = all[0];
    for (int h = 1; h < G; ++h)
      m = fmax(m, all[h * 4]);
    buf ^= 1;
    return Uniform(m);
  }

  __device__ int AnyTrue(bool local) { return Max(local ? 1.0 : 0.0) > 0.0; }
};

__device__ inline int UniformInt(int v) { return __builtin_amdgcn_readfirstlane(v); }

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
__device__ bool ElementPass(const DevicePlan &P, const Span &S, const double *a, double *elem,
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
__device__ void GatherValues(const DevicePlan &P, const Span &S, const double *elem, double *vals,
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
    const int pos = P.slot_to_sell[slot];
    vals[pos] = sum;
    if (vals32)
      vals32[pos] = static_cast<float>(sum);
  }
}

// Row-owned rhs (physics/nonlinear_assembly.cu:142-186): PM + circuit sources, then the
// ordered element terms.  Also the Jacobi inverse (:188-206).
__device__ bool GatherRhs(const DevicePlan &P, const Span &S, const double *elem,
                          const double *vals, const double *cur, double *rhs, double *inv,
                          float *inv32)
{
  bool ok = true;
  for (int row = S.row0 + threadIdx.x; row < S.row1; row += kBlockThreads) {
    double s = P.pm_rhs[row];
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

template <typename T>
struct SliceLoads {
  int c[kUnrolledWidth];
  T a[kUnrolledWidth];
  int width;
};

// Branch-free loads of one slice: entries past the slice width repeat the last column
// (a valid address in the same cache line) with a zero value, so the whole slice is one
// straight-line run of loads and later adds only +-0 terms.
template <typename T>
__device__ inline void LoadSlice(const DevicePlan &P, const T *__restrict__ vals, int s, int lane,
                                 SliceLoads<T> &L)
{
  const int base = P.slice_ptr[s] + lane;
  L.width = P.slice_width[s];
  const int last = L.width - 1;   // every slice holds a diagonal, so width >= 1
#pragma unroll
  for (int k = 0; k < kUnrolledWidth; ++k) {
    const int pos = base + max(min(k, last), 0) * kSliceRows;
    L.c[k] = P.sell_cols[pos];
    const T value = vals[pos];
    L.a[k] = k <= last ? value : T(0);
  }
}

// kPipeline = false: the plain two-round-trip loop, for rarely used products (FP64
// residuals inside the mixed form, equation residuals) where registers matter more.
template <typename T, typename V, bool kPipeline = true>
__device__ double Spmv(const DevicePlan &P, const Span &S, const T *__restrict__ vals,
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
      const int row = s * kSliceRows + lane;
      if (row < P.n) {
        y[row] = static_cast<V>(sum);
        if (w)
          local += static_cast<double>(w[row]) * static_cast<double>(sum);
      }
    }
    return local;
  }
  SliceLoads<T> cur;
  cur.width = 0;
  if (S.slice0 + wave < S.slice1)
    LoadSlice(P, vals, S.slice0 + wave, lane, cur);
  for (int s = S.slice0 + wave; s < S.slice1; s += kWaves) {
    T sum = 0;
    const int next = s + kWaves;
    if (cur.width <= kUnrolledWidth) {
      V g[kUnrolledWidth];
#pragma unroll
      for (int k = 0; k < kUnrolledWidth; ++k)
        g[k] = v[cur.c[k]];
      SliceLoads<T> nxt;
      // The last slice prefetches itself again (in bounds, never used).
      LoadSlice(P, vals, next < S.slice1 ? next : s, lane, nxt);
#pragma unroll
      for (int k = 0; k < kUnrolledWidth; ++k)
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
    const int row = s * kSliceRows + lane;
    if (row < P.n) {
      y[row] = static_cast<V>(sum);
      if (w)
        local += static_cast<double>(w[row]) * static_cast<double>(sum);
    }
  }
  return local;
}

// Vector passes keep kVecUnroll independent rows in flight per thread (all
// loads of a step are issued before any arithmetic).
constexpr int kVecUnroll = 4;

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
template <bool kPipeline = true, class Team = BlockTeam>
__device__ LinearInfo Pcg64(const DevicePlan &P, const double *vals, const double *inv,
                            const double *rhs, double *x, double *r, double *p, double *ap,
                            double tol, int maxit, bool warm, Team &R)
{
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
  for (int it = 0; it < maxit; ++it) {
    double pap[1] = {Spmv<double, double, kPipeline>(P, R.S, vals, p, ap, p)};
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
    for (int i0 = R.S.row0 + threadIdx.x; i0 < R.S.row1; i0 += kBlockThreads * kVecUnroll) {
      double xv[kVecUnroll], pv[kVecUnroll], rv[kVecUnroll], av[kVecUnroll], iv[kVecUnroll];
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          xv[u] = x[i];
          pv[u] = p[i];
          rv[u] = r[i];
          av[u] = ap[i];
          iv[u] = inv[i];
        }
      }
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          x[i] = xv[u] + alpha * pv[u];
          const double ri = rv[u] - alpha * av[u];
          r[i] = ri;
          s1[0] += ri * ri;
          s1[1] += ri * (ri * iv[u]);
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
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
        const double zi = r[i] * inv[i];
        p[i] = zi;
        rs[0] += r[i] * zi;
      }
      R.Sum(rs);
      if (!isfinite(rs[0]) || !(rs[0] > 0.0)) {
        info.status = FEMM_LINEAR_SOLVE_BREAKDOWN;
        return info;
      }
      rho = rs[0];
      continue;
    }
    if (!isfinite(s1[1]) || !(s1[1] > 0.0)) {
      info.status = FEMM_LINEAR_SOLVE_BREAKDOWN;
      return info;
    }
    const double beta = s1[1] / rho;
    for (int i0 = R.S.row0 + threadIdx.x; i0 < R.S.row1; i0 += kBlockThreads * kVecUnroll) {
      double rv[kVecUnroll], iv[kVecUnroll], pv[kVecUnroll];
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          rv[u] = r[i];
          iv[u] = inv[i];
          pv[u] = p[i];
        }
      }
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1)
          p[i] = rv[u] * iv[u] + beta * pv[u];
      }
    }
    rho = s1[1];
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
};

// r = b - A64 x in FP64; returns ||r||_2 and max|r| through the reducer.
template <class Team>
__device__ void TrueResidual(const DevicePlan &P, const double *vals, const double *rhs,
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
// from the current x.
template <class Team>
__device__ RefinedInfo ReliableSolve(const DevicePlan &P, const double *vals, const double *inv,
                                     const float *vals32, const float *inv32, const double *rhs,
                                     double *x, double *r, double *p, double *ap, float *d32,
                                     float *r32, float *p32, float *ap32, double tol, int maxit,
                                     bool warm, int max_updates, double delta, Team &R)
{
  RefinedInfo out{{FEMM_INTERNAL_ERROR, 0, INFINITY}, 0};
  const int n = P.n;
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
    for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
      const float ri = static_cast<float>(r[i] / sigma);
      const float zi = ri * inv32[i];
      r32[i] = ri;
      p32[i] = zi;
      d32[i] = 0.0f;
      s0[0] += static_cast<double>(ri) * ri;
      s0[1] += static_cast<double>(ri) * zi;
    }
    R.Sum(s0);
    rr = s0[0];
    rho = s0[1];
  }
  double rr_ref = rr;           // ||r32||^2 at the last replacement
  double last_true = rnorm;     // FP64 true residual at the last replacement
  bool fallback = false;
  int it = 0;
  while (it < maxit) {
    double pap[1] = {Spmv(P, R.S, vals32, p32, ap32, p32)};
    R.Sum(pap);
    if (!(pap[0] > 0.0) || !(rho > 0.0) || !isfinite(pap[0]) || !isfinite(rho)) {
      fallback = true;
      break;
    }
    const float alpha = static_cast<float>(rho / pap[0]);
    double s1[2] = {0.0, 0.0};
    for (int i0 = R.S.row0 + threadIdx.x; i0 < R.S.row1; i0 += kBlockThreads * kVecUnroll) {
      float dv[kVecUnroll], pv[kVecUnroll], rv[kVecUnroll], av[kVecUnroll], iv[kVecUnroll];
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          dv[u] = d32[i];
          pv[u] = p32[i];
          rv[u] = r32[i];
          av[u] = ap32[i];
          iv[u] = inv32[i];
        }
      }
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          d32[i] = dv[u] + alpha * pv[u];
          const float ri = rv[u] - alpha * av[u];
          r32[i] = ri;
          s1[0] += static_cast<double>(ri) * ri;
          s1[1] += static_cast<double>(ri) * (ri * iv[u]);
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
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads) {
        const float ri = static_cast<float>(r[i] / sigma);
        r32[i] = ri;
        p32[i] = static_cast<float>(q * static_cast<double>(p32[i]));
        s2[0] += static_cast<double>(ri) * ri;
        s2[1] += static_cast<double>(ri) * (ri * inv32[i]);
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
    const float beta = static_cast<float>(rho_next / rho);
    for (int i0 = R.S.row0 + threadIdx.x; i0 < R.S.row1; i0 += kBlockThreads * kVecUnroll) {
      float rv[kVecUnroll], iv[kVecUnroll], pv[kVecUnroll];
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1) {
          rv[u] = r32[i];
          iv[u] = inv32[i];
          pv[u] = p32[i];
        }
      }
#pragma unroll
      for (int u = 0; u < kVecUnroll; ++u) {
        const int i = i0 + u * kBlockThreads;
        if (i < R.S.row1)
          p32[i] = rv[u] * iv[u] + beta * pv[u];
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
  const LinearInfo li = Pcg64<false>(P, vals, inv, rhs, x, r, p, ap, tol,
                                     fallback ? max(0, maxit - it) : 0, true, R);
  out.linear.iterations += li.iterations;
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
  float *vals32, *r32, *p32, *ap32, *d32, *inv32;
};

__device__ ItemPointers ItemWork(const DevicePlan &P, const DeviceWork &W, int item, bool mixed)
{
  ItemPointers ip;
  const size_t n = P.n, s = P.sell_size;
  ip.elem = W.elem + static_cast<size_t>(item) * P.T * kElemStride;
  ip.vals = W.vals + item * s;
  ip.rhs = W.rhs + item * n;
  ip.inv = W.inv + item * n;
  ip.x = W.x + item * n;
  ip.r = W.r + item * n;
  ip.p = W.p + item * n;
  ip.ap = W.ap + item * n;
  ip.trial = W.trial + static_cast<size_t>(item) * P.N;
  ip.vals32 = mixed ? W.vals32 + item * s : nullptr;
  ip.r32 = mixed ? W.r32 + item * n : nullptr;
  ip.p32 = mixed ? W.p32 + item * n : nullptr;
  ip.ap32 = mixed ? W.ap32 + item * n : nullptr;
  ip.d32 = mixed ? W.d32 + item * n : nullptr;
  ip.inv32 = mixed ? W.inv32 + item * n : nullptr;
  return ip;
}

// Full device assembly for one item; returns false if any element or diagonal is invalid.
template <class Team>
__device__ bool AssembleItem(const DevicePlan &P, const ItemPointers &w, const double *a,
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
__device__ double EquationNorm(const DevicePlan &P, const ItemPointers &w, const double *a,
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

struct NewtonArgs {
  femm_newton_options_t options;
  femm_solve_config_t config;
  int32_t prior_mode;
  femm_newton_options_t warm_options;   // kModeLearned: the warm attempt's options
};

// Points the item's plan at its material variant; false for an index outside the set.
__device__ bool ApplyVariant(DevicePlan &P, const DeviceVariants &V, const int32_t *variant_of_item,
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

// True when a launch with prior results leaves this item as the earlier attempt left it.
__device__ bool KeepPrior(const femm_item_result_t &earlier, int32_t mode)
{
  if (mode == kPriorRedoFailed)
    return earlier.status == FEMM_OK;
  if (mode == kPriorRedoUnconverged)
    return earlier.status != FEMM_NONLINEAR_SOLVE_NOT_CONVERGED;
  return false;
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
__device__ void WriteFlux(const DevicePlan &P, const double *a, double *flux, Team &R,
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
template <bool kMixed, class Team>
__device__ LinearInfo LinearSolve(const DevicePlan &P, const ItemPointers &w,
                                  const femm_solve_config_t &cfg, double tol, int maxit,
                                  femm_item_result_t &res, Team &R)
{
  LinearInfo li;
  if constexpr (kMixed) {
    const RefinedInfo ri = ReliableSolve(P, w.vals, w.inv, w.vals32, w.inv32, w.rhs, w.x, w.r, w.p,
                                         w.ap, w.d32, w.r32, w.p32, w.ap32, tol, maxit,
                                         cfg.pcg_warm_start != 0, cfg.max_refinements,
                                         cfg.inner_relative_tolerance, R);
    li = ri.linear;
    res.refinements += ri.refinements;
  } else {
    li = Pcg64(P, w.vals, w.inv, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit, cfg.pcg_warm_start != 0,
               R);
  }
  R.Sync();
  res.linear_iterations += li.iterations;
  res.last_linear_tolerance = tol;
  res.linear_residual = li.residual;
  return li;
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
// share one body so that a kernel running both (the learned mode) holds one copy of the
// assembly and the linear solve.
template <bool kMixed, class Team>
__device__ femm_item_result_t ItemSolve(const DevicePlan &P, const ItemPointers &w,
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
  bool eq_current = false;
  for (int it = 0; status == FEMM_NONLINEAR_SOLVE_NOT_CONVERGED; ++it) {
    if (safeguarded) {
      // The assembly at the current state is current (start or accepted trial).
      if (it >= o.max_newton_iterations)
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
    const double tol = safeguarded ? o.linear_relative_tolerance : LinearTolerance(o, adaptive, prev);
    if (cfg.pcg_warm_start) {
      for (int i = R.S.row0 + threadIdx.x; i < R.S.row1; i += kBlockThreads)
        w.x[i] = a[P.free_nodes[i]];
      R.Sync();
    }
    const LinearInfo li = LinearSolve<kMixed>(P, w, cfg, tol, o.max_linear_iterations, res, R);
    if (li.status != FEMM_OK) {
      status = li.status;
      break;
    }
    // Undamped step ||A_cand - A|| / ||A_cand|| over all nodes (nonlinear_solver.cu:326-335).
    double s[2] = {0.0, 0.0};
    for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads) {
      const int f = P.node_to_free[v];
      const double cand = f >= 0 ? static_cast<double>(P.node_sign[v]) * w.x[f] : P.boundary[v];
      const double d = cand - a[v];
      s[0] += d * d;
      s[1] += cand * cand;
    }
    R.Sum(s);
    const double step = sqrt(s[0]) / fmax(sqrt(s[1]), DBL_MIN);
    res.newton_iterations = it + 1;
    res.relative_step = step;
    if (!isfinite(step)) {
      status = FEMM_NUMERICAL_NONFINITE;
      break;
    }
    if (safeguarded) {
      if (step < 100.0 * o.relative_tolerance && norm / scale <= o.relative_tolerance) {
        status = FEMM_OK;
        break;
      }
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
        if (!AssembleItem(P, w, w.trial, cur, true, R)) {
          status = FEMM_ASSEMBLY_FAILED;
          break;
        }
        R.Sync();
        const double tn = EquationNorm(P, w, w.trial, R, w.p);
        if (isfinite(tn) &&
            (tn <= (1.0 - 1e-4 * alpha) * norm || tn / scale <= o.relative_tolerance)) {
          for (int v = R.S.node0 + threadIdx.x; v < R.S.node1; v += kBlockThreads)
            a[v] = w.trial[v];
          R.Sync();
          norm = tn;
          accepted = true;
          break;
        }
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
  kModeLearned = 2       // safeguarded from `warm` within a budget, then the driver from the
                         // plain start in the same workgroup if that attempt is not accepted
};

// The nonlinear solve of `kMode` for one item by `R`'s team, results of workgroup 0
// written out.  With prior results, an item the earlier launch kept (KeepPrior) is left
// untouched and the others add their counters to the earlier ones.
template <bool kMixed, int kMode, class Team>
__device__ void SolveItem(const DevicePlan &P, const DeviceWork &W, int item, bool lead,
                          bool inputs_ok, const double *currents, const NewtonArgs &args,
                          const double *warm, double *A, double *flux_out,
                          femm_item_result_t *results, const femm_item_result_t *prior,
                          Team &R)
{
  femm_item_result_t earlier{};
  if (prior != nullptr)
    earlier = prior[item];
  const ItemPointers w = ItemWork(P, W, item, kMixed);
  double *a = A + static_cast<size_t>(item) * P.N;
  const double *cur = currents + static_cast<size_t>(item) * P.C;
  for (int c = 0; c < P.C; ++c)
    inputs_ok = inputs_ok && isfinite(cur[c]);
  const double *warm_item = warm ? warm + static_cast<size_t>(item) * P.N : nullptr;
  // One call site of ItemSolve (the learned mode loops over its two attempts) keeps a
  // single inlined copy of the assembly and the linear solve.
  constexpr int kAttempts = kMode == kModeLearned ? 2 : 1;
  femm_item_result_t res{}, spent{};
#pragma clang loop unroll(disable)
  for (int attempt = 0; attempt < kAttempts; ++attempt) {
    const bool first = attempt == 0;
    const bool safeguarded = kMode == kModeSafeguarded || (kMode == kModeLearned && first);
    const femm_newton_options_t &o =
        (kMode == kModeLearned && first) ? args.warm_options : args.options;
    res = ItemSolve<kMixed>(P, w, cur, a, inputs_ok, o, args.config, first ? warm_item : nullptr,
                            safeguarded, R);
    res.newton_iterations += spent.newton_iterations;
    res.linear_iterations += spent.linear_iterations;
    res.refinements += spent.refinements;
    if (res.status == FEMM_OK)
      break;
    spent = res;
  }
  WriteFlux(P, a, flux_out + static_cast<size_t>(item) * P.C, R, lead);
  if (prior != nullptr) {
    res.newton_iterations += earlier.newton_iterations;
    res.linear_iterations += earlier.linear_iterations;
    res.refinements += earlier.refinements;
  }
  if (lead && threadIdx.x == 0)
    results[item] = res;
}

// One workgroup per batch item.
template <bool kMixed, int kMode>
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
  SolveItem<kMixed, kMode>(P, W, item, true, inputs_ok, currents, args, warm, A, flux_out, results,
                           prior, R);
}

// `partitions` workgroups per batch item (grid = items x partitions, cooperative launch).
// An aborted team (a barrier wait that ran out) reports INTERNAL_ERROR for its item.
template <bool kMixed, int kMode>
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
             W.team_barrier + 2 * static_cast<size_t>(item),
             W.team_abort + item,
             partitions,
             g,
             0};
  SolveItem<kMixed, kMode>(P, W, item, g == 0, inputs_ok, currents, args, warm, A, flux_out,
                           results, prior, R);
  R.Sync();
  if (g == 0 && threadIdx.x == 0 &&
      __hip_atomic_load(R.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
    results[item].status = FEMM_INTERNAL_ERROR;
}

__global__ __launch_bounds__(kBlockThreads) void AssembleKernel(DevicePlan P, DeviceWork W,
                                                                const double *states,
                                                                const double *currents,
                                                                int newton, int32_t *valid)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  const int item = blockIdx.x;
  const ItemPointers w = ItemWork(P, W, item, false);
  const bool ok = AssembleItem(P, w, states + static_cast<size_t>(item) * P.N,
                               currents + static_cast<size_t>(item) * P.C, newton != 0, R);
  if (threadIdx.x == 0)
    valid[item] = ok ? 1 : 0;
}

__global__ __launch_bounds__(kBlockThreads) void PcgKernel(DevicePlan P, DeviceWork W, double tol,
                                                           int maxit, double *info)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  const int item = blockIdx.x;
  const ItemPointers w = ItemWork(P, W, item, false);
  for (int row = R.S.row0 + threadIdx.x; row < R.S.row1; row += kBlockThreads) {
    const double d = w.vals[P.diag_sell[row]];
    w.inv[row] = (isfinite(d) && d > 0.0) ? 1.0 / d : 0.0;
  }
  R.Sync();
  const LinearInfo li = Pcg64(P, w.vals, w.inv, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit, false, R);
  if (threadIdx.x == 0) {
    info[3 * item] = li.status;
    info[3 * item + 1] = li.iterations;
    info[3 * item + 2] = li.residual;
  }
}

// Element B (centroid basis) and flux per item (nonlinear_solver.cu:16-83).
__global__ void ElementBKernel(DevicePlan P, const double *states, int items, double *bx,
                               double *by)
{
  const size_t gid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(items) * P.T;
  if (gid >= total)
    return;
  const int item = static_cast<int>(gid / P.T), t = static_cast<int>(gid % P.T);
  const double *a = states + static_cast<size_t>(item) * P.N;
  const size_t g = static_cast<size_t>(t) * P.Q * 3;   // quadrature point 0 = centroid
  double x = 0.0, y = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double av = a[P.tri_nodes[3 * t + k]];
    x += av * P.tri_gx[g + k];
    y += av * P.tri_gy[g + k];
  }
  bx[gid] = x;
  by[gid] = y;
}

__global__ __launch_bounds__(kBlockThreads) void FluxKernel(DevicePlan P, const double *states,
                                                            double *flux)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  const int item = blockIdx.x;
  WriteFlux(P, states + static_cast<size_t>(item) * P.N, flux + static_cast<size_t>(item) * P.C, R);
}

// ----------------------------------------------------------------------------- postprocess

__device__ void ElementB(const DevicePostPlan &P, const double *a, int e, double *bx, double *by)
{
  const size_t g = static_cast<size_t>(e) * P.Q * 3;   // quadrature point 0 = centroid basis
  double x = 0.0, y = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double av = a[P.tri_nodes[3 * e + k]];
    x += av * P.tri_gx[g + k];
    y += av * P.tri_gy[g + k];
  }
  *bx = x;
  *by = y;
}

// FEMM GetNodalB subset of the reference (postprocess/field_postprocess.cpp:34-190):
// inverse-centroid-distance average in a single-label patch, otherwise the
// CCW / CW traversal to the label interface keeping Bn from A and Bt from the
// source-side element, with the sharp-corner magnitude fallback.
__device__ bool SmoothedNodalB(const DevicePostPlan &P, const double *a, int el, double nbx[3],
                               double nby[3])
{
  const int own_mat = P.tri_mat[el];
  for (int loc = 0; loc < 3; ++loc) {
    const int node = P.tri_nodes[3 * el + loc];
    const int i0 = P.inc_off[node], i1 = P.inc_off[node + 1];
    const double nx = P.node_xy[2 * node], ny = P.node_xy[2 * node + 1];
    bool homo = true;
    for (int k = i0; k < i1; ++k)
      homo = homo && own_mat == P.tri_mat[P.inc_elem[k]];
    if (homo) {
      double ws = 0.0, sx = 0.0, sy = 0.0;
      for (int k = i0; k < i1; ++k) {
        const int nb = P.inc_elem[k];
        const int *t = P.tri_nodes + 3 * nb;
        const double cx = (P.node_xy[2 * t[0]] + P.node_xy[2 * t[1]] + P.node_xy[2 * t[2]]) / 3.0;
        const double cy =
            (P.node_xy[2 * t[0] + 1] + P.node_xy[2 * t[1] + 1] + P.node_xy[2 * t[2] + 1]) / 3.0;
        const double w = 1.0 / hypot(nx - cx, ny - cy);
        if (!(w > 0.0) || !isfinite(w))
          return false;
        double ex, ey;
        ElementB(P, a, nb, &ex, &ey);
        ws += w;
        sx += w * ex;
        sy += w * ey;
      }
      nbx[loc] = sx / ws;
      nby[loc] = sy / ws;
      continue;
    }
    double ws = 0.0, sx = 0.0, sy = 0.0, v1x = 0.0, v1y = 0.0, v2x = 0.0, v2y = 0.0;
    bool special = false;
    for (int pass = 0; pass < 2 && !special; ++pass) {
      int cur = el;
      for (int scan = i0; scan < i1; ++scan) {
        int idx = -1;
        for (int q = 0; q < 3; ++q)
          if (P.tri_nodes[3 * cur + q] == node)
            idx = q;
        if (idx < 0)
          return false;
        const int other = P.tri_nodes[3 * cur + (idx + (pass == 0 ? 2 : 1)) % 3];
        int next = -1;
        for (int k = i0; k < i1 && next < 0; ++k) {
          const int cand = P.inc_elem[k];
          if (cand == cur)
            continue;
          for (int q = 0; q < 3; ++q)
            if (P.tri_nodes[3 * cand + q] == other)
              next = cand;
        }
        if (next < 0) {
          double ex, ey;
          ElementB(P, a, cur, &ex, &ey);
          sx = ex;
          sy = ey;
          special = true;
          break;
        }
        if (own_mat != P.tri_mat[next]) {
          const double dx = P.node_xy[2 * other] - nx, dy = P.node_xy[2 * other + 1] - ny;
          const double len = hypot(dx, dy);
          if (!(len > 0.0))
            return false;
          const double tx = dx / len, ty = dy / len;
          double ex, ey;
          ElementB(P, a, cur, &ex, &ey);
          const double bt = ex * tx + ey * ty;
          const double bn = (a[other] - a[node]) / len;
          const double w = 0.5 / len;
          sx += w * (tx * bt + ty * bn);
          sy += w * (ty * bt - tx * bn);
          ws += w;
          if (pass == 0) {
            v1x = tx;
            v1y = ty;
          } else {
            v2x = tx;
            v2y = ty;
          }
          break;
        }
        cur = next;
      }
    }
    if (special) {
      nbx[loc] = sx;
      nby[loc] = sy;
      continue;
    }
    if (!(ws > 0.0))
      return false;
    nbx[loc] = sx / ws;
    nby[loc] = sy / ws;
    const bool simple = hypot(v1x, v1y) < 0.9 || hypot(v2x, v2y) < 0.9 ||
                        (-v1x * v2x - v1y * v2y) > 0.985;
    if (!simple) {
      double mag = 0.0;
      for (int k = i0; k < i1; ++k) {
        const int nb = P.inc_elem[k];
        if (P.tri_mat[nb] == own_mat) {
          double ex, ey;
          ElementB(P, a, nb, &ex, &ey);
          mag = fmax(mag, hypot(ex, ey));
        }
      }
      double ox, oy;
      ElementB(P, a, el, &ox, &oy);
      const double own = hypot(ox, oy);
      nbx[loc] = own > 0.0 ? mag * ox / own : 0.0;
      nby[loc] = own > 0.0 ? mag * oy / own : 0.0;
    }
  }
  return true;
}

// Per item: WST loads over the elements with a nonzero mask gradient
// (field_postprocess.cpp:369-381) and the smoothed radial B at every probe (:382-398).
__global__ __launch_bounds__(kBlockThreads) void PostprocessKernel(DevicePostPlan P,
                                                                   const double *states,
                                                                   double *loads, double *radial)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockReducer R{lds, 0};
  const int item = blockIdx.x;
  const double *a = states + static_cast<size_t>(item) * P.N;
  constexpr double kMu0 = 4.0e-7 * 3.141592653589793238462643383279502884;
  double s[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < P.nwst; k += kBlockThreads) {
    double bx, by;
    ElementB(P, a, P.wst_elem[k], &bx, &by);
    const double hx = P.wst_hx[k], hy = P.wst_hy[k];
    const double fx = ((bx * bx - by * by) * hx + 2.0 * bx * by * hy) / (2.0 * kMu0);
    const double fy = (2.0 * bx * by * hx + (by * by - bx * bx) * hy) / (2.0 * kMu0);
    const double w = P.wst_w[k];
    s[0] += w * fx;
    s[1] += w * fy;
    s[2] += w * (P.wst_cx[k] * fy - P.wst_cy[k] * fx);
  }
  R.Sum(s);
  if (threadIdx.x == 0) {
    loads[3 * item] = s[0];
    loads[3 * item + 1] = s[1];
    loads[3 * item + 2] = s[2];
  }
  for (int p = threadIdx.x; p < P.nprobe; p += kBlockThreads) {
    double nbx[3], nby[3];
    double value = NAN;
    if (SmoothedNodalB(P, a, P.probe_elem[p], nbx, nby)) {
      double bx = 0.0, by = 0.0;
      for (int k = 0; k < 3; ++k) {
        bx += P.probe_lambda[3 * p + k] * nbx[k];
        by += P.probe_lambda[3 * p + k] * nby[k];
      }
      value = bx * P.probe_cos[p] + by * P.probe_sin[p];
    }
    radial[static_cast<size_t>(item) * P.nprobe + p] = value;
  }
}

template <typename T>
T *Carve(char *&cursor, size_t count)
{
  T *p = reinterpret_cast<T *>(cursor);
  cursor += (count * sizeof(T) + 255) / 256 * 256;
  return p;
}

template <typename T>
size_t Bytes(size_t count)
{
  return (count * sizeof(T) + 255) / 256 * 256;
}

}  // namespace

int UploadPlan(const HostPlan &h, DeviceBuffers *dev)
{
  // Two passes over the same sequence of arrays: the first (null base) measures
  // the 256-byte-aligned layout, the second allocates once and copies.
  DevicePlan &P = dev->plan;
  for (int pass = 0; pass < 2; ++pass) {
    char *base = static_cast<char *>(dev->base);
    size_t offset = 0;
    bool copy_ok = true;
    auto up = [&](const auto &vec) {
      using T = typename std::decay_t<decltype(vec)>::value_type;
      T *p = reinterpret_cast<T *>(base + offset);
      offset += (vec.size() * sizeof(T) + 255) / 256 * 256;
      if (pass == 1 && !vec.empty() &&
          hipMemcpy(p, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        copy_ok = false;
      return static_cast<const T *>(p);
    };
    P.tri_nodes = up(h.tri_nodes);
    P.tri_mat = up(h.tri_mat);
    P.tri_circuit = up(h.tri_circuit);
    P.tri_w = up(h.tri_w);
    P.tri_gx = up(h.tri_gx);
    P.tri_gy = up(h.tri_gy);
    P.tri_flux = up(h.tri_flux);
    P.mat_nu0 = up(h.mat_nu0);
    P.mat_alpha = up(h.mat_alpha);
    P.mat_cold = up(h.mat_cold);
    P.mat_curve_off = up(h.mat_curve_off);
    P.mat_curve_cnt = up(h.mat_curve_cnt);
    P.curve_b = up(h.curve_b);
    P.curve_h = up(h.curve_h);
    P.curve_s = up(h.curve_s);
    P.node_to_free = up(h.node_to_free);
    P.free_nodes = up(h.free_nodes);
    P.node_sign = up(h.node_sign);
    P.boundary = up(h.boundary);
    P.slot_off = up(h.slot_off);
    P.contrib = up(h.contrib);
    P.row_off = up(h.row_off);
    P.rcontrib = up(h.rcontrib);
    P.pm_rhs = up(h.pm_rhs);
    P.circuit_rhs = up(h.circuit_rhs);
    P.slice_ptr = up(h.slice_ptr);
    P.slice_width = up(h.slice_width);
    P.sell_cols = up(h.sell_cols);
    P.slot_to_sell = up(h.slot_to_sell);
    P.diag_sell = up(h.diag_sell);
    P.row_offsets = up(h.row_offsets);
    if (pass == 0) {
      void *mem = nullptr;
      if (hipMalloc(&mem, offset) != hipSuccess)
        return FEMM_GPU_ALLOCATION_FAILED;
      dev->base = mem;
    } else if (!copy_ok) {
      return FEMM_INTERNAL_ERROR;
    }
  }
  P.N = h.N;
  P.T = h.T;
  P.n = h.n;
  P.nnz = h.nnz;
  P.C = h.C;
  P.Q = h.Q;
  P.nslices = h.nslices;
  P.sell_size = h.sell_size;
  P.planar = h.planar;
  P.any_nonlinear = h.any_nonlinear ? 1 : 0;
  P.any_boundary_value = h.any_boundary_value ? 1 : 0;
  P.depth = h.depth;
  return hipDeviceSynchronize() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

namespace
{

// Visits the variant fields in layout order: fn(member pointer, per-variant element count).
template <typename Fn>
void ForEachVariantField(const HostPlan &topo, Fn &&fn)
{
  fn(&HostPlan::tri_mat, topo.tri_mat.size());
  fn(&HostPlan::tri_circuit, topo.tri_circuit.size());
  fn(&HostPlan::tri_flux, topo.tri_flux.size());
  fn(&HostPlan::mat_nu0, topo.mat_nu0.size());
  fn(&HostPlan::mat_alpha, topo.mat_alpha.size());
  fn(&HostPlan::mat_cold, topo.mat_cold.size());
  fn(&HostPlan::mat_curve_off, topo.mat_curve_off.size());
  fn(&HostPlan::mat_curve_cnt, topo.mat_curve_cnt.size());
  fn(&HostPlan::curve_b, topo.curve_b.size());
  fn(&HostPlan::curve_h, topo.curve_h.size());
  fn(&HostPlan::curve_s, topo.curve_s.size());
  fn(&HostPlan::pm_rhs, topo.pm_rhs.size());
  fn(&HostPlan::circuit_rhs, topo.circuit_rhs.size());
}

// Device array of a variant field.
const int32_t *&FieldSlot(DeviceVariants &D, std::vector<int32_t> HostPlan::*m)
{
  if (m == &HostPlan::tri_mat)
    return D.tri_mat;
  if (m == &HostPlan::tri_circuit)
    return D.tri_circuit;
  if (m == &HostPlan::mat_curve_off)
    return D.mat_curve_off;
  return D.mat_curve_cnt;
}

const double *&FieldSlot(DeviceVariants &D, std::vector<double> HostPlan::*m)
{
  if (m == &HostPlan::tri_flux)
    return D.tri_flux;
  if (m == &HostPlan::mat_nu0)
    return D.mat_nu0;
  if (m == &HostPlan::mat_alpha)
    return D.mat_alpha;
  if (m == &HostPlan::mat_cold)
    return D.mat_cold;
  if (m == &HostPlan::curve_b)
    return D.curve_b;
  if (m == &HostPlan::curve_h)
    return D.curve_h;
  if (m == &HostPlan::curve_s)
    return D.curve_s;
  if (m == &HostPlan::pm_rhs)
    return D.pm_rhs;
  return D.circuit_rhs;
}

}  // namespace

int AllocateVariants(const HostPlan &topo, int32_t count, DeviceBuffers *dev)
{
  if (dev->variants_base) {
    if (hipDeviceSynchronize() != hipSuccess)
      return FEMM_INTERNAL_ERROR;
    (void)hipFree(dev->variants_base);
    dev->variants_base = nullptr;
  }
  dev->variants = DeviceVariants{};
  if (count <= 0)
    return FEMM_OK;
  const size_t V = static_cast<size_t>(count);
  size_t bytes = 0;
  ForEachVariantField(topo, [&](auto member, size_t stride) {
    using T = typename std::decay_t<decltype(topo.*member)>::value_type;
    bytes += (V * stride * sizeof(T) + 255) / 256 * 256;
  });
  bytes += (V * sizeof(int32_t) + 255) / 256 * 256;
  void *mem = nullptr;
  if (hipMalloc(&mem, bytes) != hipSuccess)
    return FEMM_GPU_ALLOCATION_FAILED;
  if (hipMemset(mem, 0, bytes) != hipSuccess) {
    (void)hipFree(mem);
    return FEMM_INTERNAL_ERROR;
  }
  dev->variants_base = mem;
  DeviceVariants &D = dev->variants;
  char *cursor = static_cast<char *>(mem);
  ForEachVariantField(topo, [&](auto member, size_t stride) {
    using T = typename std::decay_t<decltype(topo.*member)>::value_type;
    FieldSlot(D, member) = reinterpret_cast<const T *>(cursor);
    cursor += (V * stride * sizeof(T) + 255) / 256 * 256;
  });
  D.any_nonlinear = reinterpret_cast<const int32_t *>(cursor);
  D.count = count;
  D.mats = static_cast<int32_t>(topo.mat_nu0.size());
  D.knots = static_cast<int32_t>(topo.curve_b.size());
  return FEMM_OK;
}

int CopyVariant(const HostPlan &topo, int32_t index, const HostPlan &part, DeviceBuffers *dev)
{
  DeviceVariants &D = dev->variants;
  if (index < 0 || index >= D.count)
    return FEMM_INVALID_ARGUMENT;
  bool ok = true;
  ForEachVariantField(topo, [&](auto member, size_t stride) {
    using T = typename std::decay_t<decltype(topo.*member)>::value_type;
    const std::vector<T> &src = part.*member;
    if (src.size() != stride) {
      ok = false;
      return;
    }
    T *dst = const_cast<T *>(FieldSlot(D, member)) + static_cast<size_t>(index) * stride;
    if (ok && stride &&
        hipMemcpy(dst, src.data(), stride * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      ok = false;
  });
  if (!ok)
    return FEMM_INVALID_ARGUMENT;
  const int32_t flag = part.any_nonlinear ? 1 : 0;
  if (hipMemcpy(const_cast<int32_t *>(D.any_nonlinear) + index, &flag, sizeof(int32_t),
                hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  return FEMM_OK;
}

int EnsureWork(const HostPlan &h, int32_t items, bool mixed, DeviceBuffers *dev)
{
  const bool have_mixed = dev->work.vals32 != nullptr;
  if (dev->work_base && dev->work.capacity >= items && (!mixed || have_mixed))
    return FEMM_OK;
  const int32_t cap = std::max(items, dev->work.capacity);
  const bool want_mixed = mixed || have_mixed;
  if (dev->work_base) {
    if (hipDeviceSynchronize() != hipSuccess)
      return FEMM_INTERNAL_ERROR;
    (void)hipFree(dev->work_base);
    dev->work_base = nullptr;
  }
  const size_t ci = static_cast<size_t>(cap);
  size_t bytes = Bytes<double>(ci * h.T * kElemStride) + Bytes<double>(ci * h.sell_size) +
                 6 * Bytes<double>(ci * h.n) + Bytes<double>(ci * h.N);
  if (want_mixed)
    bytes += Bytes<float>(ci * h.sell_size) + 5 * Bytes<float>(ci * h.n);
  bytes += Bytes<double>(ci * 2 * kMaxPartitions * 4) + Bytes<uint32_t>(ci * 2) +
           Bytes<int32_t>(ci);
  void *base = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess)
    return FEMM_GPU_ALLOCATION_FAILED;
  // SELL padding entries must read as zero; they are never written afterwards.
  if (hipMemset(base, 0, bytes) != hipSuccess) {
    (void)hipFree(base);
    return FEMM_INTERNAL_ERROR;
  }
  char *cur = static_cast<char *>(base);
  DeviceWork &w = dev->work;
  w = DeviceWork{};
  w.capacity = cap;
  w.elem = Carve<double>(cur, ci * h.T * kElemStride);
  w.vals = Carve<double>(cur, ci * h.sell_size);
  w.rhs = Carve<double>(cur, ci * h.n);
  w.inv = Carve<double>(cur, ci * h.n);
  w.x = Carve<double>(cur, ci * h.n);
  w.r = Carve<double>(cur, ci * h.n);
  w.p = Carve<double>(cur, ci * h.n);
  w.ap = Carve<double>(cur, ci * h.n);
  w.trial = Carve<double>(cur, ci * h.N);
  if (want_mixed) {
    w.vals32 = Carve<float>(cur, ci * h.sell_size);
    w.r32 = Carve<float>(cur, ci * h.n);
    w.p32 = Carve<float>(cur, ci * h.n);
    w.ap32 = Carve<float>(cur, ci * h.n);
    w.d32 = Carve<float>(cur, ci * h.n);
    w.inv32 = Carve<float>(cur, ci * h.n);
  }
  w.team_partials = Carve<double>(cur, ci * 2 * kMaxPartitions * 4);
  w.team_barrier = Carve<uint32_t>(cur, ci * 2);
  w.team_abort = Carve<int32_t>(cur, ci);
  dev->work_base = base;
  dev->work_bytes = bytes;
  return hipDeviceSynchronize() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

void ReleaseDevice(DeviceBuffers *dev)
{
  if (dev->variants_base)
    (void)hipFree(dev->variants_base);
  dev->variants_base = nullptr;
  dev->variants = DeviceVariants{};
  if (dev->work_base)
    (void)hipFree(dev->work_base);
  if (dev->base)
    (void)hipFree(dev->base);
  dev->work_base = nullptr;
  dev->base = nullptr;
}

namespace
{

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

}  // namespace

int AutoPartitions(const DeviceBuffers &dev, int32_t items)
{
  for (const auto &t : dev.tuned)
    if (t.first == items)
      return t.second;
  int device = 0, cus = 0;
  if (items <= 0 || hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 1;
  // Untuned rule: one workgroup per item once the batch fills half the CUs; otherwise
  // split the CUs over the items with about 32 slices (2,048 rows) per workgroup, where
  // team barriers and per-CU bandwidth balanced on the 58k-node motor.
  if (items * 2 >= cus)
    return 1;
  const int by_cus = cus / items;
  const int by_rows = std::max(1, dev.plan.nslices / 32);
  return std::max(1, std::min({by_cus, by_rows, kMaxPartitions}));
}

namespace
{

template <bool kMixed, int kMode>
int LaunchTeams(const DeviceBuffers &dev, int32_t items, int32_t partitions,
                const double *currents, const NewtonArgs &args, const double *warm, double *a,
                double *flux, femm_item_result_t *results, hipStream_t stream,
                const LaunchExtras &x)
{
  const void *kernel = reinterpret_cast<const void *>(&TeamSolveKernel<kMixed, kMode>);
  const long grid = static_cast<long>(items) * partitions;
  if (partitions > kMaxPartitions || grid > CoResidentBlocks(kernel))
    return FEMM_INVALID_ARGUMENT;
  // Fresh barrier words and abort flags for the launched items.
  if (hipMemsetAsync(dev.work.team_barrier, 0, sizeof(uint32_t) * 2 * items, stream) !=
          hipSuccess ||
      hipMemsetAsync(dev.work.team_abort, 0, sizeof(int32_t) * items, stream) != hipSuccess)
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

template <int kMode>
int LaunchSolve(const DeviceBuffers &dev, int32_t items, const double *currents,
                const NewtonArgs &args_in, const double *warm, double *a, double *flux,
                femm_item_result_t *results, hipStream_t stream, const LaunchExtras &x)
{
  if (items <= 0 || items > dev.work.capacity)
    return FEMM_INVALID_ARGUMENT;
  NewtonArgs args = args_in;
  if (args.config.partitions == 0)
    args.config.partitions = AutoPartitions(dev, items);
  const int32_t partitions = args.config.partitions;
  const bool mixed = args.config.precision == FEMM_PRECISION_MIXED;
  if (partitions > 1) {
    if constexpr (kMode == kModeLearned) {
      // Partitioned: the warm attempt and the plain-start fallback as two launches.
      NewtonArgs warm_args = args;
      warm_args.options = args.warm_options;
      int st = mixed ? LaunchTeams<true, kModeSafeguarded>(dev, items, partitions, currents,
                                                         warm_args, warm, a, flux, results,
                                                         stream, x)
                     : LaunchTeams<false, kModeSafeguarded>(dev, items, partitions, currents,
                                                          warm_args, warm, a, flux, results,
                                                          stream, x);
      if (st != FEMM_OK)
        return st;
      LaunchExtras redo = x;
      redo.prior = results;
      NewtonArgs plain_args = args;
      plain_args.prior_mode = kPriorRedoFailed;
      return mixed ? LaunchTeams<true, kModeNewton>(dev, items, partitions, currents, plain_args,
                                                   nullptr, a, flux, results, stream, redo)
                   : LaunchTeams<false, kModeNewton>(dev, items, partitions, currents,
                                                    plain_args, nullptr, a, flux, results, stream,
                                                    redo);
    } else {
      return mixed ? LaunchTeams<true, kMode>(dev, items, partitions, currents, args, warm, a,
                                             flux, results, stream, x)
                   : LaunchTeams<false, kMode>(dev, items, partitions, currents, args, warm, a,
                                              flux, results, stream, x);
    }
  }
  if (partitions < 0)
    return FEMM_INVALID_ARGUMENT;
  if (mixed)
    hipLaunchKernelGGL((SolveKernel<true, kMode>), dim3(items), dim3(kBlockThreads), 0, stream,
                       dev.plan, dev.work, currents, args, warm, a, flux, results, x.prior,
                       dev.variants, x.variant_of_item);
  else
    hipLaunchKernelGGL((SolveKernel<false, kMode>), dim3(items), dim3(kBlockThreads), 0, stream,
                       dev.plan, dev.work, currents, args, warm, a, flux, results, x.prior,
                       dev.variants, x.variant_of_item);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace

int LaunchNewton(const DeviceBuffers &dev, int32_t items, const double *currents,
                 const femm_newton_options_t &options, const femm_solve_config_t &config,
                 const double *warm, double *a, double *flux, femm_item_result_t *results,
                 hipStream_t stream, const LaunchExtras &x)
{
  const NewtonArgs args{options, config, x.prior ? x.prior_mode : kPriorNone, options};
  return LaunchSolve<kModeNewton>(dev, items, currents, args, warm, a, flux, results, stream, x);
}

int LaunchSafeguarded(const DeviceBuffers &dev, int32_t items, const double *currents,
                      const femm_newton_options_t &options, const femm_solve_config_t &config,
                      const double *warm, double *a, double *flux, femm_item_result_t *results,
                      hipStream_t stream, const LaunchExtras &x)
{
  const NewtonArgs args{options, config, x.prior ? x.prior_mode : kPriorNone, options};
  return LaunchSolve<kModeSafeguarded>(dev, items, currents, args, warm, a, flux, results, stream,
                                       x);
}

int LaunchLearned(const DeviceBuffers &dev, int32_t items, const double *currents,
                  const femm_newton_options_t &warm_options,
                  const femm_newton_options_t &plain_options, const femm_solve_config_t &config,
                  const double *warm, double *a, double *flux, femm_item_result_t *results,
                  hipStream_t stream, const LaunchExtras &x)
{
  if (warm == nullptr)
    return FEMM_INVALID_ARGUMENT;
  const NewtonArgs args{plain_options, config, x.prior ? x.prior_mode : kPriorNone, warm_options};
  return LaunchSolve<kModeLearned>(dev, items, currents, args, warm, a, flux, results, stream, x);
}

int LaunchAssemble(const DeviceBuffers &dev, int32_t items, const double *states,
                   const double *currents, int32_t newton, int32_t *valid, hipStream_t stream)
{
  if (items <= 0 || items > dev.work.capacity)
    return FEMM_INVALID_ARGUMENT;
  hipLaunchKernelGGL(AssembleKernel, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan,
                     dev.work, states, currents, newton, valid);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

int LaunchPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit, double *info,
              hipStream_t stream)
{
  if (items <= 0 || items > dev.work.capacity)
    return FEMM_INVALID_ARGUMENT;
  hipLaunchKernelGGL(PcgKernel, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan, dev.work,
                     tol, maxit, info);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

int LaunchFinalize(const DeviceBuffers &dev, int32_t items, const double *states, double *bx,
                   double *by, double *flux, hipStream_t stream)
{
  if (items <= 0)
    return FEMM_INVALID_ARGUMENT;
  if (bx && by) {
    const size_t total = static_cast<size_t>(items) * dev.plan.T;
    hipLaunchKernelGGL(ElementBKernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256),
                       0, stream, dev.plan, states, items, bx, by);
  }
  if (flux)
    hipLaunchKernelGGL(FluxKernel, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan, states,
                       flux);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm

namespace femm
{

int LaunchPostprocess(const DevicePostPlan &post, int32_t items, const double *states,
                      double *loads, double *radial_b, hipStream_t stream)
{
  if (items <= 0)
    return FEMM_INVALID_ARGUMENT;
  hipLaunchKernelGGL(PostprocessKernel, dim3(items), dim3(kBlockThreads), 0, stream, post, states,
                     loads, radial_b);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm
