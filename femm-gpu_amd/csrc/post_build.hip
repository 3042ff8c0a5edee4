// Weighted-stress postprocess plan built on the GPU: PrepareWeightedStressMask and
// SolveWeightedStressMask (postprocess/weighted_stress.cu:24-183) and
// BuildFrozenPostprocessPlanWithMask (postprocess/field_postprocess.cpp:221-339).
//
// The host checks the options, classifies the materials (selected / air / other, one byte per
// material) and takes cos / sin of the probe angles.  Everything per node and per element runs
// in the kernels below, in this order:
//  * element classes: the first selected and the first other element touching each node, and
//    node incidence counts;
//  * node incidence lists: counts scanned, lists filled by atomics, then each list sorted, so
//    it is in element order whatever order the atomics ran in;
//  * the edge rule: a selected element may share an edge only with selected or air elements
//    (weighted_stress.cu:49-78);
//  * fixed mask values: Dirichlet nodes 0, nodes of selected elements 1, other nodes of
//    non-air elements 0.  The reference visits the elements in order and rejects a selected
//    element meeting a node already fixed at a value other than 1 (:80-98); per node that is
//    "a Dirichlet node of a selected element, or a node whose first other element comes
//    before its first selected element";
//  * the mask Laplacian (weight sqrt(area), :100-124) on this plan's own CSR pattern and DOF
//    classes.  The mask ties the periodic pairs with sign +1 (:134-151): its classes are the
//    plan's unions with every sign +1, so each CSR slot sums its contributions with sign +1
//    (a Dirichlet class of the plan holds the mask at 0).  A class holding a fixed node is fixed: its row
//    becomes an identity row with a zero right-hand side and its couplings move to the free
//    rows' right-hand sides, so the library's PCG (split over co-resident workgroups) solves
//    the reduced system of the reference with the fixed rows inert (their residual, direction
//    and update stay exactly zero);
//  * node mask values thresholded at 0.5 (FEMM WeightingScheme 0);
//  * per-element mask gradients, kept where nonzero, compacted in element order by a scan
//    (field_postprocess.cpp:248-269);
//  * probe location: the first element, in element order, whose box passes the radius test
//    and which contains the point (:274-336), found by atomicMin over (probe, element).
// The integer work is exact and the floating-point formulas are the host's, term by term
// (device code is compiled without FMA contraction), so gradients, weights, centroids and
// barycentric coordinates are bitwise those of a host build.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>
#include <vector>

#include "post.hpp"

namespace femm
{
namespace
{

constexpr int kThreads = 256;
constexpr int kScanThreads = 1024;
constexpr int kProbeRows = 1024;   // probes per launch row (grid y); larger counts loop
constexpr int8_t kAir = 0, kSelected = 1, kOther = 2;

// flags[0]: rule bits; flags[1]: nodes without a fixed value; flags[2]: free mask classes.
constexpr uint32_t kSeen = 1u, kEdgeShared = 2u, kFixedConflict = 4u, kClassConflict = 8u,
                   kBadElement = 16u, kProbeMissing = 32u;

int Blocks(size_t count) { return static_cast<int>((count + kThreads - 1) / kThreads); }

__global__ void FillInt(int32_t *p, int count, int32_t value)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count)
    p[i] = value;
}

__global__ void MarkDirichlet(int count, const int32_t *nodes, int N, int8_t *dir)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const int v = nodes[i];
    if (v >= 0 && v < N)
      dir[v] = 1;
  }
}

__global__ void ClassifyElements(int T, const int32_t *tri, const int32_t *tri_mat,
                                 const int8_t *mat_class, int32_t *first_sel,
                                 int32_t *first_other, int32_t *count, uint32_t *flags)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T)
    return;
  const int8_t c = mat_class[tri_mat[e]];
  for (int k = 0; k < 3; ++k) {
    const int v = tri[3 * e + k];
    atomicAdd(&count[v], 1);
    if (c == kSelected)
      atomicMin(&first_sel[v], e);
    else if (c == kOther)
      atomicMin(&first_other[v], e);
  }
  if (c == kSelected)
    atomicOr(&flags[0], kSeen);
}

// out[0..L) = exclusive prefix sums of in[0..L), out[L] = total; one workgroup, each thread a
// contiguous chunk.  A plan-build step run once per geometry.
__global__ __launch_bounds__(kScanThreads) void ExclusiveScan(const int32_t *in, int32_t *out,
                                                              int L)
{
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  const int chunk = (L + kScanThreads - 1) / kScanThreads;
  const int lo = min(L, t * chunk), hi = min(L, lo + chunk);
  int32_t s = 0;
  for (int i = lo; i < hi; ++i)
    s += in[i];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < kScanThreads; d <<= 1) {
    const int32_t add = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int32_t run = t > 0 ? part[t - 1] : 0;
  for (int i = lo; i < hi; ++i) {
    const int32_t x = in[i];
    out[i] = run;
    run += x;
  }
  if (t == kScanThreads - 1)
    out[L] = part[t];
}

__global__ void FillIncidence(int T, const int32_t *tri, const int32_t *off, int32_t *cursor,
                              int32_t *inc)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T)
    return;
  for (int k = 0; k < 3; ++k) {
    const int v = tri[3 * e + k];
    inc[off[v] + atomicAdd(&cursor[v], 1)] = e;
  }
}

// Each node's incident elements in increasing element order (insertion sort; lists are short).
__global__ void SortIncidence(int N, const int32_t *off, int32_t *inc)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N)
    return;
  const int lo = off[v], hi = off[v + 1];
  for (int i = lo + 1; i < hi; ++i) {
    const int32_t key = inc[i];
    int j = i - 1;
    while (j >= lo && inc[j] > key) {
      inc[j + 1] = inc[j];
      --j;
    }
    inc[j + 1] = key;
  }
}

__global__ void CheckEdges(int T, const int32_t *tri, const int32_t *tri_mat,
                           const int8_t *mat_class, const int32_t *off, const int32_t *inc,
                           uint32_t *flags)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T || mat_class[tri_mat[e]] != kSelected)
    return;
  for (int k = 0; k < 3; ++k) {
    const int a = tri[3 * e + k], b = tri[3 * e + (k + 1) % 3];
    for (int q = off[a]; q < off[a + 1]; ++q) {
      const int o = inc[q];
      if (mat_class[tri_mat[o]] != kOther)
        continue;
      const int32_t *to = tri + 3 * o;
      if (to[0] == b || to[1] == b || to[2] == b) {   // o holds both ends: it has edge {a, b}
        atomicOr(&flags[0], kEdgeShared);
        return;
      }
    }
  }
}

// code[v]: -1 no fixed value, else the fixed mask value (0 or 1).  class_bits[f] collects
// bit (1 << value) over the fixed nodes of free plan class f.
__global__ void FixNodes(int N, const int8_t *dir, const int32_t *first_sel,
                         const int32_t *first_other, const int32_t *node_to_free, int8_t *code,
                         int32_t *class_bits, uint32_t *flags)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N)
    return;
  const bool sel = first_sel[v] != INT_MAX, oth = first_other[v] != INT_MAX;
  int8_t c = -1;
  if (dir[v]) {
    c = 0;
    if (sel)
      atomicOr(&flags[0], kFixedConflict);
  } else if (sel) {
    c = 1;
    if (oth && first_other[v] < first_sel[v])
      atomicOr(&flags[0], kFixedConflict);
  } else if (oth) {
    c = 0;
  }
  code[v] = c;
  if (c < 0) {
    atomicAdd(&flags[1], 1u);
    return;
  }
  const int f = node_to_free[v];
  if (f >= 0)
    atomicOr(&class_bits[f], 1 << c);
  else if (c == 1)   // a class fixed by a Dirichlet node holds the mask at 0
    atomicOr(&flags[0], kClassConflict);
}

__global__ void CountFreeClasses(int n, const int32_t *class_bits, uint32_t *flags)
{
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n)
    return;
  if (class_bits[f] == 0)
    atomicAdd(&flags[2], 1u);
  else if (class_bits[f] == 3)
    atomicOr(&flags[0], kClassConflict);
}

// sqrt(area) * area * grad(lambda_i) . grad(lambda_j) of element e (weighted_stress.cu:108-124).
__device__ double MaskTerm(const int32_t *tri, const double *xy, int e, int i, int j,
                           uint32_t *flags)
{
  const int32_t *t = tri + 3 * e;
  const double x[3] = {xy[2 * t[0]], xy[2 * t[1]], xy[2 * t[2]]};
  const double y[3] = {xy[2 * t[0] + 1], xy[2 * t[1] + 1], xy[2 * t[2] + 1]};
  const double det = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
  if (!(det > 0.0) || !isfinite(det)) {
    atomicOr(&flags[0], kBadElement);
    return 0.0;
  }
  const double area = 0.5 * det;
  const double gxi = (y[(i + 1) % 3] - y[(i + 2) % 3]) / det;
  const double gyi = (x[(i + 2) % 3] - x[(i + 1) % 3]) / det;
  const double gxj = (y[(j + 1) % 3] - y[(j + 2) % 3]) / det;
  const double gyj = (x[(j + 2) % 3] - x[(j + 1) % 3]) / det;
  return sqrt(area) * area * (gxi * gxj + gyi * gyj);
}

// The mask system on the plan's pattern, one thread per row: slot values (every contribution
// with sign +1), fixed rows as identity rows, couplings to fixed classes moved to the rhs.
__global__ void MaskSystem(DevicePlan P, const double *xy, const int32_t *class_bits,
                           double *vals, double *rhs, uint32_t *flags)
{
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= P.n)
    return;
  const bool row_fixed = class_bits[f] != 0;
  double b = 0.0;
  for (int slot = P.row_offsets[f]; slot < P.row_offsets[f + 1]; ++slot) {
    const int pos = P.slot_to_sell[slot];
    const int g = P.sell_cols[pos];
    double k = 0.0;
    for (int q = P.slot_off[slot]; q < P.slot_off[slot + 1]; ++q) {
      const int c = P.contrib[q];
      k += MaskTerm(P.tri_nodes, xy, c >> 5, (c >> 3) & 3, (c >> 1) & 3, flags);
    }
    double value = k;
    if (row_fixed) {
      value = g == f ? 1.0 : 0.0;
    } else if (class_bits[g] != 0) {
      b -= k * static_cast<double>(class_bits[g] >> 1);
      value = 0.0;
    }
    vals[pos] = value;
  }
  rhs[f] = row_fixed ? 0.0 : b;
}

__global__ void ExpandMask(int N, const int8_t *code, const int32_t *node_to_free,
                           const int32_t *class_bits, const double *x, double *mask)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N)
    return;
  double value;
  if (code[v] >= 0) {
    value = code[v];
  } else {
    const int f = node_to_free[v];
    value = f < 0 ? 0.0 : (class_bits[f] != 0 ? static_cast<double>(class_bits[f] >> 1) : x[f]);
  }
  mask[v] = value > 0.5 ? 1.0 : 0.0;
}

struct Gradient {
  double hx, hy, det;
};

__device__ Gradient MaskGradient(const int32_t *tri, const double *xy, const double *mask, int e)
{
  const int32_t *t = tri + 3 * e;
  const double *p0 = xy + 2 * t[0], *p1 = xy + 2 * t[1], *p2 = xy + 2 * t[2];
  const double det = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]);
  const double dx[3] = {p1[1] - p2[1], p2[1] - p0[1], p0[1] - p1[1]};
  const double dy[3] = {p2[0] - p1[0], p0[0] - p2[0], p1[0] - p0[0]};
  double hx = 0.0, hy = 0.0;
  for (int k = 0; k < 3; ++k) {
    hx += mask[t[k]] * (-dx[k] / det);
    hy -= mask[t[k]] * (dy[k] / det);
  }
  return {hx, hy, det};
}

__global__ void GradientFlags(int T, const int32_t *tri, const double *xy, const double *mask,
                              int32_t *keep)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T)
    return;
  const Gradient g = MaskGradient(tri, xy, mask, e);
  keep[e] = (g.hx != 0.0 || g.hy != 0.0) ? 1 : 0;
}

struct WstArrays {
  int32_t *elem;
  double *hx, *hy, *w, *cx, *cy;
};

__global__ void WriteGradients(int T, const int32_t *tri, const double *xy, const double *mask,
                               const int32_t *keep, const int32_t *pos, double depth, WstArrays W)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T || !keep[e])
    return;
  const Gradient g = MaskGradient(tri, xy, mask, e);
  const int32_t *t = tri + 3 * e;
  const double *p0 = xy + 2 * t[0], *p1 = xy + 2 * t[1], *p2 = xy + 2 * t[2];
  const int k = pos[e];
  W.elem[k] = e;
  W.hx[k] = g.hx;
  W.hy[k] = g.hy;
  W.w[k] = 0.5 * g.det * depth;
  W.cx[k] = (p0[0] + p1[0] + p2[0]) / 3.0;
  W.cy[k] = (p0[1] + p1[1] + p2[1]) / 3.0;
}

__device__ double Lo(double a, double b) { return b < a ? b : a; }   // std::min
__device__ double Hi(double a, double b) { return a < b ? b : a; }   // std::max

// Barycentric coordinates of (x, y) in element e; false unless det > 0 and all finite.
__device__ bool Locate(const int32_t *tri, const double *xy, int e, double x, double y,
                       double lam[3])
{
  const int32_t *t = tri + 3 * e;
  const double *p0 = xy + 2 * t[0], *p1 = xy + 2 * t[1], *p2 = xy + 2 * t[2];
  const double det = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]);
  if (!(det > 0.0))
    return false;
  const double l1 = ((x - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (y - p0[1])) / det;
  const double l2 = ((p1[0] - p0[0]) * (y - p0[1]) - (x - p0[0]) * (p1[1] - p0[1])) / det;
  const double l0 = 1.0 - l1 - l2;
  lam[0] = l0;
  lam[1] = l1;
  lam[2] = l2;
  return isfinite(l0) && isfinite(l1) && isfinite(l2);
}

// found[s] = the smallest candidate element containing probe s (INT_MAX if none).
__global__ void LocateProbes(int T, int S, const int32_t *tri, const double *xy, double r2,
                             double tol, const double *px, const double *py, int32_t *found)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= T)
    return;
  const int32_t *t = tri + 3 * e;
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 2; ++d) {
      lo[d] = Lo(lo[d], xy[2 * t[k] + d]);
      hi[d] = Hi(hi[d], xy[2 * t[k] + d]);
    }
  const double cx = lo[0] > 0.0 ? lo[0] : (hi[0] < 0.0 ? hi[0] : 0.0);
  const double cy = lo[1] > 0.0 ? lo[1] : (hi[1] < 0.0 ? hi[1] : 0.0);
  const double fx = Hi(fabs(lo[0]), fabs(hi[0]));
  const double fy = Hi(fabs(lo[1]), fabs(hi[1]));
  if (!(r2 + tol >= cx * cx + cy * cy && r2 <= fx * fx + fy * fy + tol))
    return;
  for (int s = blockIdx.y; s < S; s += gridDim.y) {
    double lam[3];
    if (Locate(tri, xy, e, px[s], py[s], lam) && lam[0] >= -1e-12 && lam[1] >= -1e-12 &&
        lam[2] >= -1e-12)
      atomicMin(&found[s], e);
  }
}

__global__ void ProbeOutput(int S, const int32_t *tri, const int32_t *tri_mat,
                            const int8_t *mat_class, const double *xy, const double *px,
                            const double *py, const int32_t *found, int32_t *probe_elem,
                            double *lambda, uint32_t *flags)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S)
    return;
  const int e = found[s];
  double lam[3] = {0.0, 0.0, 0.0};
  if (e == INT_MAX || mat_class[tri_mat[e]] != kAir) {
    atomicOr(&flags[0], kProbeMissing);
    probe_elem[s] = 0;
  } else {
    Locate(tri, xy, e, px[s], py[s], lam);
    probe_elem[s] = e;
  }
  for (int k = 0; k < 3; ++k)
    lambda[3 * s + k] = lam[k];
}

// Bump allocator over one device allocation (256-byte aligned pieces).
struct Arena {
  char *base = nullptr;
  size_t offset = 0;
  template <class T>
  T *Take(size_t count)
  {
    T *p = reinterpret_cast<T *>(base + offset);
    offset += (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
    return p;
  }
};

struct Scratch {
  int8_t *mat_class, *dir, *code;
  int32_t *dnodes, *first_sel, *first_other, *count, *cursor, *class_bits, *keep, *pos, *found;
  int32_t *inc_off, *inc_elem, *probe_elem;
  uint32_t *flags;
  double *xy, *mask, *info, *px, *py, *pcos, *psin, *lambda;
  WstArrays wst;
};

Scratch CarveScratch(Arena &a, int32_t M, int32_t D, int32_t N, int32_t T, int32_t n, int32_t S)
{
  Scratch s{};
  s.mat_class = a.Take<int8_t>(M);
  s.dir = a.Take<int8_t>(N);
  s.code = a.Take<int8_t>(N);
  s.dnodes = a.Take<int32_t>(D);
  s.first_sel = a.Take<int32_t>(N);
  s.first_other = a.Take<int32_t>(N);
  s.count = a.Take<int32_t>(N);
  s.cursor = a.Take<int32_t>(N);
  s.class_bits = a.Take<int32_t>(n);
  s.keep = a.Take<int32_t>(T);
  s.pos = a.Take<int32_t>(static_cast<size_t>(T) + 1);
  s.found = a.Take<int32_t>(S);
  s.inc_off = a.Take<int32_t>(static_cast<size_t>(N) + 1);
  s.inc_elem = a.Take<int32_t>(3 * static_cast<size_t>(T));
  s.probe_elem = a.Take<int32_t>(S);
  s.flags = a.Take<uint32_t>(4);
  s.xy = a.Take<double>(2 * static_cast<size_t>(N));
  s.mask = a.Take<double>(N);
  s.info = a.Take<double>(3);
  s.px = a.Take<double>(S);
  s.py = a.Take<double>(S);
  s.pcos = a.Take<double>(S);
  s.psin = a.Take<double>(S);
  s.lambda = a.Take<double>(3 * static_cast<size_t>(S));
  s.wst.elem = a.Take<int32_t>(T);
  s.wst.hx = a.Take<double>(T);
  s.wst.hy = a.Take<double>(T);
  s.wst.w = a.Take<double>(T);
  s.wst.cx = a.Take<double>(T);
  s.wst.cy = a.Take<double>(T);
  return s;
}

bool Up(void *dst, const void *src, size_t bytes)
{
  return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
}

bool Copy(void *dst, const void *src, size_t bytes)
{
  return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess;
}

bool Launched() { return hipGetLastError() == hipSuccess; }

// The plan stages after the options check; `scratch` is freed by the caller.
int Build(const femm_model_t *m, const HostPlan &h, const DeviceBuffers &dev,
          const femm_postprocess_options_t &o, const std::vector<int8_t> &mat_class,
          bool probes_valid, HostPostPlan *host, DevicePostPlan *out, void **base,
          void **scratch, DeviceBuffers *mask_dev)
{
  const int32_t N = h.N, T = h.T, n = h.n, M = m->material_count, D = m->dirichlet_count;
  const int32_t S = probes_valid ? o.angle_count : 0;
  const DevicePlan &P = dev.plan;
  Arena sizing;
  CarveScratch(sizing, M, D, N, T, n, S);
  if (hipMalloc(scratch, sizing.offset) != hipSuccess) {
    *scratch = nullptr;
    return FEMM_GPU_ALLOCATION_FAILED;
  }
  if (hipMemset(*scratch, 0, sizing.offset) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  Arena arena{static_cast<char *>(*scratch), 0};
  const Scratch s = CarveScratch(arena, M, D, N, T, n, S);

  std::vector<double> px(S), py(S), pc(S), ps(S);
  const double R = o.airgap_radius_m;
  for (int32_t k = 0; k < S; ++k) {
    pc[k] = std::cos(o.angles_rad[k]);
    ps[k] = std::sin(o.angles_rad[k]);
    px[k] = R * pc[k];
    py[k] = R * ps[k];
  }
  if (!Up(s.mat_class, mat_class.data(), M) || !Up(s.dnodes, m->dirichlet_nodes, sizeof(int32_t) * D) ||
      !Up(s.xy, m->node_xy, sizeof(double) * 2 * N) || !Up(s.px, px.data(), sizeof(double) * S) ||
      !Up(s.py, py.data(), sizeof(double) * S) || !Up(s.pcos, pc.data(), sizeof(double) * S) ||
      !Up(s.psin, ps.data(), sizeof(double) * S))
    return FEMM_INTERNAL_ERROR;

  // Classes, incidence, edge rule, fixed values.
  hipLaunchKernelGGL(FillInt, dim3(Blocks(N)), dim3(kThreads), 0, nullptr, s.first_sel, N, INT_MAX);
  hipLaunchKernelGGL(FillInt, dim3(Blocks(N)), dim3(kThreads), 0, nullptr, s.first_other, N, INT_MAX);
  if (D > 0)
    hipLaunchKernelGGL(MarkDirichlet, dim3(Blocks(D)), dim3(kThreads), 0, nullptr, D, s.dnodes, N, s.dir);
  hipLaunchKernelGGL(ClassifyElements, dim3(Blocks(T)), dim3(kThreads), 0, nullptr, T, P.tri_nodes,
                     P.tri_mat, s.mat_class, s.first_sel, s.first_other, s.count, s.flags);
  hipLaunchKernelGGL(ExclusiveScan, dim3(1), dim3(kScanThreads), 0, nullptr, s.count, s.inc_off, N);
  hipLaunchKernelGGL(FillIncidence, dim3(Blocks(T)), dim3(kThreads), 0, nullptr, T, P.tri_nodes,
                     s.inc_off, s.cursor, s.inc_elem);
  hipLaunchKernelGGL(SortIncidence, dim3(Blocks(N)), dim3(kThreads), 0, nullptr, N, s.inc_off, s.inc_elem);
  hipLaunchKernelGGL(CheckEdges, dim3(Blocks(T)), dim3(kThreads), 0, nullptr, T, P.tri_nodes,
                     P.tri_mat, s.mat_class, s.inc_off, s.inc_elem, s.flags);
  hipLaunchKernelGGL(FixNodes, dim3(Blocks(N)), dim3(kThreads), 0, nullptr, N, s.dir, s.first_sel,
                     s.first_other, P.node_to_free, s.code, s.class_bits, s.flags);
  hipLaunchKernelGGL(CountFreeClasses, dim3(Blocks(n)), dim3(kThreads), 0, nullptr, n,
                     s.class_bits, s.flags);
  uint32_t flags[3] = {0, 0, 0};
  if (!Launched() || hipMemcpy(flags, s.flags, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if (!(flags[0] & kSeen) || (flags[0] & kEdgeShared))
    return FEMM_INVALID_MATERIAL;
  if (flags[0] & kFixedConflict)
    return FEMM_BOUNDARY_INVALID;

  // Mask: fixed values alone when every node has one, else the Laplace solve.
  const bool all_fixed = flags[1] == 0;
  if (!all_fixed) {
    if ((flags[0] & kClassConflict) || flags[2] == 0)
      return FEMM_BOUNDARY_INVALID;
    mask_dev->plan = P;
    int st = EnsureWork(h, 1, false, mask_dev);
    if (st != FEMM_OK)
      return st;
    hipLaunchKernelGGL(MaskSystem, dim3(Blocks(n)), dim3(kThreads), 0, nullptr, P, s.xy,
                       s.class_bits, mask_dev->work.vals, mask_dev->work.rhs, s.flags);
    if (!Launched())
      return FEMM_INTERNAL_ERROR;
    st = LaunchTeamPcg(*mask_dev, o.mask_relative_tolerance, o.max_mask_iterations, s.info,
                       nullptr);
    if (st != FEMM_OK)
      return st;
    double info[3] = {0.0, 0.0, 0.0};
    if (hipMemcpy(info, s.info, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(flags, s.flags, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
      return FEMM_INTERNAL_ERROR;
    if (flags[0] & kBadElement)
      return FEMM_MESH_INVALID;
    host->mask_iterations = static_cast<int32_t>(info[1]);
    if (static_cast<int>(info[0]) != FEMM_OK)
      return static_cast<int>(info[0]);
  }
  hipLaunchKernelGGL(ExpandMask, dim3(Blocks(N)), dim3(kThreads), 0, nullptr, N, s.code,
                     P.node_to_free, s.class_bits, all_fixed ? nullptr : mask_dev->work.x, s.mask);

  // Mask gradients, compacted in element order.
  const double depth = m->problem_type == FEMM_PLANAR ? m->depth_m : h.depth;
  hipLaunchKernelGGL(GradientFlags, dim3(Blocks(T)), dim3(kThreads), 0, nullptr, T, P.tri_nodes,
                     s.xy, s.mask, s.keep);
  hipLaunchKernelGGL(ExclusiveScan, dim3(1), dim3(kScanThreads), 0, nullptr, s.keep, s.pos, T);
  hipLaunchKernelGGL(WriteGradients, dim3(Blocks(T)), dim3(kThreads), 0, nullptr, T, P.tri_nodes,
                     s.xy, s.mask, s.keep, s.pos, depth, s.wst);
  int32_t nwst = 0;
  if (!Launched() ||
      hipMemcpy(&nwst, s.pos + T, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  host->mask.assign(N, 0.0);
  if (hipMemcpy(host->mask.data(), s.mask, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess)
    return FEMM_INTERNAL_ERROR;

  // Probes.
  if (!probes_valid)
    return FEMM_INVALID_ARGUMENT;
  if (S > 0) {
    const double r2 = R * R;
    const double tol = 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, r2);
    hipLaunchKernelGGL(FillInt, dim3(Blocks(S)), dim3(kThreads), 0, nullptr, s.found, S, INT_MAX);
    hipLaunchKernelGGL(LocateProbes, dim3(Blocks(T), std::min(S, kProbeRows)), dim3(kThreads), 0,
                       nullptr, T, S, P.tri_nodes, s.xy, r2, tol, s.px, s.py, s.found);
    hipLaunchKernelGGL(ProbeOutput, dim3(Blocks(S)), dim3(kThreads), 0, nullptr, S, P.tri_nodes,
                       P.tri_mat, s.mat_class, s.xy, s.px, s.py, s.found, s.probe_elem, s.lambda,
                       s.flags);
    if (!Launched() || hipMemcpy(flags, s.flags, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
      return FEMM_INTERNAL_ERROR;
    if (flags[0] & kProbeMissing)
      return FEMM_INVALID_ARGUMENT;
  }

  // The plan's own allocation, sized to what was found.
  Arena plan_sizing;
  auto carve = [&](Arena &a, DevicePostPlan &p) {
    p.wst_elem = a.Take<int32_t>(nwst);
    p.wst_hx = a.Take<double>(nwst);
    p.wst_hy = a.Take<double>(nwst);
    p.wst_w = a.Take<double>(nwst);
    p.wst_cx = a.Take<double>(nwst);
    p.wst_cy = a.Take<double>(nwst);
    p.probe_elem = a.Take<int32_t>(S);
    p.probe_lambda = a.Take<double>(3 * static_cast<size_t>(S));
    p.probe_cos = a.Take<double>(S);
    p.probe_sin = a.Take<double>(S);
    p.inc_off = a.Take<int32_t>(static_cast<size_t>(N) + 1);
    p.inc_elem = a.Take<int32_t>(3 * static_cast<size_t>(T));
    p.node_xy = a.Take<double>(2 * static_cast<size_t>(N));
  };
  DevicePostPlan p{};
  carve(plan_sizing, p);
  if (hipMalloc(base, plan_sizing.offset) != hipSuccess) {
    *base = nullptr;
    return FEMM_GPU_ALLOCATION_FAILED;
  }
  Arena plan_arena{static_cast<char *>(*base), 0};
  carve(plan_arena, p);
  const size_t w = sizeof(double) * nwst;
  if (!Copy(const_cast<int32_t *>(p.wst_elem), s.wst.elem, sizeof(int32_t) * nwst) ||
      !Copy(const_cast<double *>(p.wst_hx), s.wst.hx, w) ||
      !Copy(const_cast<double *>(p.wst_hy), s.wst.hy, w) ||
      !Copy(const_cast<double *>(p.wst_w), s.wst.w, w) ||
      !Copy(const_cast<double *>(p.wst_cx), s.wst.cx, w) ||
      !Copy(const_cast<double *>(p.wst_cy), s.wst.cy, w) ||
      !Copy(const_cast<int32_t *>(p.probe_elem), s.probe_elem, sizeof(int32_t) * S) ||
      !Copy(const_cast<double *>(p.probe_lambda), s.lambda, sizeof(double) * 3 * S) ||
      !Copy(const_cast<double *>(p.probe_cos), s.pcos, sizeof(double) * S) ||
      !Copy(const_cast<double *>(p.probe_sin), s.psin, sizeof(double) * S) ||
      !Copy(const_cast<int32_t *>(p.inc_off), s.inc_off, sizeof(int32_t) * (N + 1)) ||
      !Copy(const_cast<int32_t *>(p.inc_elem), s.inc_elem, sizeof(int32_t) * 3 * T) ||
      !Copy(const_cast<double *>(p.node_xy), s.xy, sizeof(double) * 2 * N) ||
      hipDeviceSynchronize() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  p.nwst = nwst;
  p.nprobe = S;
  p.N = P.N;
  p.T = P.T;
  p.Q = P.Q;
  p.tri_nodes = P.tri_nodes;
  p.tri_mat = P.tri_mat;
  p.tri_gx = P.tri_gx;
  p.tri_gy = P.tri_gy;
  *out = p;
  return FEMM_OK;
}

}  // namespace

int BuildPostPlanDevice(const femm_model_t *m, const HostPlan &h, const DeviceBuffers &dev,
                        const femm_postprocess_options_t &o, HostPostPlan *host,
                        DevicePostPlan *out, void **base)
{
  *host = HostPostPlan{};
  *out = DevicePostPlan{};
  *base = nullptr;
  auto done = [host](int st) { return host->status = st; };
  const int32_t M = m->material_count;
  if ((o.selected_group < 0 && (o.selected_label < 0 || o.selected_label >= M)) ||
      (o.air_group < 0 && (o.air_label < 0 || o.air_label >= M)) ||
      (o.selected_group >= 0 && o.selected_group == o.air_group) ||
      (o.selected_group < 0 && o.air_group < 0 && o.selected_label == o.air_label) ||
      !(o.mask_relative_tolerance > 0.0) || !std::isfinite(o.mask_relative_tolerance) ||
      o.max_mask_iterations <= 0 || o.angle_count < 0 || (o.angle_count > 0 && !o.angles_rad))
    return done(FEMM_INVALID_ARGUMENT);
  std::vector<int8_t> mat_class(M);
  for (int32_t k = 0; k < M; ++k) {
    const bool sel = o.selected_group >= 0 ? m->materials[k].group == o.selected_group
                                           : k == o.selected_label;
    const bool air = o.air_group >= 0 ? m->materials[k].group == o.air_group : k == o.air_label;
    mat_class[k] = sel ? kSelected : (air ? kAir : kOther);
  }
  // A bad radius or angle is reported after the mask stages, as the host plan build of the
  // reference reports it (field_postprocess.cpp:274-300 runs after the mask solve).
  bool probes_valid = o.airgap_radius_m >= 0.0 && std::isfinite(o.airgap_radius_m);
  for (int32_t k = 0; probes_valid && k < o.angle_count; ++k)
    probes_valid = std::isfinite(o.angles_rad[k]);
  void *scratch = nullptr;
  DeviceBuffers mask_dev;
  int st = Build(m, h, dev, o, mat_class, probes_valid, host, out, base, &scratch, &mask_dev);
  (void)hipDeviceSynchronize();
  mask_dev.base = nullptr;   // the pattern belongs to the solve plan
  ReleaseDevice(&mask_dev);
  if (scratch)
    (void)hipFree(scratch);
  if (st != FEMM_OK) {
    if (*base)
      (void)hipFree(*base);
    *base = nullptr;
    *out = DevicePostPlan{};
    host->mask.clear();
  }
  return done(st);
}

}  // namespace femm
