// HIP kernels of the MI355X nonlinear magnetostatic solve path (gfx950, wave64):
// the point-Jacobi solve forms (FP64 and mixed precision), assembly / PCG test kernels,
// field and flux kernels, the weighted-stress postprocess, and the host-side plan
// upload, workspaces and launch dispatch.  The block-Jacobi forms are compiled in
// device_block.hip from the same templates (solve_kernels.hpp).

#include "solve_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>
#include <vector>

namespace femm
{
namespace
{

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

// `reps` products y = A x per item with the operator values in the workspace (the last
// assembled Newton operator), the same pipelined SELL-64 SpMV as the PCG: the product on
// its own, for measuring it against HBM bandwidth.
template <typename T>
__global__ __launch_bounds__(kBlockThreads) void SpmvKernel(DevicePlan P, DeviceWork W, int reps,
                                                            const T *x, T *y)
{
  const int item = blockIdx.x;
  const T *vals = nullptr;
  if constexpr (sizeof(T) == 8)
    vals = W.vals + static_cast<size_t>(item) * P.sell_size;
  else
    vals = W.vals32 + static_cast<size_t>(item) * P.sell_size;
  const size_t n = P.n;
  for (int rep = 0; rep < reps; ++rep) {
    // The PCG's own product: two slices per step in FP32, one in FP64.
    if constexpr (sizeof(T) == 4)
      Spmv2Fit<T, T>(P, WholeItem(P), vals, x + item * n, y + item * n,
                  static_cast<const T *>(nullptr));
    else
      SpmvFit<T, T>(P, WholeItem(P), vals, x + item * n, y + item * n,
                       static_cast<const T *>(nullptr));
    __syncthreads();
  }
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
  const BlockPrec<double> B{nullptr, 1, P.n};
  const LinearInfo li =
      Pcg64(P, w.vals, w.inv, B, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit, false, R);
  if (threadIdx.x == 0) {
    info[3 * item] = li.status;
    info[3 * item + 1] = li.iterations;
    info[3 * item + 2] = li.residual;
  }
}

// One PCG (as PcgKernel, item 0) split over `partitions` co-resident workgroups of a
// cooperative launch, for one-off solves on large meshes (the weighted-stress mask).
__global__ __launch_bounds__(kBlockThreads) void TeamPcgKernel(DevicePlan P, DeviceWork W,
                                                               double tol, int maxit,
                                                               double *info, int partitions)
{
  __shared__ double lds[2 * kWaves * 4];
  const int g = UniformInt(static_cast<int>(blockIdx.x));
  GridTeam R{BlockReducer{lds, 0}, Partition(P, partitions, g), W.team_partials, W.team_barrier,
             W.team_abort, partitions, g, 0};
  const ItemPointers w = ItemWork(P, W, 0, false);
  for (int row = R.S.row0 + threadIdx.x; row < R.S.row1; row += kBlockThreads) {
    const double d = w.vals[P.diag_sell[row]];
    w.inv[row] = (isfinite(d) && d > 0.0) ? 1.0 / d : 0.0;
  }
  R.Sync();
  const BlockPrec<double> B{nullptr, 1, P.n};
  const LinearInfo li = Pcg64<true, 0>(P, w.vals, w.inv, B, w.rhs, w.x, w.r, w.p, w.ap, tol,
                                       maxit, false, R, w.p_alt, w.q_buf);
  if (g == 0 && threadIdx.x == 0) {
    info[0] = li.status;
    info[1] = li.iterations;
    info[2] = li.residual;
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

// Per item: the air-gap element postprocess (ComputeAirGapElementPostprocess,
// postprocess/airgap_postprocess.cpp:41-160).  Each cell's centre field comes from its
// four virtual corners (Catmull-Rom interpolants of the rings at the element's shifts):
// Br = d(A)/ds across the arc and Bt = -dA/dr across the gap; the Maxwell stress on the
// mean circle gives the force and torque; the first element's Br per cell (kept in LDS)
// gives the discrete Fourier series evaluated at the probe angles.
__global__ __launch_bounds__(kBlockThreads) void AgePostprocessKernel(DevicePostPlan P,
                                                                      const double *states,
                                                                      double *loads,
                                                                      double *radial)
{
  extern __shared__ double age_lds[];   // [cells0] Br, [H] cos, [H] sin coefficients
  __shared__ double lds[2 * kWaves * 4];
  BlockReducer R{lds, 0};
  const int item = blockIdx.x;
  const double *a = states + static_cast<size_t>(item) * P.N;
  constexpr double kMu0 = 4.0e-7 * 3.141592653589793238462643383279502884;
  double s[3] = {0.0, 0.0, 0.0};
  for (int c = threadIdx.x; c < P.age_cells; c += kBlockThreads) {
    const int g = P.age_cell_element[c];
    const int k = c - P.age_first[g];
    const double *par = P.age_par + static_cast<size_t>(g) * kAgeParams;
    double v[10];
    for (int q = 0; q < 10; ++q)
      v[q] = a[P.age_nodes[10 * c + q]] * P.age_w[10 * c + q];
    double in0 = 0.0, in1 = 0.0, out0 = 0.0, out1 = 0.0;
    for (int q = 0; q < 4; ++q) {
      in0 += par[q] * v[q];
      in1 += par[q] * v[q + 1];
      out0 += par[4 + q] * v[q + 5];
      out1 += par[4 + q] * v[q + 6];
    }
    const double dt = par[8], radius = par[9], dr = par[10];
    const double br = ((in1 - in0) + (out1 - out0)) / (2.0 * dt * radius);
    const double bt = ((in0 + in1) - (out0 + out1)) / (2.0 * dr);
    const double theta = (static_cast<double>(k) + 0.5) * dt;
    const double normal = br * br - bt * bt, shear = 2.0 * br * bt;
    s[0] += par[11] * (normal * cos(theta) - shear * sin(theta));
    s[1] += par[11] * (normal * sin(theta) + shear * cos(theta));
    s[2] += par[12] * br * bt / kMu0;
    if (g == 0)
      age_lds[k] = br;
  }
  R.Sum(s);   // ends with a barrier: the first element's Br is complete
  if (threadIdx.x == 0) {
    loads[3 * item] = s[0];
    loads[3 * item + 1] = s[1];
    loads[3 * item + 2] = s[2];
  }
  const int cells0 = P.age_cells0, H = P.age_harmonics;
  const double dt0 = P.age_par[8];
  for (int h = threadIdx.x; h < H; h += kBlockThreads) {
    const double order = static_cast<double>(P.age_order[h]);
    double cs = 0.0, sn = 0.0;
    for (int k = 0; k < cells0; ++k) {
      const double phase = (static_cast<double>(k) + 0.5) * dt0 * order;
      cs += age_lds[k] * cos(phase);
      sn += age_lds[k] * sin(phase);
    }
    age_lds[cells0 + h] = cs / P.age_norm[h];
    age_lds[cells0 + H + h] = sn / P.age_norm[h];
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P.nprobe; p += kBlockThreads) {
    const double angle = P.probe_angle[p];
    double value = 0.0;
    for (int h = 0; h < H; ++h) {
      const double phase = static_cast<double>(P.age_order[h]) * angle;
      value += age_lds[cells0 + h] * cos(phase) + age_lds[cells0 + H + h] * sin(phase);
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
    P.age_vals = h.age_vals.empty() ? nullptr : up(h.age_vals);
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
  P.max_slice_width = 0;
  for (const int32_t w : h.slice_width)
    P.max_slice_width = std::max(P.max_slice_width, w);
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

size_t WorkBytes(const HostPlan &h, int32_t capacity, bool mixed, bool blocks)
{
  const size_t ci = static_cast<size_t>(capacity);
  const size_t nv = static_cast<size_t>(VectorStride(h.n));   // per-item vector stride
  size_t bytes = Bytes<double>(ci * h.T * kElemStride) + Bytes<double>(ci * h.sell_size) +
                 8 * Bytes<double>(ci * nv) + Bytes<double>(ci * h.N);
  if (mixed)
    bytes += Bytes<float>(ci * h.sell_size) + 7 * Bytes<float>(ci * nv);
  if (blocks)
    bytes += Bytes<double>(ci * kMaxBlock * nv) + (mixed ? Bytes<float>(ci * kMaxBlock * nv) : 0);
  bytes += Bytes<double>(ci * 2 * kMaxPartitions * 4) + Bytes<uint32_t>(ci * kBarrierWords) +
           Bytes<int32_t>(ci) + Bytes<int32_t>(ci + 1);
  return bytes;
}

DeviceWork CarveWork(const HostPlan &h, int32_t capacity, bool mixed, bool blocks, void *base)
{
  const size_t ci = static_cast<size_t>(capacity);
  const size_t nv = static_cast<size_t>(VectorStride(h.n));
  char *cur = static_cast<char *>(base);
  DeviceWork w{};
  w.capacity = capacity;
  w.elem = Carve<double>(cur, ci * h.T * kElemStride);
  w.vals = Carve<double>(cur, ci * h.sell_size);
  w.rhs = Carve<double>(cur, ci * nv);
  w.inv = Carve<double>(cur, ci * nv);
  w.x = Carve<double>(cur, ci * nv);
  w.r = Carve<double>(cur, ci * nv);
  w.p = Carve<double>(cur, ci * nv);
  w.ap = Carve<double>(cur, ci * nv);
  w.p_alt = Carve<double>(cur, ci * nv);
  w.q_buf = Carve<double>(cur, ci * nv);
  w.trial = Carve<double>(cur, ci * h.N);
  if (mixed) {
    w.vals32 = Carve<float>(cur, ci * h.sell_size);
    w.r32 = Carve<float>(cur, ci * nv);
    w.p32 = Carve<float>(cur, ci * nv);
    w.ap32 = Carve<float>(cur, ci * nv);
    w.d32 = Carve<float>(cur, ci * nv);
    w.inv32 = Carve<float>(cur, ci * nv);
    w.p32_alt = Carve<float>(cur, ci * nv);
    w.q32_buf = Carve<float>(cur, ci * nv);
  }
  if (blocks) {
    w.bjac = Carve<double>(cur, ci * kMaxBlock * nv);
    if (mixed)
      w.bjac32 = Carve<float>(cur, ci * kMaxBlock * nv);
  }
  w.team_partials = Carve<double>(cur, ci * 2 * kMaxPartitions * 4);
  w.team_barrier = Carve<uint32_t>(cur, ci * kBarrierWords);
  w.team_abort = Carve<int32_t>(cur, ci);
  w.cont = Carve<int32_t>(cur, ci + 1);
  return w;
}

int EnsureWork(const HostPlan &h, int32_t items, bool mixed, DeviceBuffers *dev, bool blocks)
{
  const bool have_mixed = dev->work.vals32 != nullptr;
  const bool have_blocks = dev->work.bjac != nullptr;
  if (dev->work_base && dev->work.capacity >= items && (!mixed || have_mixed) &&
      (!blocks || have_blocks))
    return FEMM_OK;
  const int32_t cap = std::max(items, dev->work.capacity);
  const bool want_mixed = mixed || have_mixed;
  const bool want_blocks = blocks || have_blocks;
  if (dev->work_base) {
    if (hipDeviceSynchronize() != hipSuccess)
      return FEMM_INTERNAL_ERROR;
    (void)hipFree(dev->work_base);
    dev->work_base = nullptr;
  }
  const size_t bytes = WorkBytes(h, cap, want_mixed, want_blocks);
  void *base = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess)
    return FEMM_GPU_ALLOCATION_FAILED;
  // SELL padding entries must read as zero; they are never written afterwards.
  if (hipMemset(base, 0, bytes) != hipSuccess) {
    (void)hipFree(base);
    return FEMM_INTERNAL_ERROR;
  }
  dev->work = CarveWork(h, cap, want_mixed, want_blocks, base);
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

int FormOf(const DeviceBuffers &dev, const femm_solve_config_t &config)
{
  return (config.precision == FEMM_PRECISION_MIXED ? kFormMixed : 0) |
         (config.block_size > 1 && dev.work.bjac != nullptr ? kFormBlock : 0);
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
  if (args.config.block_size > 1 &&
      (dev.work.bjac == nullptr || (args.config.precision == FEMM_PRECISION_MIXED &&
                                    dev.work.bjac32 == nullptr)))
    return FEMM_INVALID_ARGUMENT;   // workspace without block-Jacobi planes
  const int form = FormOf(dev, args.config);
  if (form & kFormBlock)
    return LaunchBlockSolve(form, kMode, dev, items, partitions, currents, args, warm, a, flux,
                            results, stream, x);
  if (form & kFormMixed)
    return LaunchFormSolve<kFormMixed, kMode>(dev, items, partitions, currents, args, warm, a,
                                              flux, results, stream, x);
  return LaunchFormSolve<0, kMode>(dev, items, partitions, currents, args, warm, a, flux, results,
                                   stream, x);
}

const void *GroupKernel(int form, bool recovery)
{
  if (form & kFormBlock)
    return BlockGroupKernel(form, recovery);
  return (form & kFormMixed) ? FormGroupKernel<kFormMixed>(recovery) : FormGroupKernel<0>(recovery);
}

}  // namespace

int GroupCapacity(bool mixed)
{
  int cap = 1 << 30;
  for (int form = 0; form < 4; ++form)
    if (((form & kFormMixed) != 0) == mixed)
      cap = std::min({cap, CoResidentBlocks(GroupKernel(form, false)),
                      CoResidentBlocks(GroupKernel(form, true))});
  return cap;
}

int LaunchGroup(const GroupEntry *entries_device, const std::vector<GroupEntry> &entries,
                const femm_newton_options_t &options, const femm_solve_config_t &config,
                bool safeguarded, bool prior_results, hipStream_t stream)
{
  if (entries.empty())
    return FEMM_INVALID_ARGUMENT;
  const bool mixed = config.precision == FEMM_PRECISION_MIXED;
  const bool block = config.block_size > 1;
  for (const GroupEntry &E : entries)
    if (block && (E.work.bjac == nullptr || (mixed && E.work.bjac32 == nullptr)))
      return FEMM_INVALID_ARGUMENT;
  const void *kernel = GroupKernel((mixed ? kFormMixed : 0) | (block ? kFormBlock : 0), safeguarded);
  long grid = 0;
  for (const GroupEntry &E : entries) {
    if (E.items <= 0 || E.partitions <= 0 || E.partitions > kMaxPartitions ||
        E.items > E.work.capacity || E.first_block != grid)
      return FEMM_INVALID_ARGUMENT;
    grid += static_cast<long>(E.items) * E.partitions;
  }
  if (grid > CoResidentBlocks(kernel))
    return FEMM_INVALID_ARGUMENT;
  const int count = static_cast<int>(entries.size());
  ClearGroupKernel<<<count, 256, 0, stream>>>(entries_device, count);
  if (hipGetLastError() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  const GroupEntry *table = entries_device;
  int entry_count = count;
  NewtonArgs args{options, config, prior_results ? kPriorRedoUnconverged : kPriorNone, options};
  int prior = prior_results ? 1 : 0;
  void *params[] = {&table, &entry_count, &args, &prior};
  return hipLaunchCooperativeKernel(kernel, dim3(static_cast<unsigned>(grid)), dim3(kBlockThreads),
                                    params, 0, stream) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
}

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

int LaunchSpmv(const DeviceBuffers &dev, int32_t items, bool fp32, int32_t reps, const void *x,
               void *y, hipStream_t stream)
{
  if (items <= 0 || items > dev.work.capacity || reps < 0 || (fp32 && dev.work.vals32 == nullptr))
    return FEMM_INVALID_ARGUMENT;
  if (fp32)
    hipLaunchKernelGGL(SpmvKernel<float>, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan,
                       dev.work, reps, static_cast<const float *>(x), static_cast<float *>(y));
  else
    hipLaunchKernelGGL(SpmvKernel<double>, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan,
                       dev.work, reps, static_cast<const double *>(x), static_cast<double *>(y));
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

int LaunchPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit, double *info,
              hipStream_t stream, int32_t block_size)
{
  if (items <= 0 || items > dev.work.capacity)
    return FEMM_INVALID_ARGUMENT;
  if (block_size > 1) {
    if (dev.work.bjac == nullptr || block_size > kMaxBlock || (block_size & (block_size - 1)) != 0)
      return FEMM_INVALID_ARGUMENT;
    return LaunchBlockPcg(dev, items, tol, maxit, block_size, info, stream);
  }
  hipLaunchKernelGGL(PcgKernel, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan, dev.work,
                     tol, maxit, info);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

int LaunchTeamPcg(const DeviceBuffers &dev, double tol, int32_t maxit, double *info,
                  hipStream_t stream)
{
  if (dev.work.capacity < 1)
    return FEMM_INVALID_ARGUMENT;
  // About 32 slices (2,048 rows) per workgroup, as the untuned team split.
  const void *kernel = reinterpret_cast<const void *>(&TeamPcgKernel);
  const int parts = std::max(1, std::min({dev.plan.nslices / 32, kMaxPartitions,
                                          CoResidentBlocks(kernel)}));
  ClearTeamKernel<<<1, 256, 0, stream>>>(dev.work.team_barrier, dev.work.team_abort, 1);
  if (hipGetLastError() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  DevicePlan plan = dev.plan;
  DeviceWork work = dev.work;
  int partitions = parts;
  void *params[] = {&plan, &work, &tol, &maxit, &info, &partitions};
  return hipLaunchCooperativeKernel(kernel, dim3(static_cast<unsigned>(parts)),
                                    dim3(kBlockThreads), params, 0, stream) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
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
  if (post.age_mode) {
    const size_t lds = sizeof(double) * (static_cast<size_t>(post.age_cells0) +
                                         2 * static_cast<size_t>(post.age_harmonics));
    if (lds > kAgeLdsLimit)
      return FEMM_UNSUPPORTED_FEATURE;
    hipLaunchKernelGGL(AgePostprocessKernel, dim3(items), dim3(kBlockThreads), lds, stream, post,
                       states, loads, radial_b);
  } else {
    hipLaunchKernelGGL(PostprocessKernel, dim3(items), dim3(kBlockThreads), 0, stream, post,
                       states, loads, radial_b);
  }
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm
