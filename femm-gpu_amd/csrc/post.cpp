// Weighted-stress (WST) force/torque and smoothed air-gap B: host plan and mask solve.

#include "post.hpp"

#include "age.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace femm
{
namespace
{

struct Selector {
  const femm_model_t *m;
  int32_t label, group;
  bool operator()(int32_t mat) const
  {
    return group >= 0 ? m->materials[mat].group == group : mat == label;
  }
};

double Det(const femm_model_t *m, int32_t e)
{
  const int32_t *t = m->triangle_nodes + 3 * e;
  const double *p0 = m->node_xy + 2 * t[0], *p1 = m->node_xy + 2 * t[1], *p2 = m->node_xy + 2 * t[2];
  return (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]);
}

// Mask Laplace solve: the Laplacian weighted by sqrt(area) over the mesh, fixed
// values on the selected body (1), other non-air regions and Dirichlet nodes (0),
// periodic pairs folded with sign +1 (weighted_stress.cu:102-151).  Assembled and
// solved with this library's own device assembly and PCG on a second plan whose
// per-element reluctivity is sqrt(area).
int SolveMask(const femm_model_t *m, const std::vector<double> &fixed, int max_iterations,
              double tolerance, std::vector<double> *mask, int32_t *iterations)
{
  const int32_t N = m->node_count, T = m->triangle_count;
  std::vector<femm_material_t> mats(T);
  std::vector<int32_t> tri_mat(T);
  for (int32_t e = 0; e < T; ++e) {
    femm_material_t mt{};
    mt.reluctivity_zero = std::sqrt(0.5 * Det(m, e));
    mt.bh_curve = -1;
    mt.circuit = -1;
    mt.group = -1;
    mats[e] = mt;
    tri_mat[e] = e;
  }
  std::vector<int32_t> dn;
  std::vector<double> dv;
  for (int32_t v = 0; v < N; ++v)
    if (std::isfinite(fixed[v])) {
      dn.push_back(v);
      dv.push_back(fixed[v]);
    }
  std::vector<femm_node_constraint_t> cons(m->constraints, m->constraints + m->constraint_count);
  for (auto &c : cons)
    c.sign = 1;
  femm_model_t mm{};
  mm.problem_type = FEMM_PLANAR;
  mm.node_count = N;
  mm.node_xy = m->node_xy;
  mm.triangle_count = T;
  mm.triangle_nodes = m->triangle_nodes;
  mm.triangle_material = tri_mat.data();
  mm.material_count = T;
  mm.materials = mats.data();
  mm.curve_count = 0;
  std::vector<int32_t> no_curves(1, 0);
  mm.curve_offsets = no_curves.data();
  mm.dirichlet_count = static_cast<int32_t>(dn.size());
  mm.dirichlet_nodes = dn.data();
  mm.dirichlet_values = dv.data();
  mm.constraint_count = static_cast<int32_t>(cons.size());
  mm.constraints = cons.data();
  mm.circuit_count = 0;
  mm.depth_m = 1.0;
  HostPlan hp;
  int st = BuildHostPlan(&mm, &hp, nullptr);
  if (st != FEMM_OK)
    return st;
  DeviceBuffers db;
  st = UploadPlan(hp, &db);
  if (st == FEMM_OK)
    st = EnsureWork(hp, 1, false, &db);
  void *scratch = nullptr;
  const size_t bytes = sizeof(double) * (N + 4) + 64;
  if (st == FEMM_OK && hipMalloc(&scratch, bytes) != hipSuccess)
    st = FEMM_GPU_ALLOCATION_FAILED;
  std::vector<double> x(hp.n), info(3, 0.0);
  if (st == FEMM_OK) {
    double *d_state = static_cast<double *>(scratch);
    double *d_info = d_state + N;
    int32_t *d_valid = reinterpret_cast<int32_t *>(d_state + N + 3);
    if (hipMemcpy(d_state, hp.boundary.data(), sizeof(double) * N, hipMemcpyHostToDevice) !=
        hipSuccess)
      st = FEMM_INTERNAL_ERROR;
    if (st == FEMM_OK)
      st = LaunchAssemble(db, 1, d_state, d_state, 0, d_valid, nullptr);
    if (st == FEMM_OK)
      st = LaunchPcg(db, 1, tolerance, max_iterations, d_info, nullptr);
    if (st == FEMM_OK &&
        (hipDeviceSynchronize() != hipSuccess ||
         hipMemcpy(x.data(), db.work.x, sizeof(double) * hp.n, hipMemcpyDeviceToHost) !=
             hipSuccess ||
         hipMemcpy(info.data(), d_info, sizeof(double) * 3, hipMemcpyDeviceToHost) != hipSuccess))
      st = FEMM_INTERNAL_ERROR;
  }
  if (scratch)
    (void)hipFree(scratch);
  ReleaseDevice(&db);
  if (st != FEMM_OK)
    return st;
  *iterations = static_cast<int32_t>(info[1]);
  if (static_cast<int>(info[0]) != FEMM_OK)
    return static_cast<int>(info[0]);
  mask->assign(N, 0.0);
  for (int32_t v = 0; v < N; ++v) {
    const int32_t f = hp.node_to_free[v];
    const double value = f >= 0 ? static_cast<double>(hp.node_sign[v]) * x[f] : hp.boundary[v];
    (*mask)[v] = value > 0.5 ? 1.0 : 0.0;   // FEMM WeightingScheme=0 threshold
  }
  return FEMM_OK;
}

}  // namespace

int BuildPostPlan(const femm_model_t *m, const HostPlan &plan,
                  const femm_postprocess_options_t &o, HostPostPlan *out)
{
  *out = HostPostPlan{};
  auto done = [out](int st) { return out->status = st; };
  const int32_t N = m->node_count, T = m->triangle_count, M = m->material_count;
  if ((o.selected_group < 0 && (o.selected_label < 0 || o.selected_label >= M)) ||
      (o.air_group < 0 && (o.air_label < 0 || o.air_label >= M)) ||
      (o.selected_group >= 0 && o.selected_group == o.air_group) ||
      (o.selected_group < 0 && o.air_group < 0 && o.selected_label == o.air_label) ||
      !(o.mask_relative_tolerance > 0.0) || !std::isfinite(o.mask_relative_tolerance) ||
      o.max_mask_iterations <= 0 || o.angle_count < 0 || (o.angle_count > 0 && !o.angles_rad))
    return done(FEMM_INVALID_ARGUMENT);
  const Selector selected{m, o.selected_label, o.selected_group};
  const Selector air{m, o.air_label, o.air_group};
  // Edges shared by a selected element and a non-air, non-selected element are
  // rejected, as FEMM's MakeMask does (weighted_stress.cu:49-78).
  std::vector<std::pair<uint64_t, int32_t>> edges;
  edges.reserve(static_cast<size_t>(T) * 3);
  bool seen = false;
  for (int32_t e = 0; e < T; ++e) {
    seen = seen || selected(m->triangle_material[e]);
    for (int k = 0; k < 3; ++k) {
      const int32_t a = m->triangle_nodes[3 * e + k], b = m->triangle_nodes[3 * e + (k + 1) % 3];
      edges.push_back({(static_cast<uint64_t>(std::min(a, b)) << 32) | static_cast<uint32_t>(std::max(a, b)), e});
    }
  }
  if (!seen)
    return done(FEMM_INVALID_MATERIAL);
  std::sort(edges.begin(), edges.end());
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    bool sel = false, other = false;
    for (; j < edges.size() && edges[j].first == edges[i].first; ++j) {
      const int32_t mat = m->triangle_material[edges[j].second];
      sel = sel || selected(mat);
      other = other || (!selected(mat) && !air(mat));
    }
    if (sel && other)
      return done(FEMM_INVALID_MATERIAL);
    i = j;
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> fixed(N, nan);
  for (int i = 0; i < m->dirichlet_count; ++i)
    fixed[m->dirichlet_nodes[i]] = 0.0;
  for (int32_t e = 0; e < T; ++e) {
    const int32_t mat = m->triangle_material[e];
    const int32_t *t = m->triangle_nodes + 3 * e;
    if (selected(mat)) {
      for (int k = 0; k < 3; ++k) {
        if (std::isfinite(fixed[t[k]]) && fixed[t[k]] != 1.0)
          return done(FEMM_BOUNDARY_INVALID);
        fixed[t[k]] = 1.0;
      }
    } else if (!air(mat)) {
      for (int k = 0; k < 3; ++k)
        if (!std::isfinite(fixed[t[k]]))
          fixed[t[k]] = 0.0;
    }
  }
  const bool all_fixed =
      std::all_of(fixed.begin(), fixed.end(), [](double v) { return std::isfinite(v); });
  if (all_fixed) {
    out->mask.resize(N);
    for (int32_t v = 0; v < N; ++v)
      out->mask[v] = fixed[v] > 0.5 ? 1.0 : 0.0;
  } else {
    const int st = SolveMask(m, fixed, o.max_mask_iterations, o.mask_relative_tolerance,
                             &out->mask, &out->mask_iterations);
    if (st != FEMM_OK)
      return done(st);
  }
  // Per-element mask gradient h = (sum mask b_y, -sum mask b_x), weight area * depth,
  // centroid (field_postprocess.cpp:248-269); only nonzero gradients can load the body.
  const double depth = m->problem_type == FEMM_PLANAR ? m->depth_m : plan.depth;
  out->inc_off.assign(N + 1, 0);
  for (int32_t e = 0; e < T; ++e) {
    const int32_t *t = m->triangle_nodes + 3 * e;
    for (int k = 0; k < 3; ++k)
      out->inc_off[t[k] + 1]++;
    const double det = Det(m, e);
    const double *p0 = m->node_xy + 2 * t[0], *p1 = m->node_xy + 2 * t[1],
                 *p2 = m->node_xy + 2 * t[2];
    const double dx[3] = {p1[1] - p2[1], p2[1] - p0[1], p0[1] - p1[1]};
    const double dy[3] = {p2[0] - p1[0], p0[0] - p2[0], p1[0] - p0[0]};
    double hx = 0.0, hy = 0.0;
    for (int k = 0; k < 3; ++k) {
      hx += out->mask[t[k]] * (-dx[k] / det);
      hy -= out->mask[t[k]] * (dy[k] / det);
    }
    if (hx != 0.0 || hy != 0.0) {
      out->wst_elem.push_back(e);
      out->wst_hx.push_back(hx);
      out->wst_hy.push_back(hy);
      out->wst_w.push_back(0.5 * det * depth);
      out->wst_cx.push_back((p0[0] + p1[0] + p2[0]) / 3.0);
      out->wst_cy.push_back((p0[1] + p1[1] + p2[1]) / 3.0);
    }
  }
  for (int32_t v = 0; v < N; ++v)
    out->inc_off[v + 1] += out->inc_off[v];
  out->inc_elem.assign(out->inc_off[N], 0);
  std::vector<int32_t> fill(out->inc_off.begin(), out->inc_off.end() - 1);
  for (int32_t e = 0; e < T; ++e)
    for (int k = 0; k < 3; ++k)
      out->inc_elem[fill[m->triangle_nodes[3 * e + k]]++] = e;
  out->node_xy.assign(m->node_xy, m->node_xy + 2 * static_cast<size_t>(N));
  // Probes: first containing element in element order (field_postprocess.cpp:274-336).
  const double R = o.airgap_radius_m;
  if (!(R >= 0.0) || !std::isfinite(R))
    return done(FEMM_INVALID_ARGUMENT);
  const double r2 = R * R;
  const double tol = 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, r2);
  std::vector<int32_t> candidates;
  for (int32_t e = 0; e < T; ++e) {
    const int32_t *t = m->triangle_nodes + 3 * e;
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 2; ++d) {
        lo[d] = std::min(lo[d], m->node_xy[2 * t[k] + d]);
        hi[d] = std::max(hi[d], m->node_xy[2 * t[k] + d]);
      }
    const double cx = lo[0] > 0.0 ? lo[0] : (hi[0] < 0.0 ? hi[0] : 0.0);
    const double cy = lo[1] > 0.0 ? lo[1] : (hi[1] < 0.0 ? hi[1] : 0.0);
    const double fx = std::max(std::abs(lo[0]), std::abs(hi[0]));
    const double fy = std::max(std::abs(lo[1]), std::abs(hi[1]));
    if (r2 + tol >= cx * cx + cy * cy && r2 <= fx * fx + fy * fy + tol)
      candidates.push_back(e);
  }
  for (int32_t s = 0; s < o.angle_count; ++s) {
    const double ang = o.angles_rad[s];
    if (!std::isfinite(ang))
      return done(FEMM_INVALID_ARGUMENT);
    const double c = std::cos(ang), sn = std::sin(ang), x = R * c, y = R * sn;
    int32_t found = -1;
    double lam[3] = {};
    for (int32_t e : candidates) {
      const int32_t *t = m->triangle_nodes + 3 * e;
      const double *p0 = m->node_xy + 2 * t[0], *p1 = m->node_xy + 2 * t[1],
                   *p2 = m->node_xy + 2 * t[2];
      const double det = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]);
      if (!(det > 0.0))
        continue;
      const double l1 = ((x - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (y - p0[1])) / det;
      const double l2 = ((p1[0] - p0[0]) * (y - p0[1]) - (x - p0[0]) * (p1[1] - p0[1])) / det;
      const double l0 = 1.0 - l1 - l2;
      if (std::isfinite(l0) && std::isfinite(l1) && std::isfinite(l2) && l0 >= -1e-12 &&
          l1 >= -1e-12 && l2 >= -1e-12) {
        found = e;
        lam[0] = l0;
        lam[1] = l1;
        lam[2] = l2;
        break;
      }
    }
    if (found < 0 || !air(m->triangle_material[found]))
      return done(FEMM_INVALID_ARGUMENT);
    out->probe_elem.push_back(found);
    out->probe_lambda.insert(out->probe_lambda.end(), lam, lam + 3);
    out->probe_cos.push_back(c);
    out->probe_sin.push_back(sn);
  }
  return done(FEMM_OK);
}

int UploadPostPlan(const HostPostPlan &h, const DeviceBuffers &dev, DevicePostPlan *out,
                   void **base)
{
  DevicePostPlan &p = *out;
  *base = nullptr;
  for (int pass = 0; pass < 2; ++pass) {
    char *mem = static_cast<char *>(*base);
    size_t offset = 0;
    bool ok = true;
    auto up = [&](const auto &vec) {
      using T = typename std::decay_t<decltype(vec)>::value_type;
      T *ptr = reinterpret_cast<T *>(mem + offset);
      offset += (std::max<size_t>(vec.size(), 1) * sizeof(T) + 255) / 256 * 256;
      if (pass == 1 && !vec.empty() &&
          hipMemcpy(ptr, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        ok = false;
      return static_cast<const T *>(ptr);
    };
    p.wst_elem = up(h.wst_elem);
    p.wst_hx = up(h.wst_hx);
    p.wst_hy = up(h.wst_hy);
    p.wst_w = up(h.wst_w);
    p.wst_cx = up(h.wst_cx);
    p.wst_cy = up(h.wst_cy);
    p.probe_elem = up(h.probe_elem);
    p.probe_lambda = up(h.probe_lambda);
    p.probe_cos = up(h.probe_cos);
    p.probe_sin = up(h.probe_sin);
    p.inc_off = up(h.inc_off);
    p.inc_elem = up(h.inc_elem);
    p.node_xy = up(h.node_xy);
    p.age_nodes = up(h.age_nodes);
    p.age_cell_element = up(h.age_cell_element);
    p.age_first = up(h.age_first);
    p.age_order = up(h.age_order);
    p.age_w = up(h.age_w);
    p.age_par = up(h.age_par);
    p.age_norm = up(h.age_norm);
    p.probe_angle = up(h.probe_angle);
    if (pass == 0) {
      void *m = nullptr;
      if (hipMalloc(&m, offset) != hipSuccess)
        return FEMM_GPU_ALLOCATION_FAILED;
      *base = m;
    } else if (!ok) {
      return FEMM_INTERNAL_ERROR;
    }
  }
  p.nwst = static_cast<int32_t>(h.wst_elem.size());
  p.nprobe = static_cast<int32_t>(h.age_mode ? h.probe_angle.size() : h.probe_elem.size());
  p.age_mode = h.age_mode;
  p.age_cells = static_cast<int32_t>(h.age_cell_element.size());
  p.age_cells0 = h.age_first.size() > 1 ? h.age_first[1] - h.age_first[0] : 0;
  p.age_harmonics = static_cast<int32_t>(h.age_order.size());
  p.N = dev.plan.N;
  p.T = dev.plan.T;
  p.Q = dev.plan.Q;
  p.tri_nodes = dev.plan.tri_nodes;
  p.tri_mat = dev.plan.tri_mat;
  p.tri_gx = dev.plan.tri_gx;
  p.tri_gy = dev.plan.tri_gy;
  return hipDeviceSynchronize() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

int BuildAgePostPlan(const HostPlan &plan, double depth_m, const femm_postprocess_options_t &o,
                     HostPostPlan *out)
{
  *out = HostPostPlan{};
  auto done = [out](int st) { return out->status = st; };
  if (plan.ages.empty() || o.angle_count < 0 || (o.angle_count > 0 && !o.angles_rad))
    return done(FEMM_INVALID_ARGUMENT);
  for (int32_t k = 0; k < o.angle_count; ++k)
    if (!std::isfinite(o.angles_rad[k]))
      return done(FEMM_INVALID_ARGUMENT);
  constexpr double kPi = 3.141592653589793238462643383279502884;
  constexpr double kMu0 = 4.0e-7 * kPi;
  out->age_mode = 1;
  for (size_t g = 0; g < plan.ages.size(); ++g) {
    const femm_air_gap_t &age = plan.ages[g];
    const int32_t cells = age.cells;
    const double dt = (kPi / 180.0) * age.arc_length_deg / static_cast<double>(cells);
    const double radius = 0.5 * (age.inner_radius_m + age.outer_radius_m);
    const double dr = age.outer_radius_m - age.inner_radius_m;
    if (cells < 2 || !(dt > 0.0) || !(radius > 0.0) || !(dr > 0.0))
      return done(FEMM_MESH_INVALID);
    out->age_first.push_back(static_cast<int32_t>(out->age_cell_element.size()));
    double par[kAgeParams] = {};
    // The field of a cell uses the elements' own shifts (airgap_postprocess.cpp:81).
    CatmullRomWeights(1.0 - age.inner_shift, par);
    CatmullRomWeights(1.0 - age.outer_shift, par + 4);
    par[8] = dt;
    par[9] = radius;
    par[10] = dr;
    par[11] = depth_m * radius * dt / (2.0 * kMu0);
    par[12] = depth_m * radius * radius * dt;
    out->age_par.insert(out->age_par.end(), par, par + kAgeParams);
    for (int32_t k = 0; k < cells; ++k) {
      int32_t node[10];
      double weight[10];
      if (!AgeStencil(age, k, node, weight))
        return done(FEMM_MESH_INVALID);
      out->age_nodes.insert(out->age_nodes.end(), node, node + 10);
      out->age_w.insert(out->age_w.end(), weight, weight + 10);
      out->age_cell_element.push_back(static_cast<int32_t>(g));
    }
  }
  out->age_first.push_back(static_cast<int32_t>(out->age_cell_element.size()));
  // The first element's discrete Fourier series of Br (airgap_postprocess.cpp:117-139).
  const femm_air_gap_t &first = plan.ages.front();
  const int32_t cells = first.cells;
  const size_t harmonics = first.antiperiodic ? static_cast<size_t>(cells + 1) / 2
                                              : static_cast<size_t>(cells) / 2 + 1;
  const int32_t base = static_cast<int32_t>(
      std::lround((first.antiperiodic ? 180.0 : 360.0) / first.arc_length_deg));
  for (size_t h = 0; h < harmonics; ++h) {
    const int32_t order = first.antiperiodic ? base * static_cast<int32_t>(2 * h + 1)
                                             : base * static_cast<int32_t>(h);
    const bool dc_or_nyquist =
        order == 0 || (!first.antiperiodic && h + 1 == harmonics && (cells % 2) == 0);
    out->age_order.push_back(order);
    out->age_norm.push_back(dc_or_nyquist ? static_cast<double>(cells)
                                          : static_cast<double>(cells) / 2.0);
  }
  out->probe_angle.assign(o.angles_rad, o.angles_rad + o.angle_count);
  if (AgePostLdsBytes(*out) > kAgeLdsLimit)
    return done(FEMM_UNSUPPORTED_FEATURE);
  return done(FEMM_OK);
}

size_t AgePostLdsBytes(const HostPostPlan &post)
{
  const size_t cells0 = post.age_first.size() > 1
                            ? static_cast<size_t>(post.age_first[1] - post.age_first[0])
                            : 0;
  return sizeof(double) * (cells0 + 2 * post.age_order.size());
}

}  // namespace femm
