// Postprocess plans: upload of a host-built plan, and the air-gap element postprocess plan.
// The weighted-stress plan is built on the GPU (post_build.hip).

#include "post.hpp"

#include "age.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace femm
{

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
