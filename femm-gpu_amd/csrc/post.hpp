// Weighted-stress force/torque and air-gap radial B: geometry-only plan
// (mask, mask gradient, probe location) on the device.
#pragma once

#include <cstdint>
#include <vector>

#include "device.hpp"
#include "femm_gpu_amd.h"
#include "plan.hpp"

namespace femm
{

struct HostPostPlan {
  int32_t status = FEMM_INTERNAL_ERROR;
  std::vector<double> mask;                         // [N] 0/1
  // elements with a nonzero mask gradient: index, gradient, weight (area * depth), centroid
  std::vector<int32_t> wst_elem;
  std::vector<double> wst_hx, wst_hy, wst_w, wst_cx, wst_cy;
  // probes: containing element, barycentric coordinates, direction
  std::vector<int32_t> probe_elem;
  std::vector<double> probe_lambda, probe_cos, probe_sin;
  // node -> incident elements (element order), node coordinates, triangle labels
  std::vector<int32_t> inc_off, inc_elem;
  std::vector<double> node_xy;
  int32_t mask_iterations = 0;
  // Air-gap element postprocess (plans with air-gap elements): per cell (elements
  // concatenated) its ten nodes, weights and element index; per element kAgeParams values
  // (Catmull-Rom weights of both rings at the element's shifts, dt, mean radius, radial
  // width, force and torque scales) and its first cell; the first element's Fourier
  // harmonics (order, normalisation) and the probe angles.
  int32_t age_mode = 0;
  std::vector<int32_t> age_nodes, age_cell_element, age_first;
  std::vector<double> age_w, age_par;
  std::vector<int32_t> age_order;
  std::vector<double> age_norm, probe_angle;
};

constexpr int kAgeParams = 16;   // wi[4], wo[4], dt, radius, dr, force scale, torque scale

struct DevicePostPlan {
  int32_t nwst, nprobe, N, T, Q;
  int32_t age_mode, age_cells, age_cells0, age_harmonics;
  const int32_t *age_nodes, *age_cell_element, *age_first, *age_order;
  const double *age_w, *age_par, *age_norm, *probe_angle;
  const int32_t *wst_elem, *probe_elem, *inc_off, *inc_elem, *tri_nodes, *tri_mat;
  const double *wst_hx, *wst_hy, *wst_w, *wst_cx, *wst_cy;
  const double *probe_lambda, *probe_cos, *probe_sin, *node_xy, *tri_gx, *tri_gy;
};

// Builds the weighted-stress plan on the GPU (PrepareWeightedStressMask +
// SolveWeightedStressMask + BuildFrozenPostprocessPlanWithMask of the reference,
// weighted_stress.cu:24-183 and field_postprocess.cpp:221-339) from the solve plan's device
// arrays: `out` points into one allocation returned in `*base`; `host` receives the status,
// the node mask and the mask solve's iterations.  On failure nothing stays allocated.
int BuildPostPlanDevice(const femm_model_t *model, const HostPlan &plan, const DeviceBuffers &dev,
                        const femm_postprocess_options_t &options, HostPostPlan *host,
                        DevicePostPlan *out, void **base);

int UploadPostPlan(const HostPostPlan &host, const DeviceBuffers &dev, DevicePostPlan *out,
                   void **base);

// The air-gap element postprocess plan (ComputeAirGapElementPostprocess's geometry,
// postprocess/airgap_postprocess.cpp:41-139) for a plan with air-gap elements.
int BuildAgePostPlan(const HostPlan &plan, double depth_m, const femm_postprocess_options_t &o,
                     HostPostPlan *out);

// Dynamic LDS bytes of the air-gap element postprocess (the first element's cell values and
// Fourier coefficients); plans above kAgeLdsLimit are refused.
size_t AgePostLdsBytes(const HostPostPlan &post);
constexpr size_t kAgeLdsLimit = 60 * 1024;

// loads[items][3] = {Fx, Fy, torque}; radial_b[items][nprobe].
int LaunchPostprocess(const DevicePostPlan &post, int32_t items, const double *states,
                      double *loads, double *radial_b, hipStream_t stream);

}  // namespace femm
