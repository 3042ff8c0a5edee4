// Definition of the opaque femm_plan handle, shared by the C entry points.
#pragma once

#include <vector>

#include "device.hpp"
#include "femm_gpu_amd.h"
#include "plan.hpp"
#include "post.hpp"

// Retained copy of the model arrays the postprocess plan needs.
struct ModelCopy {
  std::vector<double> node_xy, dirichlet_values;
  std::vector<int32_t> tri_nodes, tri_mat, dirichlet_nodes, curve_offsets{0};
  std::vector<femm_material_t> materials;
  std::vector<femm_node_constraint_t> constraints;
  std::vector<int32_t> all_curve_offsets;   // the model's full curve layout
  int32_t curve_count = 0;
  femm_model_t view{};
};

struct femm_plan {
  femm::HostPlan host;
  femm::DeviceBuffers dev;
  ModelCopy model;
  femm::HostPostPlan post;
  femm::DevicePostPlan post_dev{};
  void *post_base = nullptr;
  bool post_ready = false;
  // staging for the host-pointer entry points
  void *io = nullptr;
  size_t io_bytes = 0;
};

namespace femm
{
// Uploads the immutable plan on first use (femm_plan_create works without a GPU).
int EnsurePlanDevice(femm_plan *plan);
// Probe solve of the autotuners: two Newton iterations of 200 PCG iterations each (a
// tolerance no solve reaches) of `items` items at unit circuit currents with `partitions`
// workgroups per item; *ms is the second of two timed launches.
int ProbeSolve(femm_plan *plan, int32_t items, const femm_solve_config_t &config,
               int32_t partitions, float *ms);
// Precision and preconditioner block size of a solve config are supported values.
bool ConfigValid(const femm_solve_config_t &c);
femm_newton_options_t DefaultNewtonOptions();
femm_solve_config_t DefaultSolveConfig();
bool DevicePresent();
}  // namespace femm
