// Request documents of the wire contract and their responses.
//
// * Single sample, neutral artifacts: gpu_femm_planar_dc_sample_v1 /
//   gpu_femm_magnetostatic_sample_v1, exactly 11 members (reference
//   gpu_solver/src/io/sample_request.cpp:152-210), response of
//   app/sample_adapter.cpp:226-327.
// * Single motor sample: gpu_femm_motor_sample_v1, exactly 12 members (:212-300), response
//   of app/sample_adapter.cpp:67-123.
// * Shared-geometry batch: gpu_femm_motor_batch_v1, {protocol, max_items_per_chunk 1..4096,
//   items 1..4096 of {task_id, request}}, unique task ids of [A-Za-z0-9_.-]{1,256}, item
//   protocol matching the batch version (app/batch_prepare.cpp:74-131); response of
//   app/batch_response.cpp:140-371 with the batch-level PARTIAL_FAILURE.
// Numbers are written with 17 significant digits (the reference's setprecision(17)).
#pragma once

#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "artifact.hpp"
#include "json.hpp"

namespace femm
{
namespace host
{

struct SampleRequest {
  std::string protocol, artifact_path, artifact_sha256;
  std::vector<double> currents;
  int32_t force_group = -1, air_group = -1;
  double airgap_radius_mm = 0.0;
  std::vector<double> airgap_angles_deg;
  double sliding_band_angle_deg = 0.0;
  bool compute_force_torque = false, include_field_solution = false;
  bool Generic() const { return protocol == "gpu_femm_magnetostatic_sample_v1"; }
};

struct MotorRequest {
  std::string protocol, artifact_path, artifact_sha256, base_sha256, source_sha256;
  std::vector<double> currents;
  int32_t selected_group = -1, air_group = -1;
  double airgap_radius_mm = 0.0;
  std::vector<double> airgap_angles_deg;
  double rotor_angle_deg = 0.0, displacement_mm[2] = {0.0, 0.0};
};

struct BatchItem {
  std::string task_id;
  MotorRequest request;
};

struct BatchRequest {
  std::string protocol;
  int32_t max_items_per_chunk = 1;
  std::vector<BatchItem> items;
};

bool ParseSampleRequest(const std::string &text, SampleRequest *out, std::string *error);
bool ParseMotorRequest(const Json &root, MotorRequest *out, std::string *error);
bool ParseMotorRequestText(const std::string &text, MotorRequest *out, std::string *error);
bool ParseBatchRequest(const std::string &text, BatchRequest *out, std::string *error);

// PlanarDcRequestMatchesArtifact (app/sample_adapter.cpp:183-224) for a conforming artifact.
bool SampleMatchesArtifact(const SampleRequest &r, const Artifact &a);
// RequestMatchesArtifact (io/sample_request.cpp:396-446) for a posed gpu_femm_mesh_v1 or a
// sliding-band gpu_femm_mesh_v2 artifact.
bool MotorMatchesArtifact(const MotorRequest &r, const Artifact &a);

// The solve's outputs for one item.
struct Solution {
  std::vector<double> a, bx, by, flux, currents;
  int32_t iterations = 0;            // Newton iterations
  double residual_l2 = 0.0;          // last relative undamped step
};

struct Loads {
  double fx = 0.0, fy = 0.0, torque = 0.0;
  std::vector<double> radial_b;
};

// Response writers.  `request` / `solution` / `loads` may be null when the stage was not
// reached.
void WriteSampleResponse(std::ostream &out, int status, int solve_status, bool post_requested,
                         int post_status, const SampleRequest *request, const Solution *solution,
                         const Loads *loads, const std::string &diagnostic,
                         const Artifact *artifact);
void WriteMotorResponse(std::ostream &out, int status, const MotorRequest *request,
                        const Solution *solution, const Loads *loads, const char *indent);

struct BatchCounters {
  int32_t requested_chunk = 0, effective_chunk = 0, mesh_cache_hits = 0, chunk_count = 0;
  int32_t parallel_width = 0, batched_launches = 0;
  int64_t symbolic_cache_hits = 0;
};

struct BatchItemOutcome {
  int status = 2;
  Solution solution;
  bool solved = false;
  Loads loads;
};

void WriteBatchResponse(std::ostream &out, const BatchRequest &request,
                        const std::vector<BatchItemOutcome> &items, const BatchCounters &c);

}  // namespace host
}  // namespace femm
