// GPU FEMM mesh artifacts: strict parse into a femm_model_t (reference
// gpu_solver/src/io/mesh_artifact.cpp:14-440).
//
// Schemas: gpu_femm_planar_dc_mesh_v1 and gpu_femm_magnetostatic_mesh_v1 (neutral: lower-case
// identities, canonical identity of `resolved` checked), gpu_femm_mesh_v1 (a posed,
// conforming motor mesh with base / pose identities) and gpu_femm_mesh_v2 (a sliding-band
// motor mesh at rotor angle 0 whose air gap is one or more native air-gap elements,
// mesh_artifact.cpp:594-705).  Sector motor schemas (v3 / v4) are rejected as unsupported.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "femm_model.h"

namespace femm
{
namespace host
{

struct Artifact {
  std::string schema;
  bool neutral = false;        // planar-DC or magnetostatic neutral schema
  bool generic = false;        // gpu_femm_magnetostatic_mesh_v1
  std::string source_sha256;   // resolved.source_fem_sha256 (the pose's FEM file for motors)
  std::string base_sha256;     // base_motor_fem_sha256 (= source for neutral schemas)
  std::string identity_sha256; // canonical_identity_sha256
  double rotor_angle_deg = 0.0, displacement_mm[2] = {0.0, 0.0};

  int32_t problem_type = FEMM_PLANAR;
  double depth_m = 0.0;
  std::vector<double> node_xy;                  // [N][2] metres
  std::vector<int32_t> tri_nodes, tri_mat;      // [T][3], [T]
  std::vector<femm_material_t> materials;       // base materials, then one per region
  std::vector<int32_t> curve_offsets{0};
  std::vector<double> curve_b, curve_h;
  std::vector<int32_t> dirichlet_nodes;
  std::vector<double> dirichlet_values;
  std::vector<femm_node_constraint_t> constraints;
  std::vector<double> currents;                 // circuits' current_A
  std::vector<int32_t> triangle_groups;         // FEMM group number per triangle
  bool sliding_band = false;                    // gpu_femm_mesh_v2
  std::vector<femm_air_gap_t> air_gaps;         // points: see age_offsets
  std::vector<femm_age_point_t> age_points;
  std::vector<size_t> age_offsets;              // first record of each element

  // A view of the arrays above (valid while this object lives and is not modified).
  femm_model_t View() const;
  // View() with the air-gap elements replaced by `gaps` (e.g. turned to a rotor angle).
  femm_model_t ViewWith(const std::vector<femm_air_gap_t> &gaps) const;
  // Points every air-gap element at its records in age_points (after a copy or move).
  void Repoint();
  bool HasGroup(int32_t group) const;
};

// Parses `text`; on failure returns false with *error.
bool ParseArtifact(const std::string &text, Artifact *out, std::string *error);

}  // namespace host
}  // namespace femm
