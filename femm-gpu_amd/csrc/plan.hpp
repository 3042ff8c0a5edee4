// Host-side plan of the nonlinear P1 solve path: everything that depends only on
// the mesh (validation, B-H splines, signed DOF map, symbolic CSR, ordered
// contributor lists, SELL-64 layout, element geometry).  Built once per
// geometry; uploaded once; shared by every batch item.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "femm_model.h"

namespace femm
{

constexpr int kSliceRows = 64;       // SELL slice height = one wave64
constexpr int kMaxQuadrature = 7;

struct Spline {
  std::vector<double> b, h, slope;
  double cold = 0.0;
  int passes = 0;
};

// FEMM DC natural spline of a B-H table (reference physics/nonlinear_model.cpp:13-113).
bool BuildSpline(Spline *s);

struct HostPlan {
  // sizes
  int32_t N = 0, T = 0, n = 0, nnz = 0, C = 0, Q = 1;
  int32_t planar = 1;
  double depth = 1.0;
  bool any_nonlinear = false;
  bool any_boundary_value = false;
  // geometry per triangle
  std::vector<int32_t> tri_nodes, tri_mat;
  std::vector<double> tri_w, tri_gx, tri_gy;   // [T][Q], [T][Q][3]
  std::vector<double> tri_flux;                // [T][3] flux weight per local node (0 if no circuit)
  std::vector<int32_t> tri_circuit;            // [T]
  // materials
  std::vector<double> mat_nu0, mat_alpha, mat_cold;
  std::vector<int32_t> mat_curve_off, mat_curve_cnt;
  std::vector<double> curve_b, curve_h, curve_s;
  // DOF map
  std::vector<int32_t> node_to_free, free_nodes;
  std::vector<int8_t> node_sign;
  std::vector<double> boundary;
  // symbolic CSR
  std::vector<int32_t> row_offsets, cols, diag_slot;
  // slot-owned contributors (packed: t << 5 | i << 3 | j << 1 | negative), element/i/j order
  std::vector<int32_t> slot_off, contrib;
  // row-owned rhs contributors (packed: t << 3 | i << 1 | negative), element order
  std::vector<int32_t> row_off, rcontrib;
  // [n] PM load (plus age_rhs: the rhs gather's first sum, pm + age, done here once),
  // [C][n] circuit loads per ampere
  std::vector<double> pm_rhs, circuit_rhs;
  // native air-gap elements: fixed stiffness pre-summed per CSR slot (added after the
  // triangle terms) and the fixed-node load per row; empty without AGEs
  std::vector<double> age_vals, age_rhs;       // [nnz], [n]
  std::vector<femm_air_gap_t> ages;            // points re-pointed into age_points
  std::vector<femm_age_point_t> age_points;
  // SELL-64
  int32_t nslices = 0, sell_size = 0;
  std::vector<int32_t> slice_ptr, slice_width, sell_cols, slot_to_sell, diag_sell;
  // element -> node incidence for the nodal state update (unused rows for fixed nodes)
  std::vector<int32_t> dirichlet_nodes;
};

// Validates like ValidateNonlinearModel (reference physics/nonlinear_model.cpp:223-319) and
// builds the plan.  Returns a femm status code.  With `topology` (a plan built from a
// model for which SameTopology holds), only the material-dependent parts are built
// (materials, splines, per-triangle material / circuit / flux weights, PM and circuit
// right-hand sides); the DOF map is taken from `topology` and the symbolic CSR and SELL
// layout are left empty.  With `validate_only` every model check runs (the reference's
// ValidateNonlinearModel and the DOF map's) and the plan stops before the symbolic CSR.
int BuildHostPlan(const femm_model_t *model, HostPlan *plan, std::string *error,
                  const HostPlan *topology = nullptr, bool validate_only = false);

// True when `variant` shares `base`'s mesh, Dirichlet nodes and values, node constraints,
// problem type, depth, circuit count, material count and B-H knot layout, so that only
// material values, the triangle-to-material map and B-H knot values may differ.
bool SameTopology(const femm_model_t *base, const femm_model_t *variant);

}  // namespace femm
