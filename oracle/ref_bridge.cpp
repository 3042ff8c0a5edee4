// Bridge from the flat femm_model_t contract to the reference's own host
// sources, compiled unmodified from the reference tree by oracle/Makefile into
// oracle/_ref/libref_bridge.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests to pin the CPU oracle against the
// reference's host assembler, symbolic CSR builder, B-H spline, artifact
// parser, sliding-band remap and air-gap element postprocess.  The reference's
// weighted-stress postprocess is not bridged: it calls a helper defined in a CUDA
// translation unit (postprocess/weighted_stress.cu).  It is loaded with RTLD_LAZY
// because the reference's AGE matrix translation unit needs build-generated code and is
// not compiled: the assembly entry points refuse models with air-gap elements
// (FEMM_UNSUPPORTED_FEATURE) instead of reaching it.

#include "physics/nonlinear_model.h"
#include "physics/nonlinear_topology.h"
#include "io/mesh_artifact.h"
#include "io/sample_request.h"
#include "postprocess/postprocess.h"

#include "femm_model.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using namespace gpu_femm;

namespace
{

bool ToReference(const femm_model_t *d, NonlinearModel *m)
{
  m->nodes.resize(d->node_count);
  for (int i = 0; i < d->node_count; ++i)
    m->nodes[i] = {d->node_xy[2 * i], d->node_xy[2 * i + 1]};
  m->triangles.resize(d->triangle_count);
  for (int e = 0; e < d->triangle_count; ++e) {
    for (int k = 0; k < 3; ++k)
      m->triangles[e].node[k] = d->triangle_nodes[3 * e + k];
    m->triangles[e].material = d->triangle_material[e];
  }
  for (int i = 0; i < d->material_count; ++i) {
    const femm_material_t &s = d->materials[i];
    NonlinearMaterial t;
    t.reluctivity_zero_m_per_h = s.reluctivity_zero;
    t.alpha_per_t2 = s.alpha;
    t.source_j_per_a = s.source_j_per_a;
    t.h_c_a_per_m = s.hc;
    t.magnetization_deg = s.magnetization_deg;
    t.bh_curve_index = s.bh_curve;
    t.circuit_index = s.circuit;
    t.group_number = s.group;
    m->materials.push_back(t);
  }
  for (int c = 0; c < d->curve_count; ++c) {
    NonlinearBhCurve curve;
    curve.b_t.assign(d->curve_b + d->curve_offsets[c], d->curve_b + d->curve_offsets[c + 1]);
    curve.h_a_per_m.assign(d->curve_h + d->curve_offsets[c], d->curve_h + d->curve_offsets[c + 1]);
    if (!BuildNonlinearBhSpline(&curve))
      return false;
    m->bh_curves.push_back(std::move(curve));
  }
  m->dirichlet_nodes.assign(d->dirichlet_nodes, d->dirichlet_nodes + d->dirichlet_count);
  m->dirichlet_a_wb_per_m.assign(d->dirichlet_values, d->dirichlet_values + d->dirichlet_count);
  for (int i = 0; i < d->constraint_count; ++i) {
    SignedNodeConstraint c;
    c.node_a = d->constraints[i].node_a;
    c.node_b = d->constraints[i].node_b;
    c.sign = static_cast<int8_t>(d->constraints[i].sign);
    m->node_constraints.push_back(c);
  }
  m->circuit_count = d->circuit_count;
  m->depth_m = d->depth_m;
  m->problem_type = d->problem_type == FEMM_AXISYMMETRIC ? MagneticProblemType::kAxisymmetric
                                                          : MagneticProblemType::kPlanar;
  for (int g = 0; g < d->air_gap_count; ++g) {
    const femm_air_gap_t &src = d->air_gaps[g];
    AirGapElement age;
    age.antiperiodic = src.antiperiodic != 0;
    age.center_x_m = src.center_xy[0];
    age.center_y_m = src.center_xy[1];
    age.inner_radius_m = src.inner_radius_m;
    age.outer_radius_m = src.outer_radius_m;
    age.arc_length_deg = src.arc_length_deg;
    age.inner_shift = src.inner_shift;
    age.outer_shift = src.outer_shift;
    for (int k = 0; k <= src.cells; ++k) {
      AirGapQuadPoint point;
      for (int q = 0; q < 4; ++q) {
        point.node[q] = src.points[k].node[q];
        point.weight[q] = src.points[k].weight[q];
      }
      age.quad_points.push_back(point);
    }
    m->air_gap_elements.push_back(std::move(age));
  }
  return true;
}

struct ParsedArtifact {
  GpuFemmMeshArtifact artifact;
  std::vector<double> node_xy, curve_b, curve_h;
  std::vector<int32_t> tri_nodes, tri_mat, curve_offsets;
  std::vector<femm_material_t> materials;
  std::vector<femm_node_constraint_t> constraints;
};

}  // namespace

extern "C" {

struct ref_assembly {
  Assembly a;
};

struct ref_symbolic {
  PreparedNonlinearAssembly p;
};

int ref_bh_spline(const double *b, const double *h, int n, double *out_b, double *out_h,
                  double *out_slope, double *cold, int *passes)
{
  NonlinearBhCurve c;
  c.b_t.assign(b, b + n);
  c.h_a_per_m.assign(h, h + n);
  const bool ok = BuildNonlinearBhSpline(&c);
  if (ok) {
    std::copy(c.b_t.begin(), c.b_t.end(), out_b);
    std::copy(c.h_a_per_m.begin(), c.h_a_per_m.end(), out_h);
    std::copy(c.slope.begin(), c.slope.end(), out_slope);
  }
  *cold = c.cold_secant_reluctivity_m_per_h;
  *passes = c.smoothing_passes;
  return ok ? 1 : 0;
}

void ref_bh_eval(const double *b, const double *h, const double *slope, int n, double field,
                 double out[4])
{
  NonlinearBhCurve c;
  c.b_t.assign(b, b + n);
  c.h_a_per_m.assign(h, h + n);
  c.slope.assign(slope, slope + n);
  const NonlinearBhEvaluation e = EvaluateNonlinearBh(c, field);
  out[0] = e.h_a_per_m;
  out[1] = e.dh_db;
  out[2] = e.reluctivity_m_per_h;
  out[3] = e.dv_db2;
}

int ref_validate(const femm_model_t *d)
{
  NonlinearModel m;
  if (!ToReference(d, &m))
    return static_cast<int>(Status::kInvalidMaterial);
  return static_cast<int>(ValidateNonlinearModel(m));
}

int ref_assemble(const femm_model_t *d, const double *a_old, const double *currents, int newton,
                 ref_assembly **out)
{
  *out = nullptr;
  if (d->air_gap_count > 0)
    return static_cast<int>(Status::kUnsupportedFeature);
  NonlinearModel m;
  if (!ToReference(d, &m))
    return static_cast<int>(Status::kInvalidMaterial);
  auto *h = new ref_assembly;
  const Status st = AssembleNonlinearNewton(
      m, std::vector<double>(a_old, a_old + m.nodes.size()),
      std::vector<double>(currents, currents + m.circuit_count), newton != 0, &h->a);
  if (st != Status::kOk) {
    delete h;
    return static_cast<int>(st);
  }
  *out = h;
  return 0;
}

void ref_assembly_sizes(const ref_assembly *h, int32_t *n, int32_t *nnz)
{
  *n = static_cast<int32_t>(h->a.free_nodes.size());
  *nnz = static_cast<int32_t>(h->a.values.size());
}

void ref_assembly_copy(const ref_assembly *h, int32_t *rows, int32_t *cols, double *values,
                       double *diagonal, double *rhs_per_amp, double *rhs_offset,
                       int32_t *free_nodes, int32_t *node_to_free, int8_t *node_sign,
                       double *boundary)
{
  const Assembly &a = h->a;
  std::copy(a.row_offsets.begin(), a.row_offsets.end(), rows);
  std::copy(a.column_indices.begin(), a.column_indices.end(), cols);
  std::copy(a.values.begin(), a.values.end(), values);
  std::copy(a.diagonal.begin(), a.diagonal.end(), diagonal);
  std::copy(a.rhs_per_amp.begin(), a.rhs_per_amp.end(), rhs_per_amp);
  std::copy(a.rhs_offset.begin(), a.rhs_offset.end(), rhs_offset);
  std::copy(a.free_nodes.begin(), a.free_nodes.end(), free_nodes);
  std::copy(a.node_to_free.begin(), a.node_to_free.end(), node_to_free);
  std::copy(a.node_sign.begin(), a.node_sign.end(), node_sign);
  std::copy(a.boundary_values.begin(), a.boundary_values.end(), boundary);
}

void ref_assembly_destroy(ref_assembly *h) { delete h; }

int ref_prepare_symbolic(const femm_model_t *d, ref_symbolic **out)
{
  *out = nullptr;
  if (d->air_gap_count > 0)
    return static_cast<int>(Status::kUnsupportedFeature);
  NonlinearModel m;
  if (!ToReference(d, &m))
    return static_cast<int>(Status::kInvalidMaterial);
  auto *h = new ref_symbolic;
  const Status st = PrepareNonlinearAssembly(m, &h->p);
  if (st != Status::kOk) {
    delete h;
    return static_cast<int>(st);
  }
  *out = h;
  return 0;
}

void ref_symbolic_sizes(const ref_symbolic *h, int32_t *n, int32_t *nnz, int32_t *me, int32_t *re)
{
  *n = static_cast<int32_t>(h->p.symbolic.free_nodes.size());
  *nnz = static_cast<int32_t>(h->p.symbolic.column_indices.size());
  *me = static_cast<int32_t>(h->p.matrix_entries.size());
  *re = static_cast<int32_t>(h->p.rhs_entries.size());
}

void ref_symbolic_copy(const ref_symbolic *h, int32_t *rows, int32_t *cols, int32_t *diag_slots,
                       int32_t *mat_offsets, int32_t *mat_entries, int32_t *rhs_offsets,
                       int32_t *rhs_entries, double *pm_rhs, double *circuit_rhs,
                       int32_t *free_nodes, int32_t *node_to_free, int8_t *node_sign,
                       double *boundary)
{
  const PreparedNonlinearAssembly &p = h->p;
  std::copy(p.symbolic.row_offsets.begin(), p.symbolic.row_offsets.end(), rows);
  std::copy(p.symbolic.column_indices.begin(), p.symbolic.column_indices.end(), cols);
  std::copy(p.diagonal_slots.begin(), p.diagonal_slots.end(), diag_slots);
  std::copy(p.matrix_offsets.begin(), p.matrix_offsets.end(), mat_offsets);
  for (size_t k = 0; k < p.matrix_entries.size(); ++k) {
    mat_entries[4 * k] = p.matrix_entries[k].triangle;
    mat_entries[4 * k + 1] = p.matrix_entries[k].local_row;
    mat_entries[4 * k + 2] = p.matrix_entries[k].local_column;
    mat_entries[4 * k + 3] = p.matrix_entries[k].sign;
  }
  std::copy(p.rhs_offsets.begin(), p.rhs_offsets.end(), rhs_offsets);
  for (size_t k = 0; k < p.rhs_entries.size(); ++k) {
    rhs_entries[3 * k] = p.rhs_entries[k].triangle;
    rhs_entries[3 * k + 1] = p.rhs_entries[k].local_row;
    rhs_entries[3 * k + 2] = p.rhs_entries[k].sign;
  }
  std::copy(p.pm_rhs.begin(), p.pm_rhs.end(), pm_rhs);
  std::copy(p.circuit_rhs.begin(), p.circuit_rhs.end(), circuit_rhs);
  std::copy(p.symbolic.free_nodes.begin(), p.symbolic.free_nodes.end(), free_nodes);
  std::copy(p.symbolic.node_to_free.begin(), p.symbolic.node_to_free.end(), node_to_free);
  std::copy(p.symbolic.node_sign.begin(), p.symbolic.node_sign.end(), node_sign);
  std::copy(p.boundary_values.begin(), p.boundary_values.end(), boundary);
}

void ref_symbolic_destroy(ref_symbolic *h) { delete h; }

// Parse a generic mesh artifact with the reference parser and expose the
// resulting model in the flat contract (curves as the parser's raw knots).
void *ref_parse_artifact(const char *text, size_t length, char *error, size_t error_capacity)
{
  auto *p = new ParsedArtifact;
  std::string err;
  if (!ParseGpuFemmMeshArtifactJson(std::string(text, length), &p->artifact, &err)) {
    std::strncpy(error, err.c_str(), error_capacity - 1);
    error[error_capacity - 1] = 0;
    delete p;
    return nullptr;
  }
  const NonlinearModel &m = p->artifact.model;
  for (const auto &n : m.nodes) {
    p->node_xy.push_back(n.x_m);
    p->node_xy.push_back(n.y_m);
  }
  for (const auto &t : m.triangles) {
    for (int k = 0; k < 3; ++k)
      p->tri_nodes.push_back(t.node[k]);
    p->tri_mat.push_back(t.material);
  }
  for (const auto &s : m.materials) {
    femm_material_t t{};
    t.reluctivity_zero = s.reluctivity_zero_m_per_h;
    t.alpha = s.alpha_per_t2;
    t.source_j_per_a = s.source_j_per_a;
    t.hc = s.h_c_a_per_m;
    t.magnetization_deg = s.magnetization_deg;
    t.bh_curve = s.bh_curve_index;
    t.circuit = s.circuit_index;
    t.group = s.group_number;
    p->materials.push_back(t);
  }
  p->curve_offsets.push_back(0);
  for (const auto &c : m.bh_curves) {
    p->curve_b.insert(p->curve_b.end(), c.b_t.begin(), c.b_t.end());
    p->curve_h.insert(p->curve_h.end(), c.h_a_per_m.begin(), c.h_a_per_m.end());
    p->curve_offsets.push_back(static_cast<int32_t>(p->curve_b.size()));
  }
  for (const auto &c : m.node_constraints)
    p->constraints.push_back({c.node_a, c.node_b, c.sign});
  return p;
}

void ref_artifact_sizes(const void *handle, int32_t sizes[6])
{
  const auto *p = static_cast<const ParsedArtifact *>(handle);
  const NonlinearModel &m = p->artifact.model;
  sizes[0] = static_cast<int32_t>(m.nodes.size());
  sizes[1] = static_cast<int32_t>(m.triangles.size());
  sizes[2] = static_cast<int32_t>(m.materials.size());
  sizes[3] = static_cast<int32_t>(m.bh_curves.size());
  sizes[4] = static_cast<int32_t>(m.dirichlet_nodes.size());
  sizes[5] = static_cast<int32_t>(m.node_constraints.size());
}

void ref_artifact_copy(const void *handle, double *node_xy, int32_t *tri_nodes, int32_t *tri_mat,
                       femm_material_t *materials, int32_t *dirichlet_nodes, double *currents,
                       femm_node_constraint_t *constraints, double *depth_m, int32_t *problem,
                       int32_t *circuits)
{
  const auto *p = static_cast<const ParsedArtifact *>(handle);
  const NonlinearModel &m = p->artifact.model;
  std::copy(p->node_xy.begin(), p->node_xy.end(), node_xy);
  std::copy(p->tri_nodes.begin(), p->tri_nodes.end(), tri_nodes);
  std::copy(p->tri_mat.begin(), p->tri_mat.end(), tri_mat);
  std::copy(p->materials.begin(), p->materials.end(), materials);
  std::copy(m.dirichlet_nodes.begin(), m.dirichlet_nodes.end(), dirichlet_nodes);
  std::copy(p->artifact.circuit_currents_a.begin(), p->artifact.circuit_currents_a.end(), currents);
  std::copy(p->constraints.begin(), p->constraints.end(), constraints);
  *depth_m = m.depth_m;
  *problem = m.problem_type == MagneticProblemType::kAxisymmetric ? FEMM_AXISYMMETRIC : FEMM_PLANAR;
  *circuits = m.circuit_count;
}

void ref_artifact_destroy(void *handle) { delete static_cast<ParsedArtifact *>(handle); }

/* The parsed artifact's air-gap elements: returns their count (-1 without a sliding band)
 * and the number of records of all elements in *total; with non-null outputs,
 * params[g][8] = {ri, ro, arc, inner shift, outer shift, centre x, centre y, antiperiodic}
 * and every element's records concatenated into points. */
int ref_artifact_air_gaps(const void *handle, double *params, femm_age_point_t *points,
                          int32_t *total)
{
  const auto *p = static_cast<const ParsedArtifact *>(handle);
  if (!p->artifact.has_sliding_band)
    return -1;
  size_t at = 0;
  const auto &ages = p->artifact.model.air_gap_elements;
  for (size_t g = 0; g < ages.size(); ++g) {
    const AirGapElement &a = ages[g];
    if (params) {
      const double v[8] = {a.inner_radius_m, a.outer_radius_m, a.arc_length_deg, a.inner_shift,
                           a.outer_shift, a.center_x_m, a.center_y_m, a.antiperiodic ? 1.0 : 0.0};
      std::copy(v, v + 8, params + 8 * g);
    }
    for (const AirGapQuadPoint &qp : a.quad_points) {
      if (points)
        for (int q = 0; q < 4; ++q) {
          points[at].node[q] = qp.node[q];
          points[at].weight[q] = qp.weight[q];
        }
      ++at;
    }
  }
  *total = static_cast<int32_t>(at);
  return static_cast<int>(ages.size());
}

/* ApplySlidingBandAngle (io/sample_request.cpp:448-491) on a sliding-band artifact posed at
 * rotor angle 0: the air-gap elements of the model rotated to angle_deg.  inner_shift[g] and
 * points (every element's cells + 1 records, concatenated) are written; returns 1 on success. */
int ref_sliding_band(const femm_model_t *d, double angle_deg, double *inner_shift,
                     femm_age_point_t *points)
{
  GpuFemmMeshArtifact artifact;
  if (!ToReference(d, &artifact.model))
    return 0;
  artifact.has_sliding_band = true;
  artifact.rotor_angle_deg = 0.0;
  NonlinearModel rotated;
  if (!ApplySlidingBandAngle(angle_deg, artifact, &rotated))
    return 0;
  size_t at = 0;
  for (size_t g = 0; g < rotated.air_gap_elements.size(); ++g) {
    const AirGapElement &age = rotated.air_gap_elements[g];
    inner_shift[g] = age.inner_shift;
    for (const AirGapQuadPoint &qp : age.quad_points) {
      for (int q = 0; q < 4; ++q) {
        points[at].node[q] = qp.node[q];
        points[at].weight[q] = qp.weight[q];
      }
      ++at;
    }
  }
  return 1;
}

/* ComputeAirGapElementPostprocess (postprocess/airgap_postprocess.cpp:41-160) for the nodal
 * potential a[node_count]: out[3] = {Fx, Fy, torque}; radial_b[angle_count] from the first
 * element's Fourier series.  Returns the reference status. */
int ref_age_postprocess(const femm_model_t *d, const double *a, const double *angles_rad,
                        int angle_count, double out[3], double *radial_b)
{
  NonlinearModel m;
  if (!ToReference(d, &m))
    return static_cast<int>(Status::kInvalidMaterial);
  NonlinearSolveResult solution;
  solution.a_wb_per_m.assign(a, a + d->node_count);
  FrozenPostprocessOptions options;
  options.airgap_angles_rad.assign(angles_rad, angles_rad + angle_count);
  const FrozenPostprocessResult r = ComputeAirGapElementPostprocess(m, solution, options);
  out[0] = r.force_x_n;
  out[1] = r.force_y_n;
  out[2] = r.torque_nm;
  for (size_t k = 0; k < r.airgap_samples.size() && k < static_cast<size_t>(angle_count); ++k)
    radial_b[k] = r.airgap_samples[k].radial_b_t;
  return static_cast<int>(r.status);
}

}  // extern "C"
