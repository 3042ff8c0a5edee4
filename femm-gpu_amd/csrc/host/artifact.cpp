// Strict mesh-artifact parser (reference gpu_solver/src/io/mesh_artifact.cpp:14-440).
//
// The rules, in the reference's order: the exact member sets of the header and `resolved`
// per schema; identity formats (neutral schemas: lower-case SHA-256 and the canonical SHA-256
// of `resolved` equal to canonical_identity_sha256); resolved / header identities equal;
// static planar (or, for the magnetostatic schema, axisymmetric) physics; materials (unique
// ids, isotropic, no laminations, B-H tables of at most 4096 non-negative non-decreasing
// points that make a valid FEMM spline); circuits listed in index order as current-driven
// series circuits; regions (turns only with a circuit, a known material); CCW triangles of
// known regions; every region with area; outer Dirichlet nodes in range with value zero;
// signed periodic / anti-periodic node pairs.  Units: mm -> m; nu0 = 1 / (mu0 mu_r); every
// region becomes its own material label (group, circuit, PM angle, J per ampere =
// turns / region area) after the base materials.

#include "artifact.hpp"

#include <algorithm>
#include <cmath>
#include <map>

#include "json.hpp"
#include "plan.hpp"

namespace femm
{
namespace host
{
namespace
{

constexpr double kMu0 = 4.0e-7 * 3.141592653589793238462643383279502884;

bool Fail(std::string *error, const char *what)
{
  if (error->empty())
    *error = what;
  return false;
}

bool Sha(const Json *v, bool lower)
{
  return v != nullptr && IsSha256Hex(v->text, lower);
}

struct Schema {
  bool neutral = false, generic = false, motor = false, sliding = false;
};

bool Header(const Json &root, Artifact *a, Schema *sc, const Json **resolved, std::string *error)
{
  const Json *schema = Member(root, "schema_version", Json::Type::kString, error);
  if (schema == nullptr)
    return false;
  const std::string &s = schema->text;
  sc->generic = s == "gpu_femm_magnetostatic_mesh_v1";
  sc->neutral = sc->generic || s == "gpu_femm_planar_dc_mesh_v1";
  sc->sliding = s == "gpu_femm_mesh_v2";
  sc->motor = s == "gpu_femm_mesh_v1" || sc->sliding;
  if (s == "gpu_femm_mesh_v3" || s == "gpu_femm_mesh_v4") {
    *error = "unsupported artifact schema: sector motor meshes";
    return false;
  }
  if (!sc->neutral && !sc->motor)
    return Fail(error, "invalid GPU FEMM mesh artifact header");
  const bool header_ok =
      sc->neutral ? ExactMembers(root, {"schema_version", "source_fem_sha256",
                                        "canonical_identity_sha256", "resolved"}, error)
                  : ExactMembers(root, {"schema_version", "base_motor_fem_sha256",
                                        "source_fem_sha256", "canonical_identity_sha256",
                                        "resolved"}, error);
  if (!header_ok)
    return false;
  const bool lower = sc->neutral;
  const Json *source = Member(root, "source_fem_sha256", Json::Type::kString, error);
  const Json *base = sc->neutral ? source
                                 : Member(root, "base_motor_fem_sha256", Json::Type::kString, error);
  const Json *identity = Member(root, "canonical_identity_sha256", Json::Type::kString, error);
  *resolved = Member(root, "resolved", Json::Type::kObject, error);
  if (!Sha(source, lower) || !Sha(base, lower) || !Sha(identity, lower) || *resolved == nullptr)
    return Fail(error, "invalid GPU FEMM mesh artifact header");
  bool body_ok = false;
  if (sc->generic)
    body_ok = ExactMembers(**resolved, {"source_fem_sha256", "model", "nodes_mm", "triangles",
                                        "regions", "materials", "circuits", "outer_dirichlet",
                                        "node_constraints"}, error);
  else if (sc->neutral)
    body_ok = ExactMembers(**resolved, {"source_fem_sha256", "model", "nodes_mm", "triangles",
                                        "regions", "materials", "circuits", "outer_dirichlet"},
                           error);
  else if (sc->sliding)
    body_ok = ExactMembers(**resolved, {"source_fem_sha256", "base_motor_fem_sha256", "model",
                                        "pose", "nodes_mm", "triangles", "regions", "materials",
                                        "circuits", "outer_dirichlet", "air_gap_elements"},
                           error);
  else
    body_ok = ExactMembers(**resolved, {"source_fem_sha256", "base_motor_fem_sha256", "model",
                                        "pose", "nodes_mm", "triangles", "regions", "materials",
                                        "circuits", "outer_dirichlet"}, error);
  if (!body_ok)
    return false;
  if (sc->neutral && CanonicalSha256(**resolved) != identity->text) {
    *error = "canonical artifact identity mismatch";
    return false;
  }
  const Json *rsource = Member(**resolved, "source_fem_sha256", Json::Type::kString, error);
  const Json *rbase = sc->neutral ? rsource
                                  : Member(**resolved, "base_motor_fem_sha256",
                                           Json::Type::kString, error);
  if (!Sha(rsource, lower) || !Sha(rbase, lower) || rsource->text != source->text ||
      rbase->text != base->text)
    return Fail(error, "source FEM SHA mismatch");
  a->schema = s;
  a->neutral = sc->neutral;
  a->generic = sc->generic;
  a->source_sha256 = source->text;
  a->base_sha256 = base->text;
  a->identity_sha256 = identity->text;
  return true;
}

bool Physics(const Json &resolved, const Schema &sc, Artifact *a, std::string *error)
{
  const Json *model = Member(resolved, "model", Json::Type::kObject, error);
  if (model == nullptr || !ExactMembers(*model, {"depth_mm", "problem_type", "frequency_hz"}, error))
    return false;
  const Json *depth = Member(*model, "depth_mm", Json::Type::kNumber, error);
  const Json *type = Member(*model, "problem_type", Json::Type::kString, error);
  const Json *frequency = Member(*model, "frequency_hz", Json::Type::kNumber, error);
  const bool planar = type != nullptr && type->text == "planar";
  const bool axisymmetric = type != nullptr && type->text == "axisymmetric";
  if (depth == nullptr || type == nullptr || frequency == nullptr || !(depth->number > 0.0) ||
      frequency->number != 0.0 || !(planar || (sc.generic && axisymmetric))) {
    *error = "unsupported model physics";
    return false;
  }
  a->depth_m = depth->number * 1e-3;
  a->problem_type = axisymmetric ? FEMM_AXISYMMETRIC : FEMM_PLANAR;
  if (sc.motor) {
    const Json *pose = Member(resolved, "pose", Json::Type::kObject, error);
    if (pose == nullptr || !ExactMembers(*pose, {"rotor_angle_deg", "displacement_mm"}, error))
      return false;
    const Json *angle = Member(*pose, "rotor_angle_deg", Json::Type::kNumber, error);
    const Json *shift = Member(*pose, "displacement_mm", Json::Type::kArray, error);
    std::vector<double> d;
    if (angle == nullptr || shift == nullptr || !NumberArray(*shift, &d) || d.size() != 2) {
      *error = "invalid pose";
      return false;
    }
    a->rotor_angle_deg = angle->number;
    a->displacement_mm[0] = d[0];
    a->displacement_mm[1] = d[1];
  }
  return true;
}

bool Nodes(const Json &list, Artifact *a, std::string *error)
{
  for (const Json &pair : list.items) {
    std::vector<double> xy;
    if (!NumberArray(pair, &xy) || xy.size() != 2)
      return Fail(error, "invalid node coordinate");
    a->node_xy.push_back(xy[0] * 1e-3);
    a->node_xy.push_back(xy[1] * 1e-3);
  }
  return true;
}

bool Materials(const Json &list, Artifact *a, std::map<int32_t, int32_t> *lookup,
               std::string *error)
{
  for (const Json &m : list.items) {
    if (!ExactMembers(m, {"id", "name", "mu_x", "mu_y", "H_c_A_per_m", "B_T", "H_A_per_m",
                          "lam_type", "lam_fill"}, error))
      return false;
    const Json *id = Member(m, "id", Json::Type::kNumber, error);
    const Json *name = Member(m, "name", Json::Type::kString, error);
    const Json *mux = Member(m, "mu_x", Json::Type::kNumber, error);
    const Json *muy = Member(m, "mu_y", Json::Type::kNumber, error);
    const Json *hc = Member(m, "H_c_A_per_m", Json::Type::kNumber, error);
    const Json *b = Member(m, "B_T", Json::Type::kArray, error);
    const Json *h = Member(m, "H_A_per_m", Json::Type::kArray, error);
    const Json *lam = Member(m, "lam_type", Json::Type::kNumber, error);
    const Json *fill = Member(m, "lam_fill", Json::Type::kNumber, error);
    int32_t mid = -1;
    if (id == nullptr || name == nullptr || mux == nullptr || muy == nullptr || hc == nullptr ||
        b == nullptr || h == nullptr || lam == nullptr || fill == nullptr || !Int32Of(*id, &mid) ||
        mid < 0 || name->text.empty() || lookup->count(mid) != 0 || !(mux->number > 0.0) ||
        mux->number != muy->number || hc->number < 0.0 || lam->number != 0.0 ||
        fill->number != 1.0 || b->items.size() != h->items.size())
      return Fail(error, "invalid material definition");
    femm_material_t mat{};
    mat.reluctivity_zero = 1.0 / (kMu0 * mux->number);
    mat.hc = hc->number;
    mat.bh_curve = -1;
    mat.circuit = -1;
    mat.group = -1;
    if (b->items.size() > 4096)
      return Fail(error, "B-H curve exceeds point cap");
    if (!b->items.empty()) {
      if (b->items.size() < 2)
        return Fail(error, "B-H curve needs two points");
      Spline curve;
      for (size_t k = 0; k < b->items.size(); ++k) {
        const Json &bk = b->items[k], &hk = h->items[k];
        if (bk.type != Json::Type::kNumber || hk.type != Json::Type::kNumber || bk.number < 0.0 ||
            hk.number < 0.0 ||
            (k > 0 && (bk.number < curve.b.back() || hk.number < curve.h.back())))
          return Fail(error, "invalid B-H point");
        curve.b.push_back(bk.number);
        curve.h.push_back(hk.number);
      }
      const std::vector<double> raw_b = curve.b, raw_h = curve.h;
      if (!BuildSpline(&curve))
        return Fail(error, "invalid B-H spline");
      mat.bh_curve = static_cast<int32_t>(a->curve_offsets.size() - 1);
      a->curve_b.insert(a->curve_b.end(), raw_b.begin(), raw_b.end());
      a->curve_h.insert(a->curve_h.end(), raw_h.begin(), raw_h.end());
      a->curve_offsets.push_back(static_cast<int32_t>(a->curve_b.size()));
    }
    lookup->emplace(mid, static_cast<int32_t>(a->materials.size()));
    a->materials.push_back(mat);
  }
  return true;
}

bool Circuits(const Json &list, Artifact *a, std::string *error)
{
  for (size_t k = 0; k < list.items.size(); ++k) {
    const Json &c = list.items[k];
    if (!ExactMembers(c, {"index", "name", "type", "current_A"}, error))
      return false;
    const Json *index = Member(c, "index", Json::Type::kNumber, error);
    const Json *name = Member(c, "name", Json::Type::kString, error);
    const Json *type = Member(c, "type", Json::Type::kString, error);
    const Json *current = Member(c, "current_A", Json::Type::kNumber, error);
    int32_t i = -1;
    if (index == nullptr || name == nullptr || type == nullptr || current == nullptr ||
        !Int32Of(*index, &i) || i != static_cast<int32_t>(k) || name->text.empty() ||
        type->text != "series")
      return Fail(error, "circuits must be ordered contiguous current-driven series circuits");
    a->currents.push_back(current->number);
  }
  return true;
}

struct Region {
  int32_t material, circuit, group;
  double turns, angle;
};

bool Regions(const Json &list, const std::map<int32_t, int32_t> &materials, size_t circuits,
             std::vector<Region> *out, std::map<int32_t, int32_t> *lookup, std::string *error)
{
  for (const Json &r : list.items) {
    if (!ExactMembers(r, {"id", "group_number", "material_id", "circuit_index", "turns",
                          "pm_magnetization_deg"}, error))
      return false;
    const Json *id = Member(r, "id", Json::Type::kNumber, error);
    const Json *group = Member(r, "group_number", Json::Type::kNumber, error);
    const Json *material = Member(r, "material_id", Json::Type::kNumber, error);
    const Json *circuit = Member(r, "circuit_index", Json::Type::kNumber, error);
    const Json *turns = Member(r, "turns", Json::Type::kNumber, error);
    const Json *angle = Member(r, "pm_magnetization_deg", Json::Type::kNumber, error);
    int32_t rid = -1, gid = -1, mid = -1, cid = -2;
    if (id == nullptr || group == nullptr || material == nullptr || circuit == nullptr ||
        turns == nullptr || angle == nullptr || !Int32Of(*id, &rid) || !Int32Of(*group, &gid) ||
        !Int32Of(*material, &mid) || !Int32Of(*circuit, &cid) || rid < 0 || gid < 0 ||
        materials.count(mid) == 0 || lookup->count(rid) != 0 || cid < -1 ||
        (cid == -1 && turns->number != 0.0) ||
        (cid >= 0 && (static_cast<size_t>(cid) >= circuits || turns->number == 0.0)))
      return Fail(error, "invalid region definition");
    lookup->emplace(rid, static_cast<int32_t>(out->size()));
    out->push_back({materials.at(mid), cid, gid, turns->number, angle->number});
  }
  return true;
}

bool Triangles(const Json &tri, const std::map<int32_t, int32_t> &regions, Artifact *a,
               std::vector<double> *region_area, std::string *error)
{
  if (!ExactMembers(tri, {"node_indices", "region_ids"}, error))
    return false;
  const Json *faces = Member(tri, "node_indices", Json::Type::kArray, error);
  const Json *ids = Member(tri, "region_ids", Json::Type::kArray, error);
  if (faces == nullptr || ids == nullptr || faces->items.empty() ||
      faces->items.size() != ids->items.size())
    return Fail(error, "invalid triangle arrays");
  const size_t N = a->node_xy.size() / 2;
  for (size_t e = 0; e < faces->items.size(); ++e) {
    const Json &f = faces->items[e];
    if (f.type != Json::Type::kArray || f.items.size() != 3)
      return Fail(error, "triangle must have three node indices");
    int32_t rid = -1;
    if (!Int32Of(ids->items[e], &rid) || regions.count(rid) == 0)
      return Fail(error, "triangle references unknown region");
    int32_t v[3];
    for (int k = 0; k < 3; ++k)
      if (!Int32Of(f.items[k], &v[k]) || v[k] < 0 || static_cast<size_t>(v[k]) >= N)
        return Fail(error, "triangle node out of range");
    const double *p0 = &a->node_xy[2 * v[0]], *p1 = &a->node_xy[2 * v[1]],
                 *p2 = &a->node_xy[2 * v[2]];
    const double det = (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]);
    if (!(det > 0.0) || !std::isfinite(det))
      return Fail(error, "triangle is not CCW");
    const int32_t r = regions.at(rid);
    (*region_area)[r] += 0.5 * det;
    a->tri_nodes.insert(a->tri_nodes.end(), v, v + 3);
    a->tri_mat.push_back(r);   // region index until the labels exist
  }
  return true;
}

bool Boundary(const Json &od, Artifact *a, std::string *error)
{
  if (!ExactMembers(od, {"node_indices", "A_Wb_per_m"}, error))
    return false;
  const Json *nodes = Member(od, "node_indices", Json::Type::kArray, error);
  const Json *values = Member(od, "A_Wb_per_m", Json::Type::kArray, error);
  if (nodes == nullptr || values == nullptr || nodes->items.empty() ||
      nodes->items.size() != values->items.size())
    return Fail(error, "invalid outer Dirichlet");
  const size_t N = a->node_xy.size() / 2;
  for (size_t k = 0; k < nodes->items.size(); ++k) {
    int32_t v = -1;
    if (!Int32Of(nodes->items[k], &v) || v < 0 || static_cast<size_t>(v) >= N ||
        values->items[k].type != Json::Type::kNumber || values->items[k].number != 0.0)
      return Fail(error, "outer Dirichlet must be zero and in range");
    a->dirichlet_nodes.push_back(v);
    a->dirichlet_values.push_back(0.0);
  }
  return true;
}

bool Constraints(const Json &list, Artifact *a, std::string *error)
{
  const size_t N = a->node_xy.size() / 2;
  for (const Json &c : list.items) {
    if (!ExactMembers(c, {"node_a", "node_b", "relation"}, error))
      return false;
    const Json *na = Member(c, "node_a", Json::Type::kNumber, error);
    const Json *nb = Member(c, "node_b", Json::Type::kNumber, error);
    const Json *rel = Member(c, "relation", Json::Type::kString, error);
    femm_node_constraint_t k{};
    if (na == nullptr || nb == nullptr || rel == nullptr || !Int32Of(*na, &k.node_a) ||
        !Int32Of(*nb, &k.node_b) || k.node_a < 0 || k.node_b < 0 ||
        static_cast<size_t>(k.node_a) >= N || static_cast<size_t>(k.node_b) >= N ||
        k.node_a == k.node_b || (rel->text != "periodic" && rel->text != "antiperiodic"))
      return Fail(error, "invalid periodic node constraint");
    k.sign = rel->text == "periodic" ? 1 : -1;
    a->constraints.push_back(k);
  }
  return true;
}

}  // namespace

// The air-gap elements of a sliding-band artifact (mesh_artifact.cpp:594-705): one element
// object, or a non-empty array of them, each with exactly its ten members; the mesh must be
// the centred reference pose at 0 degrees.
bool AirGaps(const Json &resolved, Artifact *a, std::string *error)
{
  auto it = resolved.members.find("air_gap_elements");
  const Json *ages = it == resolved.members.end() ? nullptr : &it->second;
  if (ages == nullptr ||
      (ages->type != Json::Type::kArray && ages->type != Json::Type::kObject) ||
      a->displacement_mm[0] != 0.0 || a->displacement_mm[1] != 0.0 || a->rotor_angle_deg != 0.0)
    return Fail(error, "sliding-band artifacts require a centered zero-degree reference mesh");
  std::vector<const Json *> entries;
  if (ages->type == Json::Type::kArray) {
    if (ages->items.empty())
      return Fail(error, "sliding-band requires an air-gap element");
    for (const Json &e : ages->items)
      entries.push_back(&e);
  } else {
    entries.push_back(ages);
  }
  for (const Json *entry : entries) {
    if (!ExactMembers(*entry, {"name", "periodicity", "center_mm", "ri_mm", "ro_mm",
                               "arc_length_deg", "sector_count", "inner_shift", "outer_shift",
                               "quad_points"}, error))
      return false;
    const Json *name = Member(*entry, "name", Json::Type::kString, error);
    const Json *periodicity = Member(*entry, "periodicity", Json::Type::kString, error);
    const Json *center = Member(*entry, "center_mm", Json::Type::kArray, error);
    const Json *ri = Member(*entry, "ri_mm", Json::Type::kNumber, error);
    const Json *ro = Member(*entry, "ro_mm", Json::Type::kNumber, error);
    const Json *arc = Member(*entry, "arc_length_deg", Json::Type::kNumber, error);
    const Json *sectors = Member(*entry, "sector_count", Json::Type::kNumber, error);
    const Json *inner = Member(*entry, "inner_shift", Json::Type::kNumber, error);
    const Json *outer = Member(*entry, "outer_shift", Json::Type::kNumber, error);
    const Json *points = Member(*entry, "quad_points", Json::Type::kArray, error);
    int32_t cells = 0;
    std::vector<double> c;
    if (name == nullptr || name->text.empty() || periodicity == nullptr || center == nullptr ||
        ri == nullptr || ro == nullptr || arc == nullptr || sectors == nullptr ||
        inner == nullptr || outer == nullptr || points == nullptr ||
        !NumberArray(*center, &c) || c.size() != 2 || !Int32Of(*sectors, &cells) || cells < 2 ||
        points->items.size() != static_cast<size_t>(cells) + 1 ||
        (periodicity->text != "periodic" && periodicity->text != "antiperiodic"))
      return Fail(error, "invalid sliding-band air-gap element");
    femm_air_gap_t age{};
    age.antiperiodic = periodicity->text == "antiperiodic" ? 1 : 0;
    age.cells = cells;
    age.center_xy[0] = c[0] * 1e-3;
    age.center_xy[1] = c[1] * 1e-3;
    age.inner_radius_m = ri->number * 1e-3;
    age.outer_radius_m = ro->number * 1e-3;
    age.arc_length_deg = arc->number;
    age.inner_shift = inner->number;
    age.outer_shift = outer->number;
    a->age_offsets.push_back(a->age_points.size());
    for (const Json &point : points->items) {
      if (!ExactMembers(point, {"n0", "w0", "n1", "w1", "n2", "w2", "n3", "w3"}, error))
        return false;
      femm_age_point_t rec{};
      static const char *const kNodes[4] = {"n0", "n1", "n2", "n3"};
      static const char *const kWeights[4] = {"w0", "w1", "w2", "w3"};
      for (int q = 0; q < 4; ++q) {
        const Json *n = Member(point, kNodes[q], Json::Type::kNumber, error);
        const Json *w = Member(point, kWeights[q], Json::Type::kNumber, error);
        if (n == nullptr || w == nullptr || !Int32Of(*n, &rec.node[q]))
          return Fail(error, "invalid sliding-band quadrature point");
        rec.weight[q] = w->number;
      }
      a->age_points.push_back(rec);
    }
    a->air_gaps.push_back(age);
  }
  a->sliding_band = true;
  return true;
}

void Artifact::Repoint()
{
  for (size_t g = 0; g < air_gaps.size(); ++g)
    air_gaps[g].points = age_points.data() + age_offsets[g];
}

femm_model_t Artifact::ViewWith(const std::vector<femm_air_gap_t> &gaps) const
{
  femm_model_t m = View();
  m.air_gap_count = static_cast<int32_t>(gaps.size());
  m.air_gaps = gaps.empty() ? nullptr : gaps.data();
  return m;
}

femm_model_t Artifact::View() const
{
  femm_model_t m{};
  m.problem_type = problem_type;
  m.node_count = static_cast<int32_t>(node_xy.size() / 2);
  m.node_xy = node_xy.data();
  m.triangle_count = static_cast<int32_t>(tri_mat.size());
  m.triangle_nodes = tri_nodes.data();
  m.triangle_material = tri_mat.data();
  m.material_count = static_cast<int32_t>(materials.size());
  m.materials = materials.data();
  m.curve_count = static_cast<int32_t>(curve_offsets.size() - 1);
  m.curve_offsets = curve_offsets.data();
  m.curve_b = curve_b.data();
  m.curve_h = curve_h.data();
  m.dirichlet_count = static_cast<int32_t>(dirichlet_nodes.size());
  m.dirichlet_nodes = dirichlet_nodes.data();
  m.dirichlet_values = dirichlet_values.data();
  m.constraint_count = static_cast<int32_t>(constraints.size());
  m.constraints = constraints.data();
  m.circuit_count = static_cast<int32_t>(currents.size());
  m.depth_m = depth_m;
  m.air_gap_count = static_cast<int32_t>(air_gaps.size());
  m.air_gaps = air_gaps.empty() ? nullptr : air_gaps.data();
  return m;
}

bool Artifact::HasGroup(int32_t group) const
{
  return std::any_of(materials.begin(), materials.end(),
                     [group](const femm_material_t &m) { return m.group == group; });
}

bool ParseArtifact(const std::string &text, Artifact *out, std::string *error)
{
  *out = Artifact{};
  error->clear();
  Json root;
  if (!ParseJson(text, &root, error))
    return false;
  Schema sc;
  const Json *res = nullptr;
  Artifact a;
  if (!Header(root, &a, &sc, &res, error) || !Physics(*res, sc, &a, error))
    return false;
  const Json *nodes = Member(*res, "nodes_mm", Json::Type::kArray, error);
  const Json *tris = Member(*res, "triangles", Json::Type::kObject, error);
  const Json *regions = Member(*res, "regions", Json::Type::kArray, error);
  const Json *materials = Member(*res, "materials", Json::Type::kArray, error);
  const Json *circuits = Member(*res, "circuits", Json::Type::kArray, error);
  const Json *od = Member(*res, "outer_dirichlet", Json::Type::kObject, error);
  const Json *cons = sc.generic ? Member(*res, "node_constraints", Json::Type::kArray, error)
                                : nullptr;
  if (nodes == nullptr || tris == nullptr || regions == nullptr || materials == nullptr ||
      circuits == nullptr || od == nullptr || (sc.generic && cons == nullptr))
    return false;
  if (nodes->items.size() < 3 || materials->items.empty() || regions->items.empty()) {
    *error = "empty mesh/material input";
    return false;
  }
  std::map<int32_t, int32_t> material_lookup, region_lookup;
  std::vector<Region> region_list;
  if (!Nodes(*nodes, &a, error) || !Materials(*materials, &a, &material_lookup, error) ||
      !Circuits(*circuits, &a, error) ||
      !Regions(*regions, material_lookup, a.currents.size(), &region_list, &region_lookup, error))
    return false;
  std::vector<double> area(region_list.size(), 0.0);
  if (!Triangles(*tris, region_lookup, &a, &area, error))
    return false;
  // One material label per region, after the base materials.
  const int32_t first_label = static_cast<int32_t>(a.materials.size());
  for (size_t r = 0; r < region_list.size(); ++r) {
    if (!(area[r] > 0.0)) {
      *error = "region has no triangle area";
      return false;
    }
    const Region &reg = region_list[r];
    femm_material_t label = a.materials[reg.material];
    label.group = reg.group;
    label.circuit = reg.circuit;
    label.magnetization_deg = reg.angle;
    label.source_j_per_a = reg.circuit < 0 ? 0.0 : reg.turns / area[r];
    a.materials.push_back(label);
  }
  for (int32_t &m : a.tri_mat) {
    m += first_label;
    a.triangle_groups.push_back(a.materials[m].group);
  }
  if (!Boundary(*od, &a, error) || (sc.generic && !Constraints(*cons, &a, error)))
    return false;
  if (sc.sliding && !AirGaps(*res, &a, error))
    return false;
  *out = std::move(a);
  out->Repoint();
  // ValidateNonlinearModel of the mapped model (mesh_artifact.cpp:706-709): every model
  // check of the plan build, without its symbolic CSR.
  const femm_model_t view = out->View();
  HostPlan check;
  std::string why;
  if (BuildHostPlan(&view, &check, &why, nullptr, true) != FEMM_OK) {
    *out = Artifact{};
    *error = "mapped mesh model validation failed";
    return false;
  }
  return true;
}

}  // namespace host
}  // namespace femm
