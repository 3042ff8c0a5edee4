// Request parsing, request / artifact matching and response writing of the wire contract.

#include "protocol.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>

#include "femm_gpu_amd.h"

namespace femm
{
namespace host
{
namespace
{

std::string Num(double v)
{
  char text[40];
  std::snprintf(text, sizeof(text), "%.17g", v);
  return text;
}

std::string List(const std::vector<double> &values)
{
  std::string out;
  for (size_t i = 0; i < values.size(); ++i) {
    if (i)
      out += ", ";
    out += Num(values[i]);
  }
  return out;
}

std::string Name(int status)
{
  const char *name = femm_status_name(status);
  return name ? name : "INTERNAL_ERROR";
}

bool Fail(std::string *error, const char *what)
{
  if (error->empty())
    *error = what;
  return false;
}

bool TaskIdValid(const std::string &id)
{
  if (id.empty() || id.size() > 256)
    return false;
  for (const char c : id) {
    const bool ok = (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
                    c == '_' || c == '-' || c == '.';
    if (!ok)
      return false;
  }
  return true;
}

}  // namespace

bool ParseSampleRequest(const std::string &text, SampleRequest *r, std::string *error)
{
  error->clear();
  Json root;
  if (!ParseJson(text, &root, error) ||
      !ExactMembers(root, {"protocol", "mesh_artifact_path", "mesh_artifact_sha256",
                           "circuit_currents_A", "force_group_number", "stress_air_group_number",
                           "airgap_radius_mm", "airgap_angles_deg", "sliding_band_angle_deg",
                           "compute_force_torque", "include_field_solution"}, error))
    return false;
  const Json *protocol = Member(root, "protocol", Json::Type::kString, error);
  const Json *path = Member(root, "mesh_artifact_path", Json::Type::kString, error);
  const Json *sha = Member(root, "mesh_artifact_sha256", Json::Type::kString, error);
  const Json *currents = Member(root, "circuit_currents_A", Json::Type::kArray, error);
  const Json *force = Member(root, "force_group_number", Json::Type::kNumber, error);
  const Json *air = Member(root, "stress_air_group_number", Json::Type::kNumber, error);
  const Json *radius = Member(root, "airgap_radius_mm", Json::Type::kNumber, error);
  const Json *angles = Member(root, "airgap_angles_deg", Json::Type::kArray, error);
  const Json *sliding = Member(root, "sliding_band_angle_deg", Json::Type::kNumber, error);
  const Json *compute = Member(root, "compute_force_torque", Json::Type::kBool, error);
  const Json *include = Member(root, "include_field_solution", Json::Type::kBool, error);
  int32_t selected = -2, air_group = -2;
  if (protocol == nullptr || path == nullptr || sha == nullptr || currents == nullptr ||
      force == nullptr || air == nullptr || radius == nullptr || angles == nullptr ||
      sliding == nullptr || compute == nullptr || include == nullptr ||
      (protocol->text != "gpu_femm_planar_dc_sample_v1" &&
       protocol->text != "gpu_femm_magnetostatic_sample_v1") ||
      path->text.empty() || !IsSha256Hex(sha->text, true) || !Int32Of(*force, &selected) ||
      !Int32Of(*air, &air_group) || selected < -1 || air_group < -1 ||
      (selected == -1) != (air_group == -1) || (selected >= 0 && selected == air_group) ||
      !(radius->number >= 0.0) || !NumberArray(*currents, &r->currents) ||
      !NumberArray(*angles, &r->airgap_angles_deg) ||
      (!r->airgap_angles_deg.empty() && !(radius->number > 0.0)))
    return Fail(error, "invalid gpu_femm_planar_dc_sample_v1 request");
  r->protocol = protocol->text;
  r->artifact_path = path->text;
  r->artifact_sha256 = sha->text;
  r->force_group = selected;
  r->air_group = air_group;
  r->airgap_radius_mm = radius->number;
  r->sliding_band_angle_deg = sliding->number;
  r->compute_force_torque = compute->boolean;
  r->include_field_solution = include->boolean;
  return true;
}

bool ParseMotorRequest(const Json &root, MotorRequest *r, std::string *error)
{
  const Json *protocol = Member(root, "protocol", Json::Type::kString, error);
  if (protocol != nullptr && protocol->text != "gpu_femm_motor_sample_v1" &&
      (protocol->text == "gpu_femm_motor_sample_v2" || protocol->text == "gpu_femm_motor_sample_v3" ||
       protocol->text == "gpu_femm_motor_sample_v4")) {
    *error = "unsupported motor sample protocol: tangent, sliding-band and sector samples";
    return false;
  }
  if (protocol == nullptr || protocol->text != "gpu_femm_motor_sample_v1" ||
      !ExactMembers(root, {"protocol", "mesh_artifact_path", "mesh_artifact_sha256",
                           "base_motor_fem_sha256", "source_fem_sha256", "circuit_currents_A",
                           "selected_group_number", "air_group_number", "airgap_radius_mm",
                           "airgap_angles_deg", "rotor_angle_deg", "displacement_mm"}, error))
    return Fail(error, "invalid gpu_femm_motor_sample_v1 request");
  const Json *path = Member(root, "mesh_artifact_path", Json::Type::kString, error);
  const Json *sha = Member(root, "mesh_artifact_sha256", Json::Type::kString, error);
  const Json *base = Member(root, "base_motor_fem_sha256", Json::Type::kString, error);
  const Json *source = Member(root, "source_fem_sha256", Json::Type::kString, error);
  const Json *currents = Member(root, "circuit_currents_A", Json::Type::kArray, error);
  const Json *selected = Member(root, "selected_group_number", Json::Type::kNumber, error);
  const Json *air = Member(root, "air_group_number", Json::Type::kNumber, error);
  const Json *radius = Member(root, "airgap_radius_mm", Json::Type::kNumber, error);
  const Json *angles = Member(root, "airgap_angles_deg", Json::Type::kArray, error);
  const Json *rotor = Member(root, "rotor_angle_deg", Json::Type::kNumber, error);
  const Json *shift = Member(root, "displacement_mm", Json::Type::kArray, error);
  std::vector<double> d;
  if (path == nullptr || sha == nullptr || base == nullptr || source == nullptr ||
      currents == nullptr || selected == nullptr || air == nullptr || radius == nullptr ||
      angles == nullptr || rotor == nullptr || shift == nullptr || path->text.empty() ||
      !IsSha256Hex(sha->text, false) || !IsSha256Hex(base->text, false) ||
      !IsSha256Hex(source->text, false) || !Int32Of(*selected, &r->selected_group) ||
      !Int32Of(*air, &r->air_group) || r->selected_group < 0 || r->air_group < 0 ||
      r->selected_group == r->air_group || !(radius->number >= 0.0) || !NumberArray(*shift, &d) ||
      d.size() != 2 || !NumberArray(*currents, &r->currents) ||
      !NumberArray(*angles, &r->airgap_angles_deg) || r->currents.empty() ||
      (!r->airgap_angles_deg.empty() && !(radius->number > 0.0)))
    return Fail(error, "invalid gpu_femm_motor_sample_v1 request");
  r->protocol = protocol->text;
  r->artifact_path = path->text;
  r->artifact_sha256 = sha->text;
  r->base_sha256 = base->text;
  r->source_sha256 = source->text;
  r->airgap_radius_mm = radius->number;
  r->rotor_angle_deg = rotor->number;
  r->displacement_mm[0] = d[0];
  r->displacement_mm[1] = d[1];
  return true;
}

bool ParseMotorRequestText(const std::string &text, MotorRequest *out, std::string *error)
{
  error->clear();
  Json root;
  return ParseJson(text, &root, error) && ParseMotorRequest(root, out, error);
}

bool ParseBatchRequest(const std::string &text, BatchRequest *out, std::string *error)
{
  error->clear();
  Json root;
  if (!ParseJson(text, &root, error) ||
      !ExactMembers(root, {"protocol", "max_items_per_chunk", "items"}, error))
    return false;
  const Json *protocol = Member(root, "protocol", Json::Type::kString, error);
  const Json *chunk = Member(root, "max_items_per_chunk", Json::Type::kNumber, error);
  const Json *items = Member(root, "items", Json::Type::kArray, error);
  if (protocol != nullptr && (protocol->text == "gpu_femm_motor_batch_v2" ||
                              protocol->text == "gpu_femm_motor_batch_v3" ||
                              protocol->text == "gpu_femm_motor_batch_v4")) {
    *error = "unsupported motor batch protocol: tangent, sliding-band and sector batches";
    return false;
  }
  int32_t max_items = 0;
  if (protocol == nullptr || chunk == nullptr || items == nullptr ||
      protocol->text != "gpu_femm_motor_batch_v1" || !Int32Of(*chunk, &max_items) ||
      max_items < 1 || max_items > 4096 || items->items.empty() || items->items.size() > 4096)
    return Fail(error, "invalid gpu_femm_motor_batch_v1 request");
  out->protocol = protocol->text;
  out->max_items_per_chunk = max_items;
  out->items.clear();
  std::set<std::string> ids;
  for (const Json &item : items->items) {
    if (!ExactMembers(item, {"task_id", "request"}, error))
      return false;
    const Json *id = Member(item, "task_id", Json::Type::kString, error);
    const Json *sample = Member(item, "request", Json::Type::kObject, error);
    BatchItem parsed;
    if (id == nullptr || sample == nullptr || !TaskIdValid(id->text) || ids.count(id->text) ||
        !ParseMotorRequest(*sample, &parsed.request, error))
      return Fail(error, "invalid or duplicate batch task_id");
    if (parsed.request.protocol != "gpu_femm_motor_sample_v1") {
      *error = "motor batch and item protocol versions must match";
      return false;
    }
    parsed.task_id = id->text;
    ids.insert(parsed.task_id);
    out->items.push_back(std::move(parsed));
  }
  return true;
}

bool SampleMatchesArtifact(const SampleRequest &r, const Artifact &a)
{
  if (r.Generic() != a.generic || !a.neutral)
    return false;
  if (r.currents.size() != a.currents.size())
    return false;
  const bool groups = r.force_group >= 0;
  if (groups && (!a.HasGroup(r.force_group) || !a.HasGroup(r.air_group)))
    return false;
  const bool post = r.compute_force_torque || !r.airgap_angles_deg.empty();
  // The magnetostatic protocol exposes the field solve only.
  if (r.Generic())
    return !post && r.force_group == -1 && r.air_group == -1 && r.airgap_radius_mm == 0.0 &&
           r.airgap_angles_deg.empty() && r.sliding_band_angle_deg == 0.0;
  // A conforming artifact is already posed: rotation needs a sliding band, and the
  // weighted-stress postprocess needs a selected / air group pair.
  return r.sliding_band_angle_deg == 0.0 && (!post || groups);
}

bool MotorMatchesArtifact(const MotorRequest &r, const Artifact &a)
{
  if (a.schema != "gpu_femm_mesh_v1" && a.schema != "gpu_femm_mesh_v2")
    return false;
  // A sliding-band artifact is the reference pose: any rotor angle, no displacement; its
  // postprocess needs the selected / air groups and probes inside every element's gap.
  const bool pose = a.sliding_band ? r.displacement_mm[0] == 0.0 && r.displacement_mm[1] == 0.0
                                   : r.rotor_angle_deg == a.rotor_angle_deg &&
                                         r.displacement_mm[0] == a.displacement_mm[0] &&
                                         r.displacement_mm[1] == a.displacement_mm[1];
  bool post = true;
  if (a.sliding_band) {
    post = a.HasGroup(r.selected_group) && a.HasGroup(r.air_group);
    if (post && !r.airgap_angles_deg.empty()) {
      post = !a.air_gaps.empty();
      for (const femm_air_gap_t &age : a.air_gaps) {
        const double inner_mm = 1000.0 * age.inner_radius_m, outer_mm = 1000.0 * age.outer_radius_m;
        const double tolerance_mm = 1e-9 * std::max(1.0, std::abs(outer_mm));
        post = post && r.airgap_radius_mm >= inner_mm - tolerance_mm &&
               r.airgap_radius_mm <= outer_mm + tolerance_mm;
      }
    }
  }
  return r.base_sha256 == a.base_sha256 && r.source_sha256 == a.source_sha256 && pose && post &&
         r.currents.size() == a.currents.size();
}

void WriteSampleResponse(std::ostream &out, int status, int solve_status, bool post_requested,
                         int post_status, const SampleRequest *request, const Solution *solution,
                         const Loads *loads, const std::string &diagnostic,
                         const Artifact *artifact)
{
  const bool solved = solve_status == FEMM_OK && request != nullptr && solution != nullptr;
  const bool posted = solved && post_requested && post_status == FEMM_OK && loads != nullptr;
  const std::string message = status == FEMM_OK ? "" : (diagnostic.empty() ? Name(status) : diagnostic);
  const std::string protocol = request ? request->protocol : "gpu_femm_planar_dc_sample_v1";
  const bool generic = protocol == "gpu_femm_magnetostatic_sample_v1";
  const bool axisym = artifact != nullptr && artifact->problem_type == FEMM_AXISYMMETRIC;
  out << "{\n  \"protocol\": \"" << protocol << "\",\n  \"mesh_artifact_sha256\": \""
      << (request ? request->artifact_sha256 : "") << "\",\n";
  if (generic)
    out << "  \"problem_type\": \"" << (axisym ? "axisymmetric" : "planar")
        << "\",\n  \"field_components\": [" << (axisym ? "\"Br\", \"Bz\"" : "\"Bx\", \"By\"")
        << "],\n  \"node_potential_quantity\": \""
        << (axisym ? "poloidal_flux_function" : "magnetic_vector_potential")
        << "\",\n  \"node_potential_unit\": \"" << (axisym ? "Wb" : "Wb/m") << "\",\n";
  out << "  \"status\": \"" << (status == FEMM_OK ? "PASS" : "FAIL") << "\",\n"
      << "  \"solve_status\": \"" << Name(solve_status) << "\",\n"
      << "  \"postprocess_status\": \""
      << (!post_requested ? "NOT_REQUESTED"
                          : (solve_status == FEMM_OK ? Name(post_status) : std::string("NOT_RUN")))
      << "\",\n  \"error_identifier\": \""
      << (status == FEMM_OK ? "" : "GPU_FEMM_" + Name(status)) << "\",\n"
      << "  \"error_message\": \"" << EscapeJson(message) << "\",\n";
  if (posted && request->compute_force_torque)
    out << "  \"Fx_N\": " << Num(loads->fx) << ",\n  \"Fy_N\": " << Num(loads->fy)
        << ",\n  \"torque_Nm\": " << Num(loads->torque) << ",\n";
  else
    out << "  \"Fx_N\": null,\n  \"Fy_N\": null,\n  \"torque_Nm\": null,\n";
  const bool include = solved && request->include_field_solution;
  std::vector<double> potential;
  if (include) {
    potential = solution->a;
    if (generic && axisym)   // the poloidal flux function 2 pi r A
      for (size_t v = 0; v < potential.size(); ++v)
        potential[v] *= 6.283185307179586476925286766559005768 * artifact->node_xy[2 * v];
  }
  out << "  \"actual_circuit_currents_A\": [" << (solved ? List(solution->currents) : "")
      << "],\n  \"circuit_flux_linkage_Wb\": [" << (solved ? List(solution->flux) : "")
      << "],\n  \"airgap_sample_angles_deg\": [" << (posted ? List(request->airgap_angles_deg) : "")
      << "],\n  \"airgap_radial_flux_density_T\": [" << (posted ? List(loads->radial_b) : "")
      << "],\n  \"mesh_node_count\": " << (solved ? solution->a.size() : 0)
      << ",\n  \"mesh_element_count\": " << (solved ? solution->bx.size() : 0)
      << ",\n  \"field_solution_included\": " << (include ? "true" : "false") << ",\n  \""
      << (generic ? "node_potential" : "node_A_Wb_per_m") << "\": [" << List(potential)
      << "],\n  \"" << (generic ? "element_B_component_1_T" : "element_Bx_T") << "\": ["
      << (include ? List(solution->bx) : "") << "],\n  \""
      << (generic ? "element_B_component_2_T" : "element_By_T") << "\": ["
      << (include ? List(solution->by) : "") << "],\n  \"convergence\": {\"iterations\": "
      << (solution ? solution->iterations : 0) << ", \"residual_l2\": "
      << (solution && std::isfinite(solution->residual_l2) ? Num(solution->residual_l2) : "null")
      << "}\n}\n";
}

void WriteMotorResponse(std::ostream &out, int status, const MotorRequest *request,
                        const Solution *solution, const Loads *loads, const char *)
{
  out << "{\n  \"protocol\": \"" << (request ? request->protocol : "gpu_femm_motor_sample_v1")
      << "\",\n  \"mesh_artifact_sha256\": \"" << (request ? request->artifact_sha256 : "")
      << "\",\n  \"base_motor_fem_sha256\": \"" << (request ? request->base_sha256 : "")
      << "\",\n  \"source_fem_sha256\": \"" << (request ? request->source_sha256 : "")
      << "\",\n  \"status\": \"" << (status == FEMM_OK ? "PASS" : "FAIL")
      << "\",\n  \"solve_status\": \"" << Name(status) << "\",\n  \"error_identifier\": \""
      << (status == FEMM_OK ? "" : "GPU_FEMM_" + Name(status)) << "\",\n  \"error_message\": \""
      << (status == FEMM_OK ? "" : Name(status)) << "\",\n";
  if (status == FEMM_OK && request && solution && loads) {
    out << "  \"Fx_N\": " << Num(loads->fx) << ",\n  \"Fy_N\": " << Num(loads->fy)
        << ",\n  \"torque_Nm\": " << Num(loads->torque) << ",\n  \"actual_circuit_currents_A\": ["
        << List(solution->currents) << "],\n  \"circuit_flux_linkage_Wb\": ["
        << List(solution->flux) << "],\n  \"airgap_sample_angles_deg\": ["
        << List(request->airgap_angles_deg) << "],\n  \"airgap_radial_flux_density_T\": ["
        << List(loads->radial_b) << "],\n  \"mesh_element_count\": " << solution->bx.size()
        << ",\n  \"convergence\": {\"iterations\": " << solution->iterations
        << ", \"residual_l2\": " << Num(solution->residual_l2) << "}\n";
  } else {
    out << "  \"Fx_N\": null,\n  \"Fy_N\": null,\n  \"torque_Nm\": null,\n"
        << "  \"actual_circuit_currents_A\": [],\n  \"circuit_flux_linkage_Wb\": [],\n"
        << "  \"airgap_sample_angles_deg\": [],\n  \"airgap_radial_flux_density_T\": [],\n"
        << "  \"mesh_element_count\": 0,\n  \"convergence\": {\"iterations\": "
        << (solution ? solution->iterations : 0) << ", \"residual_l2\": "
        << (solution && std::isfinite(solution->residual_l2) ? Num(solution->residual_l2) : "null")
        << "}\n";
  }
  out << "}";
}

void WriteBatchResponse(std::ostream &out, const BatchRequest &request,
                        const std::vector<BatchItemOutcome> &items, const BatchCounters &c)
{
  bool pass = request.items.size() == items.size();
  for (const BatchItemOutcome &item : items)
    pass = pass && item.status == FEMM_OK;
  out << "{\n  \"protocol\": \"" << request.protocol << "\",\n  \"status\": \""
      << (pass ? "PASS" : "FAIL") << "\",\n  \"solve_status\": \""
      << (pass ? "PASS" : "PARTIAL_FAILURE") << "\",\n  \"error_identifier\": \""
      << (pass ? "" : "GPU_FEMM_BATCH_PARTIAL_FAILURE") << "\",\n  \"error_message\": \""
      << (pass ? "" : "one or more batch items failed") << "\",\n"
      << "  \"cache\": {\"mesh_cache_hits\": " << c.mesh_cache_hits
      << ", \"chunk_count\": " << c.chunk_count << ", \"requested_chunk_size\": "
      << c.requested_chunk << ", \"effective_chunk_size\": " << c.effective_chunk
      << ", \"actual_parallel_width\": " << c.parallel_width << ", \"csr_symbolic_cache_hits\": "
      << c.symbolic_cache_hits << ", \"batched_pcg_launches\": " << c.batched_launches
      << "},\n  \"items\": [\n";
  for (size_t k = 0; k < request.items.size(); ++k) {
    const BatchItemOutcome &o = items[k];
    out << "    {\"task_id\": \"" << request.items[k].task_id << "\", \"response\": ";
    WriteMotorResponse(out, o.status, &request.items[k].request, o.solved ? &o.solution : nullptr,
                       o.status == FEMM_OK ? &o.loads : nullptr, "");
    out << "}" << (k + 1 == request.items.size() ? "\n" : ",\n");
  }
  out << "  ]\n}\n";
}

}  // namespace host
}  // namespace femm
