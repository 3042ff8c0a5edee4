// Command-line adapters of the wire contract over the library's C ABI: single sample
// (--solve), single motor sample (--motor-single-sample) and the shared-geometry motor batch
// (--motor-batch) (reference gpu_solver/src/app/cli.cpp:20-181, sample_adapter.cpp,
// motor_batch.cpp).  Every solve is the reference algorithm (FP64 Jacobi PCG, the Newton
// driver with the production sample options and the bounded recovery); a batch solves all
// items that share an artifact in one batched launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "api_guard.hpp"
#include "artifact.hpp"
#include "femm_gpu_amd.h"
#include "protocol.hpp"
#include "sliding_band.hpp"

namespace femm
{
namespace host
{
namespace
{

constexpr double kPi = 3.141592653589793238462643383279502884;

femm_newton_options_t ProductionOptions()
{
  // ProductionSampleOptions (physics/nonlinear_model.cpp:195-205).
  femm_newton_options_t o{};
  o.relative_tolerance = 1e-8;
  o.linear_relative_tolerance = 1e-10;
  o.max_newton_iterations = 128;
  o.max_linear_iterations = 16384;
  o.recover_nonconvergence = 1;
  o.adaptive_linear_tolerance = 1;
  o.require_equation_residual = 0;
  return o;
}

bool ReadFile(const std::string &path, std::string *bytes)
{
  std::ifstream in(path, std::ios::binary);
  if (!in)
    return false;
  std::ostringstream s;
  s << in.rdbuf();
  if (!in && !in.eof())
    return false;
  *bytes = s.str();
  return true;
}

std::string PathKey(const std::string &path)
{
  std::error_code ec;
  const std::filesystem::path p = std::filesystem::weakly_canonical(path, ec);
  return ec ? path : p.generic_string();
}

bool SameObject(const std::string &a, const std::string &b)
{
  std::error_code ec;
  if (std::filesystem::exists(a, ec) && std::filesystem::exists(b, ec) &&
      std::filesystem::equivalent(a, b, ec) && !ec)
    return true;
  return PathKey(a) == PathKey(b);
}

// Writes through `write` to a unique temporary beside `path`, then renames it into place.
template <class F>
bool Publish(const std::string &path, F &&write)
{
  static std::atomic<uint64_t> sequence{0};
  const std::string tmp = path + ".tmp." +
                          std::to_string(std::chrono::steady_clock::now().time_since_epoch().count()) +
                          "." + std::to_string(sequence.fetch_add(1));
  bool ok = false;
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (out) {
      write(out);
      out.flush();
      ok = static_cast<bool>(out);
    }
  }
  std::error_code ec;
  if (ok) {
    std::filesystem::rename(tmp, path, ec);
    ok = !ec;
  }
  if (!ok)
    std::filesystem::remove(tmp, ec);
  return ok;
}

// One artifact, loaded and checked once per batch.
struct Load {
  int status = FEMM_INTERNAL_ERROR;
  std::string sha256, error;
  Artifact artifact;
};

void LoadArtifact(const std::string &path, Load *load)
{
  std::string bytes;
  if (!ReadFile(path, &bytes)) {
    load->status = FEMM_INPUT_IO;
    load->error = "cannot open mesh artifact";
    return;
  }
  load->sha256 = Sha256Hex(bytes);
  load->status = ParseArtifact(bytes, &load->artifact, &load->error) ? FEMM_OK
                                                                     : FEMM_INVALID_ARGUMENT;
}

struct PlanHandle {
  femm_plan *plan = nullptr;
  ~PlanHandle()
  {
    if (plan)
      femm_plan_destroy(plan);
  }
};

// The plan of an artifact posed for a request: a sliding-band artifact's air-gap elements
// turned to the requested rotor angle (ApplySlidingBandRotorAngle,
// io/sample_request.cpp:448-496), any other artifact as parsed.
int CreatePlan(const Artifact &art, double rotor_angle_deg, PlanHandle *h)
{
  if (!art.sliding_band) {
    const femm_model_t m = art.View();
    return femm_plan_create(&m, &h->plan);
  }
  std::vector<femm_air_gap_t> gaps;
  std::vector<femm_age_point_t> points;
  if (!RotateAirGaps(art.air_gaps.data(), static_cast<int32_t>(art.air_gaps.size()),
                     art.rotor_angle_deg, rotor_angle_deg, &gaps, &points))
    return FEMM_INVALID_ARGUMENT;
  const femm_model_t m = art.ViewWith(gaps);
  return femm_plan_create(&m, &h->plan);
}

// Solves `currents.size()` items on `plan` (one batched call); fills solutions and statuses.
int SolveBatch(femm_plan *plan, const Artifact &art, const std::vector<std::vector<double>> &currents,
               std::vector<Solution> *sol, std::vector<int> *status)
{
  const femm_model_t m = art.View();
  const size_t B = currents.size(), N = m.node_count, T = m.triangle_count;
  const size_t C = std::max<int32_t>(m.circuit_count, 1);
  std::vector<double> cur(B * C, 0.0), a(B * N), bx(B * T), by(B * T), flux(B * C);
  for (size_t b = 0; b < B; ++b)
    std::copy(currents[b].begin(), currents[b].end(), cur.begin() + b * C);
  std::vector<femm_item_result_t> res(B);
  const femm_newton_options_t o = ProductionOptions();
  const int st = femm_newton_batch(plan, static_cast<int32_t>(B), cur.data(), &o, nullptr,
                                   nullptr, a.data(), bx.data(), by.data(), flux.data(),
                                   res.data());
  sol->assign(B, Solution{});
  status->assign(B, st);
  if (st != FEMM_OK)
    return st;
  for (size_t b = 0; b < B; ++b) {
    Solution &s = (*sol)[b];
    s.a.assign(a.begin() + b * N, a.begin() + (b + 1) * N);
    s.bx.assign(bx.begin() + b * T, bx.begin() + (b + 1) * T);
    s.by.assign(by.begin() + b * T, by.begin() + (b + 1) * T);
    s.flux.assign(flux.begin() + b * C, flux.begin() + b * C + m.circuit_count);
    s.currents = currents[b];
    s.iterations = res[b].newton_iterations;
    s.residual_l2 = res[b].relative_step;
    (*status)[b] = res[b].status;
  }
  return FEMM_OK;
}

// Weighted-stress loads and radial air-gap B of the given items (FrozenPostprocess with
// max_mask_iterations 4096, as the adapters configure it).
int Postprocess(femm_plan *plan, int32_t selected_group, int32_t air_group, double radius_mm,
                const std::vector<double> &angles_deg, const std::vector<const Solution *> &items,
                std::vector<Loads> *loads)
{
  std::vector<double> angles(angles_deg.size());
  for (size_t k = 0; k < angles.size(); ++k)
    angles[k] = angles_deg[k] * kPi / 180.0;
  femm_postprocess_options_t po{};
  po.selected_label = -1;
  po.air_label = -1;
  po.selected_group = selected_group;
  po.air_group = air_group;
  po.airgap_radius_m = radius_mm * 1e-3;
  po.angles_rad = angles.empty() ? nullptr : angles.data();
  po.angle_count = static_cast<int32_t>(angles.size());
  po.max_mask_iterations = 4096;
  po.mask_relative_tolerance = 1e-12;
  int st = femm_plan_set_postprocess(plan, &po);
  if (st != FEMM_OK)
    return st;
  const size_t B = items.size(), N = items.empty() ? 0 : items[0]->a.size();
  std::vector<double> states(B * N), out(B * 3), radial(B * std::max<size_t>(angles.size(), 1));
  for (size_t b = 0; b < B; ++b)
    std::copy(items[b]->a.begin(), items[b]->a.end(), states.begin() + b * N);
  st = femm_postprocess_batch(plan, static_cast<int32_t>(B), states.data(), out.data(),
                              radial.data());
  if (st != FEMM_OK)
    return st;
  loads->assign(B, Loads{});
  for (size_t b = 0; b < B; ++b) {
    Loads &l = (*loads)[b];
    l.fx = out[3 * b];
    l.fy = out[3 * b + 1];
    l.torque = out[3 * b + 2];
    l.radial_b.assign(radial.begin() + b * std::max<size_t>(angles.size(), 1),
                      radial.begin() + b * std::max<size_t>(angles.size(), 1) + angles.size());
  }
  return FEMM_OK;
}

// ---- --solve ------------------------------------------------------------------------------

int SampleDirect(const SampleRequest *request, const std::string &parse_error, std::ostream &out)
{
  if (request == nullptr) {
    WriteSampleResponse(out, FEMM_INVALID_ARGUMENT, FEMM_INVALID_ARGUMENT, false, FEMM_OK, nullptr,
                        nullptr, nullptr, parse_error, nullptr);
    std::cerr << "FAIL planar-DC request: " << parse_error << '\n';
    return 1;
  }
  Load load;
  LoadArtifact(request->artifact_path, &load);
  int status = FEMM_OK;
  std::string error;
  if (load.status == FEMM_INPUT_IO) {
    status = FEMM_INPUT_IO;
    error = load.error;
  } else if (load.sha256 != request->artifact_sha256) {
    status = FEMM_INVALID_ARGUMENT;
    error = "mesh artifact SHA mismatch";
  } else if (load.status != FEMM_OK) {
    status = FEMM_INVALID_ARGUMENT;
    error = load.error;
  } else if (!SampleMatchesArtifact(*request, load.artifact)) {
    status = FEMM_INVALID_ARGUMENT;
    error = "artifact physics, circuit count, group, or sliding-band request mismatch";
  }
  const bool post_requested = request->compute_force_torque || !request->airgap_angles_deg.empty();
  int solve_status = status, post_status = FEMM_OK;
  std::vector<Solution> sol;
  std::vector<Loads> loads;
  bool solved = false;
  if (status == FEMM_OK) {
    PlanHandle h;
    const femm_model_t m = load.artifact.View();
    solve_status = femm_plan_create(&m, &h.plan);
    std::vector<int> st;
    if (solve_status == FEMM_OK) {
      solve_status = SolveBatch(h.plan, load.artifact, {request->currents}, &sol, &st);
      solved = solve_status == FEMM_OK;
      if (solved)
        solve_status = st[0];
    }
    if (solve_status == FEMM_OK && post_requested)
      post_status = Postprocess(h.plan, request->force_group, request->air_group,
                                request->airgap_radius_mm, request->airgap_angles_deg,
                                {&sol[0]}, &loads);
    status = solve_status;
    if (solve_status == FEMM_OK && post_requested)
      status = post_status;
  }
  WriteSampleResponse(out, status, solve_status, post_requested, post_status, request,
                      solved ? &sol[0] : nullptr,
                      post_status == FEMM_OK && !loads.empty() ? &loads[0] : nullptr, error,
                      load.status == FEMM_OK ? &load.artifact : nullptr);
  if (status != FEMM_OK) {
    std::cerr << "FAIL magnetostatic solve: " << (error.empty() ? femm_status_name(status) : error)
              << '\n';
    return 1;
  }
  return 0;
}

int SampleAdapter(const std::string &exe, const std::string &request_path,
                  const std::string &response_path)
{
  if (SameObject(exe, response_path)) {
    std::cerr << "FAIL magnetostatic solve: response must not overwrite solver executable\n";
    return 1;
  }
  if (SameObject(request_path, response_path)) {
    std::cerr << "FAIL magnetostatic solve: response must not overwrite request\n";
    return 1;
  }
  std::string text, error;
  SampleRequest request;
  bool parsed = false;
  if (!ReadFile(request_path, &text))
    error = "cannot open request";
  else
    parsed = ParseSampleRequest(text, &request, &error);
  if (parsed && SameObject(request.artifact_path, response_path)) {
    std::cerr << "FAIL magnetostatic solve: response must not overwrite mesh artifact\n";
    return 1;
  }
  int code = 1;
  if (!Publish(response_path, [&](std::ostream &out) {
        code = SampleDirect(parsed ? &request : nullptr, error, out);
      })) {
    std::cerr << "FAIL magnetostatic solve: could not atomically publish response\n";
    return 1;
  }
  return code;
}

// ---- --motor-single-sample ----------------------------------------------------------------

int MotorSampleAdapter(const std::string &request_path, const std::string &response_path)
{
  std::string text, error;
  MotorRequest request;
  if (!ReadFile(request_path, &text) || !ParseMotorRequestText(text, &request, &error)) {
    Publish(response_path, [&](std::ostream &out) {
      WriteMotorResponse(out, FEMM_INVALID_ARGUMENT, nullptr, nullptr, nullptr, "");
      out << '\n';
    });
    std::cerr << "FAIL motor single-sample request: " << error << '\n';
    return 1;
  }
  Load load;
  LoadArtifact(request.artifact_path, &load);
  int status = FEMM_OK;
  if (load.status == FEMM_INPUT_IO) {
    status = FEMM_INPUT_IO;
    error = load.error;
  } else if (load.sha256 != request.artifact_sha256) {
    status = FEMM_INVALID_ARGUMENT;
    error = "mesh artifact SHA mismatch";
  } else if (load.status != FEMM_OK) {
    status = FEMM_INVALID_ARGUMENT;
    error = load.error;
  } else if (!MotorMatchesArtifact(request, load.artifact)) {
    status = FEMM_INVALID_ARGUMENT;
    error = "artifact identity, pose, or circuit-count mismatch";
  }
  std::vector<Solution> sol;
  std::vector<Loads> loads;
  bool solved = false;
  if (status == FEMM_OK) {
    PlanHandle h;
    status = CreatePlan(load.artifact, request.rotor_angle_deg, &h);
    std::vector<int> st;
    if (status == FEMM_OK) {
      status = SolveBatch(h.plan, load.artifact, {request.currents}, &sol, &st);
      solved = status == FEMM_OK;
      if (solved)
        status = st[0];
    }
    if (status == FEMM_OK)
      status = Postprocess(h.plan, request.selected_group, request.air_group,
                           request.airgap_radius_mm, request.airgap_angles_deg, {&sol[0]}, &loads);
  }
  const bool written = Publish(response_path, [&](std::ostream &out) {
    WriteMotorResponse(out, status, &request, solved ? &sol[0] : nullptr,
                       status == FEMM_OK ? &loads[0] : nullptr, "");
    out << '\n';
  });
  if (!written)
    return 1;
  if (status != FEMM_OK) {
    std::cerr << "FAIL motor single-sample: " << (error.empty() ? femm_status_name(status) : error)
              << '\n';
    return 1;
  }
  return 0;
}

// ---- --motor-batch --------------------------------------------------------------------------

// Items a batched launch may hold: the request's chunk cap, bounded by the solve workspace
// that fits in free device memory (about 2 x (8 nnz + 96 T + 72 n + 8 N) bytes per item with
// a 1 GiB reserve; on a 288 GB MI355X this rarely binds).
int32_t EffectiveChunk(int32_t requested, const Artifact &art)
{
  size_t free_bytes = 0, total_bytes = 0;
  if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess)
    return requested;
  const size_t N = art.node_xy.size() / 2, T = art.tri_mat.size();
  const size_t per_item = 2 * (8 * 7 * N + 96 * T + 72 * N + 8 * N) + 4096;
  const size_t reserve = size_t(1) << 30;
  const size_t fit = free_bytes > reserve ? (free_bytes - reserve) / per_item : 1;
  return static_cast<int32_t>(std::max<size_t>(1, std::min<size_t>(requested, fit)));
}

int BatchAdapter(const std::string &request_path, const std::string &response_path)
{
  std::string text, error;
  BatchRequest request;
  if (!ReadFile(request_path, &text) || !ParseBatchRequest(text, &request, &error)) {
    std::cerr << "FAIL motor batch request: " << error << '\n';
    return 1;   // no trusted item identity exists to write a response
  }
  const size_t items = request.items.size();
  // Artifacts: one load per canonical path, in first-seen order.
  std::map<std::string, size_t> slot_of_path;
  std::vector<std::unique_ptr<Load>> loads;
  std::vector<size_t> slot(items);
  for (size_t k = 0; k < items; ++k) {
    const std::string key = PathKey(request.items[k].request.artifact_path);
    auto found = slot_of_path.find(key);
    if (found == slot_of_path.end()) {
      found = slot_of_path.emplace(key, loads.size()).first;
      loads.push_back(std::make_unique<Load>());
      LoadArtifact(request.items[k].request.artifact_path, loads.back().get());
    }
    slot[k] = found->second;
  }
  std::vector<BatchItemOutcome> out(items);
  BatchCounters counters;
  counters.requested_chunk = request.max_items_per_chunk;
  const Load &first = *loads[slot[0]];
  counters.effective_chunk = first.status == FEMM_OK
                                 ? EffectiveChunk(request.max_items_per_chunk, first.artifact)
                                 : request.max_items_per_chunk;
  // Plans stay cached per artifact (and rotor angle, for a sliding band) across chunks (the
  // symbolic CSR is built once per operator, as the reference's operator groups,
  // app/motor_batch.cpp:190-262).
  using PlanKey = std::pair<size_t, double>;
  std::map<PlanKey, std::unique_ptr<PlanHandle>> plans;
  for (size_t lo = 0; lo < items; lo += counters.effective_chunk) {
    const size_t hi = std::min(items, lo + static_cast<size_t>(counters.effective_chunk));
    ++counters.chunk_count;
    // Items that pass their artifact checks, grouped by operator in first-seen order.
    std::vector<PlanKey> group_order;
    std::map<PlanKey, std::vector<size_t>> groups;
    for (size_t k = lo; k < hi; ++k) {
      const MotorRequest &r = request.items[k].request;
      const Load &load = *loads[slot[k]];
      int st = load.status;
      if (st == FEMM_OK && (load.sha256 != r.artifact_sha256 || !MotorMatchesArtifact(r, load.artifact)))
        st = FEMM_INVALID_ARGUMENT;
      out[k].status = st;
      if (st != FEMM_OK)
        continue;
      const PlanKey key{slot[k], load.artifact.sliding_band ? r.rotor_angle_deg : 0.0};
      if (!groups.count(key))
        group_order.push_back(key);
      groups[key].push_back(k);
    }
    for (const PlanKey &key : group_order) {
      const std::vector<size_t> &group = groups[key];
      const Artifact &art = loads[key.first]->artifact;
      auto cached = plans.find(key);
      int st = FEMM_OK;
      if (cached == plans.end()) {
        auto h = std::make_unique<PlanHandle>();
        st = CreatePlan(art, key.second, h.get());
        if (st == FEMM_OK)
          cached = plans.emplace(key, std::move(h)).first;
        counters.mesh_cache_hits += static_cast<int32_t>(group.size()) - 1;
      } else {
        counters.mesh_cache_hits += static_cast<int32_t>(group.size());
        ++counters.symbolic_cache_hits;
      }
      if (st != FEMM_OK) {
        for (const size_t k : group)
          out[k].status = st;
        continue;
      }
      femm_plan *plan = cached->second->plan;
      counters.parallel_width = std::max(counters.parallel_width, static_cast<int32_t>(group.size()));
      std::vector<std::vector<double>> currents;
      for (const size_t k : group)
        currents.push_back(request.items[k].request.currents);
      std::vector<Solution> sol;
      std::vector<int> status;
      st = SolveBatch(plan, art, currents, &sol, &status);
      ++counters.batched_launches;
      for (size_t j = 0; j < group.size(); ++j) {
        BatchItemOutcome &o = out[group[j]];
        o.status = status[j];
        if (st == FEMM_OK) {
          o.solution = std::move(sol[j]);
          o.solved = true;
        }
      }
      // Postprocess per distinct (selected, air, radius, angles) among the solved items.
      std::vector<size_t> pending;
      for (const size_t k : group)
        if (out[k].status == FEMM_OK)
          pending.push_back(k);
      while (!pending.empty()) {
        const MotorRequest &lead = request.items[pending[0]].request;
        std::vector<size_t> same, rest;
        for (const size_t k : pending) {
          const MotorRequest &r = request.items[k].request;
          (r.selected_group == lead.selected_group && r.air_group == lead.air_group &&
                   r.airgap_radius_mm == lead.airgap_radius_mm &&
                   r.airgap_angles_deg == lead.airgap_angles_deg
               ? same
               : rest)
              .push_back(k);
        }
        std::vector<const Solution *> states;
        for (const size_t k : same)
          states.push_back(&out[k].solution);
        std::vector<Loads> l;
        const int ps = Postprocess(plan, lead.selected_group, lead.air_group,
                                   lead.airgap_radius_mm, lead.airgap_angles_deg, states, &l);
        for (size_t j = 0; j < same.size(); ++j) {
          out[same[j]].status = ps;
          if (ps == FEMM_OK)
            out[same[j]].loads = std::move(l[j]);
        }
        pending.swap(rest);
      }
    }
  }
  if (!Publish(response_path, [&](std::ostream &o) { WriteBatchResponse(o, request, out, counters); }))
    return 1;
  return std::all_of(out.begin(), out.end(), [](const BatchItemOutcome &o) { return o.status == FEMM_OK; })
             ? 0
             : 1;
}

int MeshArtifactCheck(const std::string &path)
{
  std::string bytes, error;
  Artifact a;
  if (!ReadFile(path, &bytes)) {
    std::cerr << "FAIL mesh artifact: cannot open " << path << '\n';
    return 1;
  }
  if (!ParseArtifact(bytes, &a, &error)) {
    std::cerr << "FAIL mesh artifact: " << error << '\n';
    return 1;
  }
  std::cout << "{\"schema_version\": \"" << a.schema << "\", \"sha256\": \"" << Sha256Hex(bytes)
            << "\", \"nodes\": " << a.node_xy.size() / 2 << ", \"triangles\": " << a.tri_mat.size()
            << ", \"materials\": " << a.materials.size() << ", \"circuits\": " << a.currents.size()
            << "}\n";
  return 0;
}

void Capabilities()
{
  std::cout << "{\n  \"schema_version\": \"gpu_femm_capabilities_v1\",\n"
            << "  \"sample_protocol\": \"gpu_femm_planar_dc_sample_v1\",\n"
            << "  \"sample_protocols\": [\"gpu_femm_planar_dc_sample_v1\", "
               "\"gpu_femm_magnetostatic_sample_v1\"],\n"
            << "  \"motor_sample_protocols\": [\"gpu_femm_motor_sample_v1\"],\n"
            << "  \"motor_batch_protocols\": [\"gpu_femm_motor_batch_v1\"],\n"
            << "  \"mesh_schemas\": [\"gpu_femm_planar_dc_mesh_v1\", "
               "\"gpu_femm_magnetostatic_mesh_v1\", \"gpu_femm_mesh_v1\"],\n"
            << "  \"problem_types\": [\"planar\", \"axisymmetric\"],\n"
            << "  \"frequency_hz\": [0],\n  \"precision\": \"fp64\",\n"
            << "  \"linear_solvers\": [\"pcg\"],\n  \"default_linear_solver\": \"pcg\",\n"
            << "  \"preconditioners\": [\"jacobi\", \"block-jacobi\"],\n"
            << "  \"default_preconditioner\": \"jacobi\",\n"
            << "  \"nonlinear_bh\": true,\n  \"permanent_magnets\": true,\n"
            << "  \"native_sliding_band\": false,\n"
            << "  \"general_periodic_node_constraints\": true,\n"
            << "  \"full_field_output\": true,\n  \"ac_complex\": false,\n"
            << "  \"axisymmetric\": true,\n  \"axisymmetric_assembly\": \"device\",\n"
            << "  \"generic_force_airgap_postprocessing\": false,\n"
            << "  \"device\": \"gfx950\"\n}\n";
}

int Usage()
{
  std::cerr << "Usage: femm_gpu_amd --solve <request.json> <response.json>"
            << " | --motor-single-sample <request.json> <response.json>"
            << " | --motor-batch <request.json> <response.json>"
            << " | --mesh-artifact <artifact.json> | --capabilities\n";
  return 2;
}

}  // namespace
}  // namespace host
}  // namespace femm

extern "C" int femm_cli_main(int argc, const char *const *argv)
{
  using namespace femm::host;
  try {
    if (argc < 1 || argv == nullptr)
      return Usage();
    const std::string exe = argv[0] ? argv[0] : "";
    const std::string cmd = argc >= 2 && argv[1] ? argv[1] : "";
    if (argc == 4 && cmd == "--solve")
      return SampleAdapter(exe, argv[2], argv[3]);
    if (argc == 4 && cmd == "--motor-single-sample")
      return MotorSampleAdapter(argv[2], argv[3]);
    if (argc == 4 && cmd == "--motor-batch")
      return BatchAdapter(argv[2], argv[3]);
    if (argc == 3 && cmd == "--mesh-artifact")
      return MeshArtifactCheck(argv[2]);
    if (argc == 2 && cmd == "--capabilities") {
      Capabilities();
      return 0;
    }
    return Usage();
  } catch (...) {
    std::cerr << "FAIL: internal error (" << femm_status_name(femm::StatusOfCurrentException())
              << ")\n";
    return 1;
  }
}
