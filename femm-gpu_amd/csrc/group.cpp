// Group solves: meshes of uneven size in one cooperative launch (femm_group_* of
// include/femm_gpu_amd.h).  Per item the reference Newton driver runs unchanged
// (solvers/nonlinear_solver.cu:183-530); only the launch packing is new.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "api_guard.hpp"
#include "device.hpp"
#include "femm_gpu_amd.h"
#include "handle.hpp"

struct femm_group {
  std::vector<femm_group_entry_t> entries;
  std::vector<femm::GroupEntry> uploaded;   // the table as last copied to the device
  femm::GroupEntry *table = nullptr;        // device copy, entries.size() records
  // Captured launch sequence (femm_group_capture) with its own device copy of the table, so
  // later eager solves with another layout do not change what the graph runs, and the table
  // it was captured with (a replay refuses once a plan's device arrays or workspace moved).
  hipGraphExec_t graph = nullptr;
  femm::GroupEntry *graph_table = nullptr;
  std::vector<femm::GroupEntry> graph_entries;
};

namespace
{

bool SameTable(const std::vector<femm::GroupEntry> &a, const std::vector<femm::GroupEntry> &b)
{
  return a.size() == b.size() &&
         std::memcmp(a.data(), b.data(), sizeof(femm::GroupEntry) * a.size()) == 0;
}

int EntriesValid(int32_t count, const femm_group_entry_t *entries)
{
  if (count <= 0 || entries == nullptr)
    return FEMM_INVALID_ARGUMENT;
  for (int32_t e = 0; e < count; ++e) {
    const femm_group_entry_t &E = entries[e];
    if (E.plan == nullptr || E.items <= 0 || E.partitions < 1 ||
        E.partitions > femm::kMaxPartitions || E.currents == nullptr || E.a == nullptr ||
        E.flux == nullptr || E.results == nullptr)
      return FEMM_INVALID_ARGUMENT;
    for (int32_t f = 0; f < e; ++f)
      if (entries[f].plan == E.plan)   // one workspace per plan
        return FEMM_INVALID_ARGUMENT;
  }
  return FEMM_OK;
}

// Device state of every entry's plan for `mixed`, then the launch table.
int PrepareTable(femm_group *group, bool mixed, bool blocks, std::vector<femm::GroupEntry> *table)
{
  table->clear();
  int32_t first = 0;
  for (const femm_group_entry_t &E : group->entries) {
    int st = femm::EnsurePlanDevice(E.plan);
    if (st == FEMM_OK)
      st = femm::EnsureWork(E.plan->host, E.items, mixed, &E.plan->dev, blocks);
    if (st != FEMM_OK)
      return st;
    femm::GroupEntry g{};
    g.plan = E.plan->dev.plan;
    g.work = E.plan->dev.work;
    g.currents = E.currents;
    g.a = E.a;
    g.flux = E.flux;
    g.results = E.results;
    g.items = E.items;
    g.partitions = E.partitions;
    g.first_block = first;
    first += E.items * E.partitions;
    table->push_back(g);
  }
  return FEMM_OK;
}

void ReleaseGraph(femm_group *group)
{
  if (group->graph != nullptr)
    (void)hipGraphExecDestroy(group->graph);
  if (group->graph_table != nullptr)
    (void)hipFree(group->graph_table);
  group->graph = nullptr;
  group->graph_table = nullptr;
  group->graph_entries.clear();
}

// The device state and launch table of a group solve with config `c`, uploaded to
// group->table when it changed (after the stream's earlier launches finished with it).
int PrepareGroup(femm_group *group, const femm_solve_config_t &c, hipStream_t s,
                 std::vector<femm::GroupEntry> *table)
{
  int st = PrepareTable(group, c.precision == FEMM_PRECISION_MIXED, c.block_size > 1, table);
  if (st != FEMM_OK)
    return st;
  if (SameTable(*table, group->uploaded))
    return FEMM_OK;
  if (hipStreamSynchronize(s) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  if (group->table == nullptr &&
      hipMalloc(&group->table, sizeof(femm::GroupEntry) * group->entries.size()) != hipSuccess) {
    group->table = nullptr;
    return FEMM_GPU_ALLOCATION_FAILED;
  }
  if (hipMemcpy(group->table, table->data(), sizeof(femm::GroupEntry) * table->size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  group->uploaded = *table;
  return FEMM_OK;
}

// The launches of one group solve on `s`: the reference driver followed by the bounded
// recovery, or the safeguarded Newton alone.  Only kernels are enqueued (no copies or
// memsets), so the sequence can be captured into a graph.
int EnqueueGroup(const femm::GroupEntry *device_table, const std::vector<femm::GroupEntry> &table,
                 const femm_newton_options_t &o, const femm_solve_config_t &c, bool safeguarded,
                 hipStream_t s)
{
  int st = femm::LaunchGroup(device_table, table, o, c, safeguarded, false, s);
  if (safeguarded || st != FEMM_OK || !o.recover_nonconvergence || o.max_newton_iterations <= 0)
    return st;
  // Bounded recovery of exhausted items from the plain start, as femm_newton_batch.
  return femm::LaunchGroup(device_table, table, o, c, true, true, s);
}

}  // namespace

extern "C" {

int femm_group_create(int32_t count, const femm_group_entry_t *entries, femm_group **out)
{
  FEMM_API_BEGIN
  if (out == nullptr)
    return FEMM_INVALID_ARGUMENT;
  *out = nullptr;
  int st = EntriesValid(count, entries);
  if (st != FEMM_OK)
    return st;
  auto *group = new (std::nothrow) femm_group;
  if (group == nullptr)
    return FEMM_INTERNAL_ERROR;
  group->entries.assign(entries, entries + count);
  *out = group;
  return FEMM_OK;
  FEMM_API_END
}

void femm_group_destroy(femm_group *group)
{
  FEMM_API_BEGIN
  if (group == nullptr)
    return;
  if (group->table != nullptr || group->graph != nullptr)
    (void)hipDeviceSynchronize();
  if (group->table != nullptr)
    (void)hipFree(group->table);
  ReleaseGraph(group);
  delete group;
  FEMM_API_END_VOID
}

int femm_group_set_partitions(femm_group *group, const int32_t *partitions)
{
  FEMM_API_BEGIN
  if (group == nullptr || partitions == nullptr)
    return FEMM_INVALID_ARGUMENT;
  for (size_t e = 0; e < group->entries.size(); ++e)
    if (partitions[e] < 1 || partitions[e] > femm::kMaxPartitions)
      return FEMM_INVALID_ARGUMENT;
  for (size_t e = 0; e < group->entries.size(); ++e)
    group->entries[e].partitions = partitions[e];
  return FEMM_OK;
  FEMM_API_END
}

int femm_group_autotune(femm_group *group, const femm_solve_config_t *config,
                        const double *work, int32_t *partitions, int32_t *blocks)
{
  FEMM_API_BEGIN
  if (group == nullptr)
    return FEMM_INVALID_ARGUMENT;
  if (!femm::DevicePresent())
    return FEMM_GPU_UNAVAILABLE;
  const femm_solve_config_t c = config ? *config : femm::DefaultSolveConfig();
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  const int capacity = femm::GroupCapacity(c.precision == FEMM_PRECISION_MIXED);
  const size_t count = group->entries.size();
  // ms[e][k]: probe time at 2^k workgroups per item (< 0: not probed yet).
  std::vector<std::vector<float>> ms(count);
  std::vector<double> weight(count);
  std::vector<int> level(count, 0), max_level(count, 0);
  long used = 0;
  for (size_t e = 0; e < count; ++e) {
    const femm_group_entry_t &E = group->entries[e];
    weight[e] = work ? work[e] : std::sqrt(static_cast<double>(std::max(E.plan->host.n, 1)));
    if (!(weight[e] >= 0.0))
      return FEMM_INVALID_ARGUMENT;
    used += E.items;
    int k = 0;
    while ((2 << k) <= femm::kMaxPartitions && (2 << k) <= E.plan->host.nslices &&
           static_cast<long>(E.items) * (2 << k) <= capacity)
      ++k;
    max_level[e] = k;
    ms[e].assign(k + 1, -1.0f);
  }
  if (used > capacity)
    return FEMM_INVALID_ARGUMENT;
  auto probe = [&](size_t e, int k) -> int {
    if (ms[e][k] >= 0.0f)
      return FEMM_OK;
    return femm::ProbeSolve(group->entries[e].plan, group->entries[e].items, c, 1 << k,
                            &ms[e][k]);
  };
  for (size_t e = 0; e < count; ++e) {
    const int st = probe(e, 0);
    if (st != FEMM_OK)
      return st;
  }
  // Greedy min-max: double the partitions of the entry priced slowest while that makes
  // it faster and the workgroups still fit; stop when the slowest cannot improve.
  for (;;) {
    size_t worst = 0;
    double worst_cost = -1.0;
    for (size_t e = 0; e < count; ++e) {
      const double cost = weight[e] * ms[e][level[e]];
      if (cost > worst_cost) {
        worst_cost = cost;
        worst = e;
      }
    }
    const int k = level[worst];
    const long extra = static_cast<long>(group->entries[worst].items) << k;
    if (k >= max_level[worst] || used + extra > capacity)
      break;
    const int st = probe(worst, k + 1);
    if (st != FEMM_OK)
      return st;
    if (!(ms[worst][k + 1] < ms[worst][k]))
      break;
    level[worst] = k + 1;
    used += extra;
  }
  for (size_t e = 0; e < count; ++e) {
    group->entries[e].partitions = 1 << level[e];
    if (partitions != nullptr)
      partitions[e] = 1 << level[e];
  }
  if (blocks != nullptr)
    *blocks = static_cast<int32_t>(used);
  return FEMM_OK;
  FEMM_API_END
}

// One group solve: the reference driver followed by the bounded recovery, or the safeguarded
// Newton alone.
static int GroupSolve(femm_group *group, const femm_newton_options_t *options,
                      const femm_solve_config_t *config, void *stream, bool safeguarded)
{
  if (group == nullptr)
    return FEMM_INVALID_ARGUMENT;
  if (!femm::DevicePresent())
    return FEMM_GPU_UNAVAILABLE;
  const femm_newton_options_t o = options ? *options : femm::DefaultNewtonOptions();
  const femm_solve_config_t c = config ? *config : femm::DefaultSolveConfig();
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<femm::GroupEntry> table;
  const int st = PrepareGroup(group, c, s, &table);
  if (st != FEMM_OK)
    return st;
  return EnqueueGroup(group->table, table, o, c, safeguarded, s);
}

int femm_group_capture(femm_group *group, const femm_newton_options_t *options,
                       const femm_solve_config_t *config, int32_t safeguarded)
{
  FEMM_API_BEGIN
  if (group == nullptr)
    return FEMM_INVALID_ARGUMENT;
  if (!femm::DevicePresent())
    return FEMM_GPU_UNAVAILABLE;
  const femm_newton_options_t o = options ? *options : femm::DefaultNewtonOptions();
  const femm_solve_config_t c = config ? *config : femm::DefaultSolveConfig();
  if (!femm::ConfigValid(c))
    return FEMM_INVALID_ARGUMENT;
  if (hipDeviceSynchronize() != hipSuccess)
    return FEMM_INTERNAL_ERROR;
  ReleaseGraph(group);
  // Everything that allocates or copies happens before the capture.
  std::vector<femm::GroupEntry> table;
  int st = PrepareGroup(group, c, nullptr, &table);
  if (st != FEMM_OK)
    return st;
  if (hipMalloc(&group->graph_table, sizeof(femm::GroupEntry) * table.size()) != hipSuccess) {
    group->graph_table = nullptr;
    return FEMM_GPU_ALLOCATION_FAILED;
  }
  if (hipMemcpy(group->graph_table, table.data(), sizeof(femm::GroupEntry) * table.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    ReleaseGraph(group);
    return FEMM_INTERNAL_ERROR;
  }
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    ReleaseGraph(group);
    return FEMM_INTERNAL_ERROR;
  }
  hipGraph_t graph = nullptr;
  st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess
           ? FEMM_OK
           : FEMM_INTERNAL_ERROR;
  if (st == FEMM_OK) {
    const int launched = EnqueueGroup(group->graph_table, table, o, c, safeguarded != 0, s);
    const bool ended = hipStreamEndCapture(s, &graph) == hipSuccess;
    st = launched != FEMM_OK ? launched : (ended ? FEMM_OK : FEMM_INTERNAL_ERROR);
  }
  if (st == FEMM_OK && hipGraphInstantiate(&group->graph, graph, nullptr, nullptr, 0) != hipSuccess) {
    group->graph = nullptr;
    st = FEMM_INTERNAL_ERROR;
  }
  if (graph != nullptr)
    (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(s);
  if (st != FEMM_OK) {
    ReleaseGraph(group);
    return st;
  }
  group->graph_entries = table;
  return FEMM_OK;
  FEMM_API_END
}

int femm_group_replay(femm_group *group, void *stream)
{
  FEMM_API_BEGIN
  if (group == nullptr || group->graph == nullptr)
    return FEMM_INVALID_ARGUMENT;
  // The graph holds the device arrays and workspaces it was captured on: refuse once a
  // plan's were released or reallocated (a larger batch, another precision or block size).
  for (size_t e = 0; e < group->entries.size(); ++e) {
    const femm::DeviceBuffers &dev = group->entries[e].plan->dev;
    const femm::GroupEntry &g = group->graph_entries[e];
    if (std::memcmp(&dev.plan, &g.plan, sizeof(g.plan)) != 0 ||
        std::memcmp(&dev.work, &g.work, sizeof(g.work)) != 0)
      return FEMM_INVALID_ARGUMENT;
  }
  return hipGraphLaunch(group->graph, static_cast<hipStream_t>(stream)) == hipSuccess
             ? FEMM_OK
             : FEMM_INTERNAL_ERROR;
  FEMM_API_END
}

int femm_newton_group(femm_group *group, const femm_newton_options_t *options,
                      const femm_solve_config_t *config, void *stream)
{
  FEMM_API_BEGIN
  return GroupSolve(group, options, config, stream, false);
  FEMM_API_END
}

int femm_newton_group_safeguarded(femm_group *group, const femm_newton_options_t *options,
                                  const femm_solve_config_t *config, void *stream)
{
  FEMM_API_BEGIN
  return GroupSolve(group, options, config, stream, true);
  FEMM_API_END
}

}  // extern "C"
