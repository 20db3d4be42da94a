// hip_runtime.cpp -- the reference-named C++ engines over the C ABI of
// bfnew_amd.h (reference surfaces: include/bfnew/hip/runtime.hpp:19-327,
// batched_jacobi.hpp:137-200, batched_dense_chaotic_push.hpp:105-165,
// batched_frontier_push.hpp:91-144, jacobi.hpp:64-125).
#include "bfnew/hip/engines.hpp"
#include "bfnew/hip/runtime.hpp"

#include <chrono>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>

namespace bfnew::hip {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
HipRuntimeError::HipRuntimeError(std::int32_t status, std::string message)
    : std::runtime_error(std::move(message)), status_{status} {}

void throw_if_error(std::int32_t status, std::string_view what) {
  if (status == BFN_OK) return;
  std::string message{what};
  message += ": ";
  message += bfn_last_error();
  switch (status) {
    case BFN_E_INVALID_ARGUMENT:
    case BFN_E_NOT_REPRESENTABLE:
      throw std::invalid_argument{message};
    case BFN_E_STATE:
      throw std::logic_error{message};
    case BFN_E_DEVICE_RESULT:
    case BFN_E_OUT_OF_MEMORY:
      throw std::runtime_error{message};
    default:
      throw HipRuntimeError{status, message};
  }
}

// ---------------------------------------------------------------------------
// stream, graph, workspace
// ---------------------------------------------------------------------------
HipStream::HipStream(bool /*nonblocking*/, std::int32_t device_ordinal) {
  throw_if_error(bfn_device_open(device_ordinal, &device_), "HipStream");
}
HipStream::~HipStream() noexcept {
  if (device_) (void)bfn_device_close(device_);
}
HipStream::HipStream(HipStream&& other) noexcept : device_{std::exchange(other.device_, nullptr)} {}
HipStream& HipStream::operator=(HipStream&& other) noexcept {
  if (this != &other) {
    if (device_) (void)bfn_device_close(device_);
    device_ = std::exchange(other.device_, nullptr);
  }
  return *this;
}
void HipStream::synchronize() const {
  throw_if_error(bfn_device_synchronize(device_), "HipStream::synchronize");
}

ResidentGraphPlan make_resident_graph_plan(DeviceGraphLayout32 layout) {
  ResidentGraphPlan plan;
  const std::uint32_t V = layout.vertex_count;
  const std::uint32_t T = layout.tile_count;
  if (layout.owner_tiles.size() != V)
    throw std::invalid_argument{"resident graph plan: owner_tiles size differs from vertex_count"};
  plan.tile_vertex_offsets.assign(T + 1U, 0U);
  for (std::uint32_t v = 0; v < V; ++v) {
    const std::uint32_t t = layout.owner_tiles[v];
    if (t >= T || (v > 0U && t < layout.owner_tiles[v - 1U]))
      throw std::invalid_argument{"resident graph plan: vertices are not tile-major"};
    ++plan.tile_vertex_offsets[t + 1U];
  }
  for (std::uint32_t t = 0; t < T; ++t) plan.tile_vertex_offsets[t + 1U] += plan.tile_vertex_offsets[t];
  plan.memory = report_device_graph_memory(layout);
  plan.fingerprint = fingerprint_device_graph_layout32(layout);
  plan.layout = std::move(layout);
  return plan;
}

ResidentDeviceGraph::~ResidentDeviceGraph() noexcept {
  if (graph_) (void)bfn_graph_free(graph_);
}
ResidentDeviceGraph::ResidentDeviceGraph(ResidentDeviceGraph&& o) noexcept
    : graph_{std::exchange(o.graph_, nullptr)},
      vertex_count_{o.vertex_count_},
      tile_count_{o.tile_count_},
      memory_{o.memory_},
      fingerprint_{o.fingerprint_} {}
ResidentDeviceGraph& ResidentDeviceGraph::operator=(ResidentDeviceGraph&& o) noexcept {
  if (this != &o) {
    if (graph_) (void)bfn_graph_free(graph_);
    graph_ = std::exchange(o.graph_, nullptr);
    vertex_count_ = o.vertex_count_;
    tile_count_ = o.tile_count_;
    memory_ = o.memory_;
    fingerprint_ = o.fingerprint_;
  }
  return *this;
}

void ResidentDeviceGraph::upload_once_async(ResidentGraphPlan plan, const HipStream& stream) {
  if (graph_) throw std::logic_error{"ResidentDeviceGraph: already uploaded"};
  const DeviceGraphLayout32& L = plan.layout;
  bfn_graph_layout32 c{};
  c.vertex_count = L.vertex_count;
  c.edge_count = L.edge_count;
  c.tile_count = L.tile_count;
  c.csr_run_count = static_cast<std::uint32_t>(L.csr_run_destination_tiles.size());
  c.csc_run_count = static_cast<std::uint32_t>(L.csc_run_source_tiles.size());
  c.owner_tiles = L.owner_tiles.data();
  c.tile_vertex_offsets = plan.tile_vertex_offsets.data();
  c.csr_row_offsets = L.csr_row_offsets.data();
  c.csr_destinations = L.csr_destinations.data();
  c.csr_weights = L.csr_weights.data();
  c.csr_row_run_offsets = L.csr_row_run_offsets.data();
  c.csr_run_edge_offsets = L.csr_run_edge_offsets.data();
  c.csr_run_destination_tiles = L.csr_run_destination_tiles.data();
  c.csc_column_offsets = L.csc_column_offsets.data();
  c.csc_sources = L.csc_sources.data();
  c.csc_weights = L.csc_weights.data();
  c.csc_edge_ids = L.csc_edge_ids.data();
  c.csc_column_run_offsets = L.csc_column_run_offsets.data();
  c.csc_run_edge_offsets = L.csc_run_edge_offsets.data();
  c.csc_run_source_tiles = L.csc_run_source_tiles.data();
  throw_if_error(bfn_graph_upload(stream.device(), &c, &graph_), "ResidentDeviceGraph::upload");
  vertex_count_ = L.vertex_count;
  tile_count_ = L.tile_count;
  memory_ = plan.memory;
  fingerprint_ = plan.fingerprint;
}

void ResidentDeviceGraph::wait_until_ready(const HipStream& stream) const { stream.synchronize(); }
void ResidentDeviceGraph::synchronize_upload() {}

const DeviceGraphMemoryReport& ResidentDeviceGraph::memory_report() const {
  if (!graph_) throw std::logic_error{"ResidentDeviceGraph: no upload"};
  return memory_;
}
const DeviceGraphFingerprint& ResidentDeviceGraph::fingerprint() const {
  if (!graph_) throw std::logic_error{"ResidentDeviceGraph: no upload"};
  return fingerprint_;
}
bfn_graph ResidentDeviceGraph::handle() const {
  if (!graph_) throw std::logic_error{"ResidentDeviceGraph: no upload"};
  return graph_;
}

void ResidentDeviceGraph::update_node_costs(const std::vector<float>& base_cost,
                                            const std::vector<std::int32_t>& occupancy) const {
  if (base_cost.size() != vertex_count_ || (!occupancy.empty() && occupancy.size() != vertex_count_))
    throw std::invalid_argument{"update_node_costs: arrays must have one entry per vertex"};
  throw_if_error(bfn_graph_update_node_costs(handle(), base_cost.data(),
                                             occupancy.empty() ? nullptr : occupancy.data()),
                 "ResidentDeviceGraph::update_node_costs");
}

ReusableBatchedWorkspace::ReusableBatchedWorkspace(const HipStream& stream,
                                                   std::uint64_t budget_bytes) {
  throw_if_error(bfn_workspace_create(stream.device(), budget_bytes, &workspace_),
                 "ReusableBatchedWorkspace");
}
ReusableBatchedWorkspace::~ReusableBatchedWorkspace() noexcept {
  if (workspace_) (void)bfn_workspace_free(workspace_);
}

// ---------------------------------------------------------------------------
// batched engines
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

// The lanes of a description as C ABI queries: valid lanes in lane order,
// each with the union tiles whose mask holds its bit.
struct LaneBatch {
  std::vector<std::uint32_t> lane_of_query;
  std::vector<std::uint32_t> query_ids, source_offsets{0U}, sources, target_offsets{0U}, targets,
      tile_offsets{0U}, tiles;
  bfn_batch_desc desc{};

  explicit LaneBatch(const BatchDeviceDescription& d) {
    for (std::uint32_t l = 0; l < d.lane_width; ++l) {
      if (!((d.valid_lane_mask >> l) & 1U)) continue;
      lane_of_query.push_back(l);
      query_ids.push_back(d.query_ids_by_lane[l]);
      sources.insert(sources.end(), d.sources.begin() + d.source_offsets[l],
                     d.sources.begin() + d.source_offsets[l + 1U]);
      targets.insert(targets.end(), d.targets.begin() + d.target_offsets[l],
                     d.targets.begin() + d.target_offsets[l + 1U]);
      for (std::uint32_t t : d.union_tiles)
        if ((d.tile_lane_masks[t] >> l) & 1U) tiles.push_back(t);
      source_offsets.push_back(static_cast<std::uint32_t>(sources.size()));
      target_offsets.push_back(static_cast<std::uint32_t>(targets.size()));
      tile_offsets.push_back(static_cast<std::uint32_t>(tiles.size()));
    }
    desc.query_count = static_cast<std::uint32_t>(lane_of_query.size());
    desc.query_ids = query_ids.data();
    desc.source_offsets = source_offsets.data();
    desc.sources = sources.data();
    desc.target_offsets = target_offsets.data();
    desc.targets = targets.data();
    desc.tile_offsets = tile_offsets.data();
    desc.selected_tiles = tiles.data();
    desc.engine_by_query = nullptr;
  }
};

// One route query as a one-lane C ABI batch.
struct QueryBatch {
  std::uint32_t query_id;
  std::vector<std::uint32_t> source_offsets, sources, target_offsets, targets, tile_offsets, tiles;
  bfn_batch_desc desc{};

  explicit QueryBatch(const RouteQuery& q) : query_id{q.query_id.value()} {
    for (VertexId v : q.sources) sources.push_back(v.value());
    for (VertexId v : q.targets) targets.push_back(v.value());
    for (TileId t : q.selected_tiles) tiles.push_back(t.value());
    source_offsets = {0U, static_cast<std::uint32_t>(sources.size())};
    target_offsets = {0U, static_cast<std::uint32_t>(targets.size())};
    tile_offsets = {0U, static_cast<std::uint32_t>(tiles.size())};
    desc.query_count = 1U;
    desc.query_ids = &query_id;
    desc.source_offsets = source_offsets.data();
    desc.sources = sources.data();
    desc.target_offsets = target_offsets.data();
    desc.targets = targets.data();
    desc.tile_offsets = tile_offsets.data();
    desc.selected_tiles = tiles.data();
  }
};

bfn_tuning reference_tuning() {
  bfn_tuning t{};
  bfn_tuning_default(&t);
  t.tile_sweeps = 1U;  // point Jacobi: the reference's rounds
  return t;
}

void check_options(const GpuRunOptions& options, EngineKind kind, const char* engine) {
  if (validate_gpu_run_options(options) != GpuRunOptionsError::none || options.engine != kind)
    throw std::invalid_argument{std::string{engine} +
                                " options are invalid or select another engine"};
}

// Submits, waits, and reads the status records back; lane masks are
// rebuilt in lane positions from the per-query records.
struct Completed {
  DeviceRunStatus status{};
  DeviceWorkStatistics work{};
  std::vector<bfn_query_status> per_query;
};

Completed submit_and_wait(bfn_workspace ws, bfn_graph graph, const bfn_batch_desc& desc,
                          const GpuRunOptions& options, const bfn_tuning& tuning,
                          const std::vector<std::uint32_t>& lane_of_query) {
  bfn_run_options o;
  static_assert(sizeof o == sizeof options);
  std::memcpy(&o, &options, sizeof o);
  throw_if_error(bfn_batch_submit(ws, graph, &desc, &o, &tuning), "batch submit");
  throw_if_error(bfn_batch_wait(ws), "batch wait");
  Completed c;
  c.per_query.resize(desc.query_count);
  bfn_run_status s{};
  bfn_work_statistics w{};
  throw_if_error(bfn_batch_status(ws, &s, c.per_query.data(), &w), "batch status");
  std::memcpy(static_cast<void*>(&c.work), &w, sizeof w);
  DeviceRunStatus& r = c.status;
  r.final_distance_slot = s.final_distance_slot;
  r.converged = s.converged;
  r.rounds_completed = s.rounds_completed;
  r.stop_reason = s.stop_reason;
  r.error_bits = s.error_bits;
  for (std::size_t i = 0; i < c.per_query.size(); ++i) {
    const LaneMask bit = 1U << lane_of_query[i];
    const bfn_query_status& q = c.per_query[i];
    r.valid_lane_mask |= bit;
    (q.converged ? r.converged_lane_mask : r.active_lane_mask) |= bit;
    if (s.converged && q.reached_all_targets) r.reached_target_mask |= bit;
    if (s.converged && q.bounding_box_miss) r.bounding_box_miss_mask |= bit;
  }
  if (validate_device_run_status(r) != DeviceRunStatusError::none)
    throw std::runtime_error{"batched engine produced an inconsistent DeviceRunStatus"};
  return c;
}

// Both path passes; `transfer` counts the payload as it crosses the bus
// (28-byte summaries, 4-byte vertices, labels and edge ids) and
// `control_bytes` the offsets and edge-arena slots that travel beside it.
std::vector<CompactTargetPath> read_paths(bfn_workspace ws, std::size_t target_count,
                                          CompactTransferAccounting& transfer,
                                          std::uint64_t& control_bytes) {
  std::vector<bfn_target_summary> summaries(target_count);
  std::vector<std::uint64_t> offsets(target_count + 1U);
  std::uint64_t total = 0;
  throw_if_error(bfn_batch_paths(ws, summaries.data(), offsets.data(), nullptr, nullptr, nullptr,
                                 0U, &total),
                 "compact paths");
  const std::uint64_t n = total ? total : 1U;
  std::vector<std::uint32_t> vertices(n), edges(n);
  std::vector<float> labels(n);
  throw_if_error(bfn_batch_paths(ws, summaries.data(), offsets.data(), vertices.data(),
                                 labels.data(), edges.data(), n, &total),
                 "compact paths");
  std::vector<CompactTargetPath> out(target_count);
  for (std::size_t i = 0; i < target_count; ++i) {
    const bfn_target_summary& s = summaries[i];
    CompactTargetPath& p = out[i];
    p.summary.target = VertexId{s.target};
    p.summary.selected_source = VertexId{s.selected_source};
    p.summary.distance = s.distance;
    p.summary.path_length = s.path_length;
    p.summary.reached = static_cast<CompactTargetReachStatus>(s.reached);
    p.summary.reconstruction = static_cast<CompactPathStatus>(s.reconstruction);
    p.summary.has_selected_source = s.has_selected_source;
    const std::uint64_t a = offsets[i], b = offsets[i + 1U];
    for (std::uint64_t k = a; k < b; ++k) {
      p.vertices.push_back(VertexId{vertices[k]});
      p.distance_labels.push_back(labels[k]);
      if (k + 1U < b) p.edge_ids.push_back(EdgeId{edges[k]});
    }
    transfer.vertex_bytes += 4U * p.vertices.size();
    transfer.distance_label_bytes += 4U * p.distance_labels.size();
    transfer.edge_id_bytes += 4U * p.edge_ids.size();
  }
  transfer.summary_bytes = sizeof(bfn_target_summary) * target_count;
  transfer.total_bytes = transfer.summary_bytes + transfer.vertex_bytes +
                         transfer.distance_label_bytes + transfer.edge_id_bytes;
  // The edge arena is copied at the vertex arena's length.
  control_bytes = 8U * (target_count + 1U) + (total ? 4U * total - transfer.edge_id_bytes : 0U);
  return out;
}

}  // namespace

BatchedEngineBase::BatchedEngineBase(EngineKind kind, const WeightedGraph& host_graph,
                                     const TileRunLayout64& tile_runs,
                                     const ResidentDeviceGraph& resident_graph,
                                     ReusableBatchedWorkspace& workspace, const HipStream& stream)
    : kind_{kind},
      host_graph_{host_graph},
      tile_runs_{tile_runs},
      resident_graph_{resident_graph},
      workspace_{workspace},
      stream_{stream},
      tuning_{reference_tuning()} {
  if (!resident_graph.has_upload()) throw std::invalid_argument{"batched engine: graph not uploaded"};
  if (resident_graph.vertex_count() != host_graph.vertex_count() ||
      tile_runs.csr_row_run_offsets.size() != static_cast<std::size_t>(host_graph.vertex_count()) + 1U)
    throw std::invalid_argument{"batched engine: host graph, tile runs and resident graph differ"};
}

BatchedRunOutput BatchedEngineBase::run_impl(const BatchDeviceDescription& batch,
                                             const GpuRunOptions& options, Readback readback,
                                             CompactPathBatchOutput* paths) {
  const auto wall0 = Clock::now();
  static const char* names[] = {"batched Jacobi", "batched dense chaotic push",
                                "batched frontier push"};
  check_options(options, kind_, names[static_cast<std::uint32_t>(kind_)]);
  if (readback == Readback::compact_paths && options.instrumentation != InstrumentationLevel::none)
    throw std::invalid_argument{"compact-path mode requires no instrumentation"};
  const BatchLayoutValidationResult v = validate_batch_device_description(host_graph_, batch);
  if (!v.ok()) throw std::invalid_argument{"invalid BatchDeviceDescription"};

  LaneBatch lanes{batch};
  bfn_tuning tune = tuning_;
  if (options.instrumentation != InstrumentationLevel::none) tune.count_relaxations = 1U;
  Completed c = submit_and_wait(workspace_.handle(), resident_graph_.handle(), lanes.desc, options,
                                tune, lanes.lane_of_query);

  BatchedRunOutput out;
  out.result.engine_kind = static_cast<std::uint32_t>(kind_);
  out.result.control_mode = static_cast<std::uint32_t>(options.control_mode);
  out.result.status = c.status;
  out.result.work = c.work;
  const std::uint32_t W = batch.lane_width;
  out.lane_convergence_rounds.assign(W, 0U);
  out.lane_executed_rounds.assign(W, 0U);
  out.lane_relaxations.assign(W, 0U);
  for (std::size_t i = 0; i < c.per_query.size(); ++i) {
    const std::uint32_t l = lanes.lane_of_query[i];
    out.lane_executed_rounds[l] = c.per_query[i].rounds_completed;
    if (c.per_query[i].converged) out.lane_convergence_rounds[l] = c.per_query[i].rounds_completed;
    out.lane_relaxations[l] = c.per_query[i].relaxations;
  }
  float relax_ms = 0.0f, total_ms = 0.0f;
  std::uint64_t launches = 0;
  throw_if_error(bfn_batch_timing(workspace_.handle(), &relax_ms, &total_ms, &launches), "timing");
  out.metrics.sssp_device_timeline_milliseconds = total_ms;
  out.metrics.lane_width = W;
  out.metrics.valid_lane_count = lanes.desc.query_count;
  out.metrics.engine_round_dispatches = launches;
  out.metrics.convergence_host_checks = c.work.host_checks;
  if (options.instrumentation != InstrumentationLevel::none) {
    bfn_engine_counters k{};
    throw_if_error(bfn_batch_counters(workspace_.handle(), &k), "counters");
    out.metrics.tile_visits = k.tile_visits;
    out.metrics.edge_records_evaluated = k.edges_relaxed;
    out.metrics.algorithmic_bytes = k.algorithmic_bytes;
  }

  const std::uint32_t V = host_graph_.vertex_count();
  if (readback == Readback::distances) {
    out.distances.assign(static_cast<std::size_t>(V) * W, std::numeric_limits<float>::infinity());
    std::vector<float> lane(V);
    for (std::size_t i = 0; i < lanes.lane_of_query.size(); ++i) {
      throw_if_error(bfn_batch_distances(workspace_.handle(), static_cast<std::uint32_t>(i),
                                         lane.data()),
                     "distances");
      const std::uint32_t l = lanes.lane_of_query[i];
      for (std::uint32_t u = 0; u < V; ++u) out.distances[static_cast<std::size_t>(u) * W + l] = lane[u];
    }
    out.distances_downloaded = true;
  }
  if (readback == Readback::compact_paths && paths) {
    paths->status = c.status;
    const auto t0 = Clock::now();
    paths->targets = read_paths(workspace_.handle(), lanes.targets.size(), paths->transfer,
                                paths->metrics.control_bytes);
    const auto t1 = Clock::now();
    paths->metrics.sssp_device_milliseconds = total_ms;
    paths->metrics.sssp_host_milliseconds = std::chrono::duration<double, std::milli>(t0 - wall0).count();
    paths->metrics.reconstruction_host_milliseconds = std::chrono::duration<double, std::milli>(t1 - t0).count();
    paths->metrics.end_to_end_wall_milliseconds =
        std::chrono::duration<double, std::milli>(Clock::now() - wall0).count();
  }
  out.metrics.end_to_end_wall_milliseconds =
      std::chrono::duration<double, std::milli>(Clock::now() - wall0).count();
  if (out.metrics.end_to_end_wall_milliseconds > 0.0)
    out.metrics.batch_queries_per_second =
        lanes.desc.query_count * 1e3 / out.metrics.end_to_end_wall_milliseconds;
  (void)stream_;
  (void)tile_runs_;
  return out;
}

namespace {
void check_load_strategy(BatchedJacobiLoadStrategy s) {
  if (s != BatchedJacobiLoadStrategy::compiler_uniform &&
      s != BatchedJacobiLoadStrategy::explicit_wave_broadcast)
    throw std::invalid_argument{"batched Jacobi: invalid load strategy"};
}
}  // namespace

BatchedRunOutput BatchedJacobiPullEngine::run_status_only(const BatchDeviceDescription& b,
                                                          const GpuRunOptions& o,
                                                          BatchedJacobiLoadStrategy s) {
  check_load_strategy(s);
  return run_impl(b, o, Readback::status, nullptr);
}
DeviceRunStatus BatchedJacobiPullEngine::run_compact_status(const BatchDeviceDescription& b,
                                                            const GpuRunOptions& o,
                                                            BatchedJacobiLoadStrategy s) {
  check_load_strategy(s);
  return run_impl(b, o, Readback::compact_status, nullptr).result.status;
}
CompactPathBatchOutput BatchedJacobiPullEngine::run_compact_paths(
    const BatchDeviceDescription& b, const GpuRunOptions& o, ReusableCompactPathWorkspace&,
    BatchedJacobiLoadStrategy s) {
  check_load_strategy(s);
  CompactPathBatchOutput out;
  (void)run_impl(b, o, Readback::compact_paths, &out);
  return out;
}
BatchedRunOutput BatchedJacobiPullEngine::run_with_distances(const BatchDeviceDescription& b,
                                                             const GpuRunOptions& o,
                                                             BatchedJacobiLoadStrategy s) {
  check_load_strategy(s);
  return run_impl(b, o, Readback::distances, nullptr);
}

#define BFNEW_PUSH_ENGINE_METHODS(Engine)                                                       \
  BatchedRunOutput Engine::run_status_only(const BatchDeviceDescription& b,                     \
                                           const GpuRunOptions& o) {                            \
    return run_impl(b, o, Readback::status, nullptr);                                           \
  }                                                                                             \
  DeviceRunStatus Engine::run_compact_status(const BatchDeviceDescription& b,                   \
                                             const GpuRunOptions& o) {                          \
    return run_impl(b, o, Readback::compact_status, nullptr).result.status;                     \
  }                                                                                             \
  CompactPathBatchOutput Engine::run_compact_paths(const BatchDeviceDescription& b,             \
                                                   const GpuRunOptions& o,                      \
                                                   ReusableCompactPathWorkspace&) {             \
    CompactPathBatchOutput out;                                                                 \
    (void)run_impl(b, o, Readback::compact_paths, &out);                                        \
    return out;                                                                                 \
  }                                                                                             \
  BatchedRunOutput Engine::run_with_distances(const BatchDeviceDescription& b,                  \
                                              const GpuRunOptions& o) {                         \
    return run_impl(b, o, Readback::distances, nullptr);                                        \
  }
BFNEW_PUSH_ENGINE_METHODS(BatchedDenseChaoticPushEngine)
BFNEW_PUSH_ENGINE_METHODS(BatchedFrontierPushEngine)
#undef BFNEW_PUSH_ENGINE_METHODS

// ---------------------------------------------------------------------------
// single-query engines
// ---------------------------------------------------------------------------
SingleQueryEngineBase::SingleQueryEngineBase(EngineKind kind, const WeightedGraph& host_graph,
                                             const TileRunLayout64& tile_runs,
                                             const ResidentDeviceGraph& resident_graph,
                                             ReusableDeviceWorkspace& workspace,
                                             const HipStream& stream)
    : kind_{kind},
      host_graph_{host_graph},
      resident_graph_{resident_graph},
      workspace_{workspace},
      stream_{stream},
      tuning_{reference_tuning()} {
  if (!resident_graph.has_upload()) throw std::invalid_argument{"engine: graph not uploaded"};
  if (resident_graph.vertex_count() != host_graph.vertex_count() ||
      tile_runs.csr_row_run_offsets.size() != static_cast<std::size_t>(host_graph.vertex_count()) + 1U)
    throw std::invalid_argument{"engine: host graph, tile runs and resident graph differ"};
}

GpuSsspResult SingleQueryEngineBase::run(const RouteQuery& q, const GpuRunOptions& o) {
  return run_impl(q, o, Readback::none).result;
}
SingleQueryRunOutput SingleQueryEngineBase::run_status_only(const RouteQuery& q,
                                                            const GpuRunOptions& o) {
  return run_impl(q, o, Readback::none);
}
SingleQueryRunOutput SingleQueryEngineBase::run_with_distances(const RouteQuery& q,
                                                               const GpuRunOptions& o) {
  return run_impl(q, o, Readback::full_graph);
}
SingleQueryRunOutput SingleQueryEngineBase::run_with_selected_distances(const RouteQuery& q,
                                                                        const GpuRunOptions& o) {
  return run_impl(q, o, Readback::selected_ranges);
}

SingleQueryRunOutput SingleQueryEngineBase::run_impl(const RouteQuery& query,
                                                     const GpuRunOptions& options,
                                                     Readback readback) {
  static const char* names[] = {"Jacobi", "dense chaotic push", "frontier push"};
  check_options(options, kind_, names[static_cast<std::uint32_t>(kind_)]);
  if (!validate_route_query(host_graph_, query).ok())
    throw std::invalid_argument{"engine: invalid route query"};
  QueryBatch one{query};
  bfn_tuning tune = tuning_;
  tune.count_relaxations = 1U;
  Completed c = submit_and_wait(workspace_.handle(), resident_graph_.handle(), one.desc, options,
                                tune, std::vector<std::uint32_t>{0U});
  SingleQueryRunOutput out;
  out.result.engine_kind = static_cast<std::uint32_t>(kind_);
  out.result.control_mode = static_cast<std::uint32_t>(options.control_mode);
  out.result.status = c.status;
  out.result.work = c.work;
  out.relaxations = c.per_query[0].relaxations;
  if (readback != Readback::none) {
    std::vector<float> full(host_graph_.vertex_count());
    throw_if_error(bfn_batch_distances(workspace_.handle(), 0U, full.data()), "distances");
    if (readback == Readback::full_graph) {
      out.distances = std::move(full);
    } else {
      const auto offsets = host_graph_.tile_vertex_offsets();
      for (TileId t : query.selected_tiles)
        out.distances.insert(out.distances.end(), full.begin() + offsets[t.value()],
                             full.begin() + offsets[t.value() + 1U]);
    }
    out.distances_downloaded = true;
  }
  (void)stream_;
  return out;
}

}  // namespace bfnew::hip
