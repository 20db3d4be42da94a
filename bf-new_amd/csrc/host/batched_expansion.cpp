// batched_expansion.cpp -- BatchedExpansionExecutor: one planned lane batch
// through a batched engine (reference behaviour:
// src/hip/batched_expansion.hip.cpp:459 and its payload rule :117-163); and
// PassBatchedExpansionExecutor: one device run per planning pass.
#include "bfnew/hip/batched_expansion.hpp"

#include "bfnew/compact_paths.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <string_view>

namespace bfnew::hip {

void throw_if_error(std::int32_t status, std::string_view what);  // hip_runtime.cpp

namespace {

std::uint64_t nanoseconds(double milliseconds) {
  return milliseconds > 0.0 ? static_cast<std::uint64_t>(std::llround(milliseconds * 1.0e6)) : 0U;
}

const RouteQuery& lane_query(std::span<const RouteQuery> queries, QueryId id,
                             std::uint32_t generation) {
  for (const RouteQuery& q : queries)
    if (q.query_id == id && q.expansion_generation == generation) return q;
  throw std::runtime_error{"compact paths name a lane query that was not submitted"};
}

}  // namespace

class BatchedExpansionExecutor::Impl {
 public:
  Impl(const WeightedGraph& graph, const TileRunLayout64& runs, const ResidentDeviceGraph& resident,
       ReusableBatchedWorkspace& ws, const HipStream& stream, GpuRunOptions options,
       BatchedExpansionTransferMode mode)
      : graph_{graph},
        options_{options},
        mode_{mode},
        jacobi_{graph, runs, resident, ws, stream},
        dense_{graph, runs, resident, ws, stream},
        frontier_{graph, runs, resident, ws, stream} {
    if (validate_gpu_run_options(options) != GpuRunOptionsError::none)
      throw std::invalid_argument{"batched expansion executor: invalid run options"};
    if (!valid_batched_expansion_transfer_mode(mode))
      throw std::invalid_argument{"batched expansion executor: invalid transfer mode"};
    if (mode == BatchedExpansionTransferMode::compact_paths &&
        options.instrumentation != InstrumentationLevel::none)
      throw std::invalid_argument{"batched expansion executor: compact paths need no instrumentation"};
  }

  EngineKind kind() const noexcept { return options_.engine; }
  BatchedExpansionTransferMode mode() const noexcept { return mode_; }
  void tune(const bfn_tuning& t) noexcept {
    jacobi_.set_tuning(t);
    dense_.set_tuning(t);
    frontier_.set_tuning(t);
  }

  ExpansionBatchExecution run(std::span<const RouteQuery> queries,
                              std::span<const BatchQueryFeatures> features,
                              const BatchPlanEntry& batch) {
    prepare_batch_device_description(graph_, queries, features, batch, description_);
    ExpansionBatchExecution out;
    out.result.engine_kind = static_cast<std::uint32_t>(options_.engine);
    out.result.control_mode = static_cast<std::uint32_t>(options_.control_mode);
    switch (mode_) {
      case BatchedExpansionTransferMode::compact_status:
        out.result.status = compact_status();
        break;
      case BatchedExpansionTransferMode::status_and_work_evidence: {
        GpuRunOptions o = options_;
        if (o.instrumentation == InstrumentationLevel::none) o.instrumentation = InstrumentationLevel::light;
        const BatchedRunOutput r = status_only(o);
        out.result = r.result;
        // A tile visit serves one (query, tile) slot, so every edge record
        // it reads is relaxed for exactly one lane: logical work equals the
        // records the visits read.
        out.work_evidence = ExpansionWorkEvidence::measured;
        out.shared_edge_work = r.metrics.edge_records_evaluated;
        out.logical_lane_edge_work = r.metrics.edge_records_evaluated;
        break;
      }
      case BatchedExpansionTransferMode::compact_paths:
        compact_paths(queries, out);
        break;
    }
    if (mode_ == BatchedExpansionTransferMode::compact_paths) {
      out.compact_total_device_to_host_bytes =
          out.compact_transfer.total_bytes + out.compact_status_bytes + out.compact_error_bytes;
      out.compact_overall_device_to_host_bytes =
          out.compact_total_device_to_host_bytes + out.compact_controller_poll_bytes;
    }
    return out;
  }

 private:
  DeviceRunStatus compact_status() {
    switch (options_.engine) {
      case EngineKind::jacobi_pull: return jacobi_.run_compact_status(description_, options_);
      case EngineKind::dense_chaotic_push: return dense_.run_compact_status(description_, options_);
      default: return frontier_.run_compact_status(description_, options_);
    }
  }
  BatchedRunOutput status_only(const GpuRunOptions& o) {
    switch (options_.engine) {
      case EngineKind::jacobi_pull: return jacobi_.run_status_only(description_, o);
      case EngineKind::dense_chaotic_push: return dense_.run_status_only(description_, o);
      default: return frontier_.run_status_only(description_, o);
    }
  }
  CompactPathBatchOutput paths() {
    switch (options_.engine) {
      case EngineKind::jacobi_pull: return jacobi_.run_compact_paths(description_, options_, pws_);
      case EngineKind::dense_chaotic_push: return dense_.run_compact_paths(description_, options_, pws_);
      default: return frontier_.run_compact_paths(description_, options_, pws_);
    }
  }

  // One payload per valid lane of a cleanly converged batch, sorted by id;
  // the targets of lane l are description_.targets[target_offsets[l] ..).
  // A batch that stopped otherwise returns no payload and no byte record.
  void compact_paths(std::span<const RouteQuery> queries, ExpansionBatchExecution& out) {
    CompactPathBatchOutput r = paths();
    out.result.status = r.status;
    CompactPathExecutionMetrics& t = out.compact_execution;
    t.host_timing = CompactStageTimingEvidence::measured;
    t.sssp_host_nanoseconds = nanoseconds(r.metrics.sssp_host_milliseconds);
    t.reconstruction_host_nanoseconds = nanoseconds(r.metrics.reconstruction_host_milliseconds);
    // The path passes and their copies are timed together, as reconstruction.
    t.end_to_end_host_nanoseconds = std::max(nanoseconds(r.metrics.end_to_end_wall_milliseconds),
                                             t.sssp_host_nanoseconds + t.reconstruction_host_nanoseconds);
    const BatchDeviceDescription& d = description_;
    if (r.targets.size() != d.targets.size())
      throw std::runtime_error{"compact paths disagree with the batch's targets"};
    const bool clean = r.status.converged == 1U &&
                       r.status.stop_reason == static_cast<std::uint32_t>(DeviceStopReason::converged) &&
                       r.status.error_bits == 0U;
    if (!clean) return;
    out.compact_transfer = r.transfer;
    out.compact_status_bytes = sizeof(DeviceRunStatus);
    out.compact_error_bytes = r.metrics.control_bytes;
    for (std::uint32_t lane = 0; lane < d.lane_width; ++lane) {
      if (!(d.valid_lane_mask >> lane & 1U)) continue;
      const RouteQuery& q =
          lane_query(queries, QueryId{d.query_ids_by_lane[lane]}, d.expansion_generations_by_lane[lane]);
      const std::uint32_t begin = d.target_offsets[lane], end = d.target_offsets[lane + 1];
      if (end - begin != q.targets.size())
        throw std::runtime_error{"a lane's compact paths disagree with its query's targets"};
      CompactPathPayload p;
      p.query_id = q.query_id;
      p.expansion_generation = q.expansion_generation;
      p.target_terminal_to_target = q.target_terminal_to_target;
      p.targets.assign(std::make_move_iterator(r.targets.begin() + begin),
                       std::make_move_iterator(r.targets.begin() + end));
      out.compact_paths.push_back(std::move(p));
    }
    std::sort(out.compact_paths.begin(), out.compact_paths.end(),
              [](const CompactPathPayload& a, const CompactPathPayload& b) { return a.query_id < b.query_id; });
  }

  const WeightedGraph& graph_;
  GpuRunOptions options_;
  BatchedExpansionTransferMode mode_;
  BatchedJacobiPullEngine jacobi_;
  BatchedDenseChaoticPushEngine dense_;
  BatchedFrontierPushEngine frontier_;
  ReusableCompactPathWorkspace pws_;
  BatchDeviceDescription description_;
};

BatchedExpansionExecutor::BatchedExpansionExecutor(const WeightedGraph& host_graph,
                                                   const TileRunLayout64& tile_runs,
                                                   const ResidentDeviceGraph& resident_graph,
                                                   ReusableBatchedWorkspace& workspace,
                                                   const HipStream& stream, GpuRunOptions run_options,
                                                   BatchedExpansionTransferMode transfer_mode)
    : impl_{std::make_unique<Impl>(host_graph, tile_runs, resident_graph, workspace, stream,
                                   run_options, transfer_mode)} {}

BatchedExpansionExecutor::~BatchedExpansionExecutor() = default;

EngineKind BatchedExpansionExecutor::kind() const noexcept { return impl_->kind(); }
BatchedExpansionTransferMode BatchedExpansionExecutor::transfer_mode() const noexcept {
  return impl_->mode();
}
void BatchedExpansionExecutor::set_tuning(const bfn_tuning& tuning) noexcept { impl_->tune(tuning); }

ExpansionBatchExecution BatchedExpansionExecutor::operator()(
    std::span<const RouteQuery> queries, std::span<const BatchQueryFeatures> features,
    const BatchPlanEntry& batch, const ExpansionBatchContext&) {
  return impl_->run(queries, features, batch);
}

// ---------------------------------------------------------------------------
// One device run per planning pass.
// ---------------------------------------------------------------------------
class PassBatchedExpansionExecutor::Impl {
 public:
  Impl(const WeightedGraph& graph, const ResidentDeviceGraph& resident, ReusableBatchedWorkspace& ws,
       GpuRunOptions options, BatchedExpansionTransferMode mode)
      : graph_{graph}, resident_{resident}, ws_{ws}, options_{options}, mode_{mode} {
    if (validate_gpu_run_options(options) != GpuRunOptionsError::none)
      throw std::invalid_argument{"pass expansion executor: invalid run options"};
    if (mode != BatchedExpansionTransferMode::compact_status && mode != BatchedExpansionTransferMode::compact_paths)
      throw std::invalid_argument{"pass expansion executor: transfer mode must be compact_status or compact_paths"};
    if (mode == BatchedExpansionTransferMode::compact_paths && options.instrumentation != InstrumentationLevel::none)
      throw std::invalid_argument{"pass expansion executor: compact paths need no instrumentation"};
    if (!resident.has_upload() || resident.vertex_count() != graph.vertex_count())
      throw std::invalid_argument{"pass expansion executor: the resident graph is not the host graph's"};
    bfn_tuning_default(&tuning_);
    tuning_.tile_sweeps = 1U;  // point Jacobi by default, as the engines
  }

  EngineKind kind() const noexcept { return options_.engine; }
  BatchedExpansionTransferMode mode() const noexcept { return mode_; }
  void tune(const bfn_tuning& t) noexcept { tuning_ = t; }
  std::uint64_t runs() const noexcept { return runs_; }

  ExpansionBatchExecution serve(std::span<const RouteQuery> queries, const BatchPlanEntry& batch,
                                const ExpansionBatchContext& context) {
    const auto t0 = Clock::now();
    const bool run_now = context.batch_index == 0U || !same_pass(queries);
    if (run_now) run_pass(queries);
    ExpansionBatchExecution out;
    out.result.engine_kind = static_cast<std::uint32_t>(options_.engine);
    out.result.control_mode = static_cast<std::uint32_t>(options_.control_mode);
    DeviceRunStatus& st = out.result.status;
    st.final_distance_slot = run_status_.final_distance_slot;
    st.valid_lane_mask = batch.valid_lane_mask;
    std::vector<std::uint32_t> lane_query(batch.lane_width, 0U);
    bool all_converged = true;
    for (std::uint32_t lane = 0; lane < batch.lane_width; ++lane) {
      if (!(batch.valid_lane_mask >> lane & 1U)) continue;
      const std::uint32_t qi = position_of(batch.query_ids_by_lane[lane], batch.expansion_generations_by_lane[lane]);
      lane_query[lane] = qi;
      const bfn_query_status& q = per_query_[qi];
      st.rounds_completed = std::max<std::uint64_t>(st.rounds_completed, q.rounds_completed);
      const LaneMask bit = LaneMask{1U} << lane;
      (q.converged ? st.converged_lane_mask : st.active_lane_mask) |= bit;
      all_converged = all_converged && q.converged != 0U;
    }
    const bool clean = run_status_.error_bits == 0U && all_converged;
    if (clean) {
      st.converged = 1U;
      st.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::converged);
      for (std::uint32_t lane = 0; lane < batch.lane_width; ++lane) {
        if (!(batch.valid_lane_mask >> lane & 1U)) continue;
        const bfn_query_status& q = per_query_[lane_query[lane]];
        if (q.reached_all_targets) st.reached_target_mask |= LaneMask{1U} << lane;
        if (q.bounding_box_miss) st.bounding_box_miss_mask |= LaneMask{1U} << lane;
      }
    } else {
      // The pass stopped before every lane of this batch converged: its stop
      // reason (maximum rounds, or a device error for every lane).
      st.converged = 0U;
      st.stop_reason = run_status_.stop_reason == static_cast<std::uint32_t>(DeviceStopReason::converged)
                           ? static_cast<std::uint32_t>(DeviceStopReason::maximum_rounds)
                           : run_status_.stop_reason;
      st.error_bits = run_status_.error_bits;
      if (st.error_bits != 0U) {
        st.active_lane_mask |= st.converged_lane_mask;
        st.converged_lane_mask = 0U;
      }
    }
    if (validate_device_run_status(st) != DeviceRunStatusError::none)
      throw std::runtime_error{"pass expansion executor produced an inconsistent DeviceRunStatus"};
    if (mode_ != BatchedExpansionTransferMode::compact_paths) return out;

    CompactPathExecutionMetrics& t = out.compact_execution;
    t.host_timing = CompactStageTimingEvidence::measured;
    if (run_now) {
      t.sssp_host_nanoseconds = pass_sssp_ns_;
      t.reconstruction_host_nanoseconds = pass_paths_ns_;
    }
    if (!clean) {
      t.end_to_end_host_nanoseconds = elapsed_ns(t0);
      return out;
    }
    std::uint64_t control = run_now ? 8U : 0U;  // the pass's closing offset word
    for (std::uint32_t lane = 0; lane < batch.lane_width; ++lane) {
      if (!(batch.valid_lane_mask >> lane & 1U)) continue;
      const std::uint32_t qi = lane_query[lane];
      const RouteQuery& q = submitted_[qi];
      CompactPathPayload p;
      p.query_id = q.query_id;
      p.expansion_generation = q.expansion_generation;
      p.target_terminal_to_target = q.target_terminal_to_target;
      p.targets.assign(targets_.begin() + target_offsets_[qi], targets_.begin() + target_offsets_[qi + 1]);
      for (const CompactTargetPath& path : p.targets) control += 8U + (path.vertices.empty() ? 0U : 4U);
      out.compact_paths.push_back(std::move(p));
    }
    std::sort(out.compact_paths.begin(), out.compact_paths.end(),
              [](const CompactPathPayload& a, const CompactPathPayload& b) { return a.query_id < b.query_id; });
    out.compact_transfer = measure_compact_transfer(out.compact_paths);
    out.compact_status_bytes = sizeof(DeviceRunStatus);
    out.compact_error_bytes = control;
    out.compact_total_device_to_host_bytes =
        out.compact_transfer.total_bytes + out.compact_status_bytes + out.compact_error_bytes;
    out.compact_overall_device_to_host_bytes = out.compact_total_device_to_host_bytes;
    t.end_to_end_host_nanoseconds =
        std::max(elapsed_ns(t0), t.sssp_host_nanoseconds + t.reconstruction_host_nanoseconds);
    return out;
  }

 private:
  using Clock = std::chrono::steady_clock;
  static std::uint64_t elapsed_ns(Clock::time_point t0) {
    return static_cast<std::uint64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count());
  }

  bool same_pass(std::span<const RouteQuery> queries) const {
    if (queries.size() != submitted_.size()) return false;
    for (std::size_t i = 0; i < queries.size(); ++i)
      if (queries[i].query_id != submitted_[i].query_id ||
          queries[i].expansion_generation != submitted_[i].expansion_generation ||
          queries[i].selected_tiles != submitted_[i].selected_tiles)
        return false;
    return true;
  }

  std::uint32_t position_of(QueryId id, std::uint32_t generation) const {
    const auto it = std::lower_bound(by_id_.begin(), by_id_.end(), std::make_pair(id.value(), 0U));
    if (it == by_id_.end() || it->first != id.value() || submitted_[it->second].expansion_generation != generation)
      throw std::runtime_error{"a planned lane names a query of no run of this pass"};
    return it->second;
  }

  void run_pass(std::span<const RouteQuery> queries) {
    const auto t0 = Clock::now();
    if (queries.empty()) throw std::invalid_argument{"pass expansion executor: empty pass"};
    submitted_.assign(queries.begin(), queries.end());
    by_id_.clear();
    for (std::uint32_t i = 0; i < queries.size(); ++i) by_id_.push_back({queries[i].query_id.value(), i});
    std::sort(by_id_.begin(), by_id_.end());
    for (std::size_t i = 1; i < by_id_.size(); ++i)
      if (by_id_[i].first == by_id_[i - 1].first)
        throw std::invalid_argument{"pass expansion executor: query ids must be unique"};

    // The pass as one C ABI batch, in submitted order.
    std::vector<std::uint32_t> ids, source_offsets{0U}, sources, target_offsets{0U}, targets, tile_offsets{0U}, tiles;
    for (const RouteQuery& q : queries) {
      if (!validate_route_query(graph_, q).ok())
        throw std::invalid_argument{"pass expansion executor: invalid route query"};
      ids.push_back(q.query_id.value());
      for (VertexId v : q.sources) sources.push_back(v.value());
      for (VertexId v : q.targets) targets.push_back(v.value());
      for (TileId t : q.selected_tiles) tiles.push_back(t.value());
      source_offsets.push_back(static_cast<std::uint32_t>(sources.size()));
      target_offsets.push_back(static_cast<std::uint32_t>(targets.size()));
      tile_offsets.push_back(static_cast<std::uint32_t>(tiles.size()));
    }
    bfn_batch_desc desc{};
    desc.query_count = static_cast<std::uint32_t>(queries.size());
    desc.query_ids = ids.data();
    desc.source_offsets = source_offsets.data();
    desc.sources = sources.data();
    desc.target_offsets = target_offsets.data();
    desc.targets = targets.data();
    desc.tile_offsets = tile_offsets.data();
    desc.selected_tiles = tiles.data();
    desc.engine_by_query = nullptr;
    bfn_run_options o;
    static_assert(sizeof o == sizeof options_);
    std::memcpy(&o, &options_, sizeof o);
    throw_if_error(bfn_batch_submit(ws_.handle(), resident_.handle(), &desc, &o, &tuning_), "pass submit");
    throw_if_error(bfn_batch_wait(ws_.handle()), "pass wait");
    per_query_.assign(queries.size(), bfn_query_status{});
    bfn_work_statistics w{};
    throw_if_error(bfn_batch_status(ws_.handle(), &run_status_, per_query_.data(), &w), "pass status");
    ++runs_;
    const auto t1 = Clock::now();
    pass_sssp_ns_ = static_cast<std::uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count());
    pass_paths_ns_ = 0U;
    target_offsets_ = target_offsets;
    targets_.clear();
    if (mode_ != BatchedExpansionTransferMode::compact_paths) return;

    const std::size_t NT = targets.size();
    std::vector<bfn_target_summary> summaries(NT);
    std::vector<std::uint64_t> offsets(NT + 1U);
    std::uint64_t total = 0;
    throw_if_error(bfn_batch_paths(ws_.handle(), summaries.data(), offsets.data(), nullptr, nullptr, nullptr, 0U,
                                   &total),
                   "pass paths");
    const std::uint64_t n = total ? total : 1U;
    std::vector<std::uint32_t> vertices(n), edges(n);
    std::vector<float> labels(n);
    throw_if_error(bfn_batch_paths(ws_.handle(), summaries.data(), offsets.data(), vertices.data(), labels.data(),
                                   edges.data(), n, &total),
                   "pass paths");
    targets_.resize(NT);
    for (std::size_t i = 0; i < NT; ++i) {
      const bfn_target_summary& sm = summaries[i];
      CompactTargetPath& p = targets_[i];
      p.summary.target = VertexId{sm.target};
      p.summary.selected_source = VertexId{sm.selected_source};
      p.summary.distance = sm.distance;
      p.summary.path_length = sm.path_length;
      p.summary.reached = static_cast<CompactTargetReachStatus>(sm.reached);
      p.summary.reconstruction = static_cast<CompactPathStatus>(sm.reconstruction);
      p.summary.has_selected_source = sm.has_selected_source;
      for (std::uint64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
        p.vertices.push_back(VertexId{vertices[k]});
        p.distance_labels.push_back(labels[k]);
        if (k + 1U < offsets[i + 1]) p.edge_ids.push_back(EdgeId{edges[k]});
      }
    }
    pass_paths_ns_ = elapsed_ns(t1);
  }

  const WeightedGraph& graph_;
  const ResidentDeviceGraph& resident_;
  ReusableBatchedWorkspace& ws_;
  GpuRunOptions options_;
  BatchedExpansionTransferMode mode_;
  bfn_tuning tuning_{};
  std::uint64_t runs_{};
  // The last pass run: its queries (submitted order), per-query records and
  // flattened target paths.
  std::vector<RouteQuery> submitted_;
  std::vector<std::pair<std::uint32_t, std::uint32_t>> by_id_;  // (query id, position)
  bfn_run_status run_status_{};
  std::vector<bfn_query_status> per_query_;
  std::vector<std::uint32_t> target_offsets_;
  std::vector<CompactTargetPath> targets_;
  std::uint64_t pass_sssp_ns_{}, pass_paths_ns_{};
};

PassBatchedExpansionExecutor::PassBatchedExpansionExecutor(const WeightedGraph& host_graph,
                                                           const TileRunLayout64& tile_runs,
                                                           const ResidentDeviceGraph& resident_graph,
                                                           ReusableBatchedWorkspace& workspace,
                                                           const HipStream&, GpuRunOptions run_options,
                                                           BatchedExpansionTransferMode transfer_mode) {
  if (tile_runs.csr_row_run_offsets.size() != static_cast<std::size_t>(host_graph.vertex_count()) + 1U)
    throw std::invalid_argument{"pass expansion executor: tile runs are not the host graph's"};
  impl_ = std::make_unique<Impl>(host_graph, resident_graph, workspace, run_options, transfer_mode);
}

PassBatchedExpansionExecutor::~PassBatchedExpansionExecutor() = default;

EngineKind PassBatchedExpansionExecutor::kind() const noexcept { return impl_->kind(); }
BatchedExpansionTransferMode PassBatchedExpansionExecutor::transfer_mode() const noexcept { return impl_->mode(); }
void PassBatchedExpansionExecutor::set_tuning(const bfn_tuning& tuning) noexcept { impl_->tune(tuning); }
std::uint64_t PassBatchedExpansionExecutor::device_runs() const noexcept { return impl_->runs(); }

ExpansionBatchExecution PassBatchedExpansionExecutor::operator()(std::span<const RouteQuery> queries,
                                                                 std::span<const BatchQueryFeatures>,
                                                                 const BatchPlanEntry& batch,
                                                                 const ExpansionBatchContext& context) {
  return impl_->serve(queries, batch, context);
}

}  // namespace bfnew::hip
