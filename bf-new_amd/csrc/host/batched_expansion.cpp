// batched_expansion.cpp -- BatchedExpansionExecutor: one planned lane batch
// through a batched engine (reference behaviour:
// src/hip/batched_expansion.hip.cpp:459 and its payload rule :117-163).
#include "bfnew/hip/batched_expansion.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace bfnew::hip {
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

}  // namespace bfnew::hip
