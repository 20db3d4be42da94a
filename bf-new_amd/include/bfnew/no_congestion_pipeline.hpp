// no_congestion_pipeline.hpp -- what follows the relaxation path: per-query
// compact results, their transfer accounting, a stage timing ledger and a
// sampled bounded-vs-unbounded path quality record
// (reference: include/bfnew/no_congestion_pipeline.hpp:13-236,
// src/no_congestion_pipeline.cpp:62-760).
#pragma once

#include "bfnew/batched_expansion.hpp"
#include "bfnew/compact_paths.hpp"

#include <cstdint>
#include <span>
#include <vector>

namespace bfnew {

enum class PipelineTimingEvidence : std::uint8_t { unavailable, measured };

// Host-wall and device-event observations of one stage; a zero means zero
// only when its evidence is measured.
struct PipelineStageTiming {
  PipelineTimingEvidence host_evidence{PipelineTimingEvidence::unavailable};
  std::uint64_t host_nanoseconds{};
  PipelineTimingEvidence device_evidence{PipelineTimingEvidence::unavailable};
  double device_milliseconds{};
  constexpr bool operator==(const PipelineStageTiming&) const noexcept = default;
};

[[nodiscard]] constexpr PipelineStageTiming measured_host_stage(std::uint64_t nanoseconds) noexcept {
  return PipelineStageTiming{PipelineTimingEvidence::measured, nanoseconds, PipelineTimingEvidence::unavailable,
                             0.0};
}

// Warm stages (planning, sssp, expansion, controller orchestration,
// reconstruction, result transfer) partition warm_all_query when all are
// measured (else they may not exceed it); cold_pipeline is the sum of
// artifact load, graph upload and cold execution when all three are
// measured.  Planning, expansion, orchestration and the enclosing intervals
// are host work and carry no device evidence.
struct NoCongestionStageLedger {
  PipelineStageTiming cold_artifact_load;
  PipelineStageTiming graph_upload;
  PipelineStageTiming batch_planning;
  PipelineStageTiming sssp;
  PipelineStageTiming expansion;
  PipelineStageTiming controller_orchestration;
  PipelineStageTiming reconstruction;
  PipelineStageTiming result_transfer;
  PipelineStageTiming warm_all_query;
  PipelineStageTiming cold_execution;
  PipelineStageTiming cold_pipeline;
  constexpr bool operator==(const NoCongestionStageLedger&) const noexcept = default;
};

enum class NoCongestionStageLedgerError : std::uint8_t {
  none,
  invalid_evidence,
  unavailable_host_has_value,
  unavailable_device_has_value,
  invalid_device_duration,
  cpu_only_device_evidence,
  warm_evidence_mismatch,
  warm_sum_overflow,
  warm_sum_mismatch,
  cold_evidence_mismatch,
  cold_sum_overflow,
  cold_sum_mismatch,
};

// Fills cold_pipeline and validates; throws std::invalid_argument for an
// invalid ledger, std::overflow_error when the cold sum overflows.
[[nodiscard]] NoCongestionStageLedger make_no_congestion_stage_ledger(
    PipelineStageTiming cold_artifact_load, PipelineStageTiming graph_upload, PipelineStageTiming batch_planning,
    PipelineStageTiming sssp, PipelineStageTiming expansion, PipelineStageTiming controller_orchestration,
    PipelineStageTiming reconstruction, PipelineStageTiming result_transfer, PipelineStageTiming warm_all_query,
    PipelineStageTiming cold_execution);

[[nodiscard]] NoCongestionStageLedgerError validate_no_congestion_stage_ledger(
    const NoCongestionStageLedger& ledger) noexcept;

struct NoCongestionResultAccounting {
  std::uint64_t query_count{};
  std::uint64_t target_summary_count{};
  std::uint64_t complete_path_count{};
  std::uint64_t unreachable_target_count{};
  std::uint64_t terminal_failure_target_count{};
  std::uint64_t reconstruction_failure_target_count{};
  // Final results as one serialization (summaries + path arenas, 4 bytes per
  // vertex, label and edge id) plus one status record per batch: a sizing
  // model, not a measurement.
  std::uint64_t modeled_batch_status_count{};
  std::uint64_t modeled_batch_status_bytes{};
  CompactTransferAccounting final_result_serialization{};
  std::uint64_t modeled_final_transfer_bytes{};
  // The device's own byte records of the run (retries included), when the
  // runner measured them.
  CompactStageTimingEvidence device_transfer_evidence{CompactStageTimingEvidence::unavailable};
  CompactTransferAccounting actual_compact_device_transfer{};
  std::uint64_t actual_status_bytes{};
  std::uint64_t actual_error_bytes{};
  std::uint64_t actual_compact_total_device_to_host_bytes{};
  std::uint64_t actual_controller_poll_count{};
  std::uint64_t actual_controller_poll_bytes{};
  std::uint64_t actual_overall_device_to_host_bytes{};
  constexpr bool operator==(const NoCongestionResultAccounting&) const noexcept = default;
};

// Results must be in strictly increasing QueryId order, each consistent with
// its disposition (std::invalid_argument otherwise).  With device_metrics
// whose compact total is nonzero, its records must add up (payload parts to
// the payload total; payload + status + control words to the compact total;
// one status record per batch; polls and their bytes together; compact +
// polls to the overall total) and are copied as measured.
[[nodiscard]] NoCongestionResultAccounting measure_no_congestion_result_transfer(
    std::span<const CompactQueryResult> results, std::uint64_t modeled_batch_status_count,
    const BatchedExpansionMetrics* device_metrics = nullptr);

enum class PathQualitySamplingMethod : std::uint8_t { splitmix64_query_id };

struct CompactPathQualityObservation {
  QueryId query_id{};
  VertexId target{};
  float bounded_distance{};
  float unbounded_distance{};
  std::uint32_t bounded_path_length{};
  std::uint32_t unbounded_path_length{};
  double absolute_cost_inflation{};
  std::int64_t absolute_path_length_inflation{};
  double cost_inflation_ratio{};
  double path_length_inflation_ratio{};
  constexpr bool operator==(const CompactPathQualityObservation&) const noexcept = default;
};

struct CompactPathQualitySample {
  PathQualitySamplingMethod method{PathQualitySamplingMethod::splitmix64_query_id};
  std::uint64_t selection_seed{};
  std::uint64_t population_query_count{};
  std::uint64_t requested_sample_query_count{};
  std::uint64_t sampled_query_count{};
  std::uint64_t finite_target_pairs{};
  std::vector<QueryId> sampled_query_ids;
  std::vector<CompactPathQualityObservation> observations;
  double absolute_cost_inflation_p50{}, absolute_cost_inflation_p95{}, absolute_cost_inflation_p99{},
      absolute_cost_inflation_max{};
  double cost_ratio_p50{}, cost_ratio_p95{}, cost_ratio_p99{}, cost_ratio_max{};
  double absolute_path_length_inflation_p50{}, absolute_path_length_inflation_p95{},
      absolute_path_length_inflation_p99{}, absolute_path_length_inflation_max{};
  double path_length_ratio_p50{}, path_length_ratio_p95{}, path_length_ratio_p99{}, path_length_ratio_max{};
  bool operator==(const CompactPathQualitySample&) const = default;
};

// The sampled queries are the first sample_query_count by splitmix64(seed ^
// query id) (ties by id), so the sample does not depend on input order.  For
// every complete target of a sampled query its bounded cost and edge count
// are compared with the unbounded shortest path from the query's sources on
// the whole graph (fp32 Dijkstra on the host; the path takes the
// smallest-EdgeId tight predecessor, backtracking off cycles).  Percentiles
// are nearest-rank.  Inflation is evidence about region bounds, not a
// correctness check.
[[nodiscard]] CompactPathQualitySample sample_compact_path_quality(const WeightedGraph& graph,
                                                                   std::span<const ExpansionQueryOutcome> outcomes,
                                                                   std::span<const CompactQueryResult> results,
                                                                   std::size_t sample_query_count,
                                                                   std::uint64_t selection_seed);

struct NoCongestionPipelineOptions {
  PipelineStageTiming cold_artifact_load{};
  PipelineStageTiming graph_upload{};
  PipelineStageTiming cold_execution{};
  std::size_t quality_sample_query_count{};
  std::uint64_t quality_selection_seed{};
};

struct NoCongestionPipelineResult {
  BatchedExpansionRunResult expansion;
  std::vector<CompactQueryResult> compact_results;  // QueryId order
  NoCongestionResultAccounting result_accounting{};
  NoCongestionStageLedger timing{};
  CompactPathQualitySample quality{};
};

// run_batched_expansion in compact-path mode through `runner` (a device
// executor in production), then the results, their accounting, the stage
// ledger (planning, sssp, expansion, reconstruction and transfer from the
// run's metrics; orchestration is the rest of the run) and the quality
// sample.  Throws std::invalid_argument unless options.enable_compact_paths.
[[nodiscard]] NoCongestionPipelineResult run_no_congestion_pipeline(
    const WeightedGraph& graph, const TileDirectory& directory, const TileRunLayout64& tile_runs,
    std::span<const RouteQuery> queries, const BatchedExpansionOptions& options, const ExpansionBatchRunner& runner,
    const NoCongestionPipelineOptions& pipeline_options = {});

}  // namespace bfnew
