// batched_expansion.hpp -- reachability expansion over batched queries
// (reference: include/bfnew/batched_expansion.hpp:17-340,
// src/batched_expansion.cpp:364-1680).
//
// run_batched_expansion plans the active queries into lane batches, runs each
// batch through an ExpansionBatchRunner, and restarts every query whose
// targets were not reached inside its selected tiles with a larger region:
// its tile box grown by the schedule's margins, then (by policy) the whole
// graph, until it is reached or classified.  A runner receives one planned
// lane batch (BatchPlanEntry) with the queries and features it indexes, runs
// it, and returns one ExpansionBatchExecution.
#pragma once

#include "bfnew/batch_plan.hpp"
#include "bfnew/compact_path_types.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/query.hpp"
#include "bfnew/spatial.hpp"

#include <cstdint>
#include <functional>
#include <optional>
#include <span>
#include <vector>

namespace bfnew {

// There is no default schedule: a caller names one.
enum class ExpansionScheduleKind : std::uint8_t {
  unspecified,
  one_geometric_ring,         // margins +1 tile per expansion
  fixed_larger_ring,          // margins +fixed_ring_size (>= 2)
  doubling_xy_margins,        // margins 1, 2, 4, ...
  hybrid_small_then_doubling, // +1 for the first hybrid_small_expansion_count, then doubling
};

enum class ExpansionTerminalPolicy : std::uint8_t {
  full_region_fallback,  // after the scheduled expansions, one restart on every tile
  explicit_failure,      // classify as expansion_limit / region_stalled instead
};

struct ExpansionSchedulePolicy {
  ExpansionScheduleKind kind{ExpansionScheduleKind::unspecified};
  std::uint32_t fixed_ring_size{};
  std::uint32_t hybrid_small_expansion_count{};
  constexpr bool operator==(const ExpansionSchedulePolicy&) const noexcept = default;
};

[[nodiscard]] constexpr ExpansionSchedulePolicy one_ring_expansion() noexcept {
  return {ExpansionScheduleKind::one_geometric_ring, 0U, 0U};
}
[[nodiscard]] constexpr ExpansionSchedulePolicy fixed_ring_expansion(std::uint32_t ring) noexcept {
  return {ExpansionScheduleKind::fixed_larger_ring, ring, 0U};
}
[[nodiscard]] constexpr ExpansionSchedulePolicy doubling_margin_expansion() noexcept {
  return {ExpansionScheduleKind::doubling_xy_margins, 0U, 0U};
}
[[nodiscard]] constexpr ExpansionSchedulePolicy hybrid_margin_expansion(std::uint32_t small) noexcept {
  return {ExpansionScheduleKind::hybrid_small_then_doubling, 0U, small};
}

struct BatchedExpansionOptions {
  GpuRunOptions run_options{};
  BatchPlannerPolicy planner_policy{};
  // Nonzero identity of runner settings outside the options above; folded
  // into the schedule-comparison fingerprint.
  std::uint64_t execution_configuration_fingerprint{};
  ExpansionSchedulePolicy schedule{};
  std::uint32_t maximum_expansions{};  // scheduled expansions; the fallback is extra
  ExpansionTerminalPolicy terminal_policy{ExpansionTerminalPolicy::full_region_fallback};
  std::uint32_t enable_compact_paths{};  // 0 or 1

  // Field-wise; GpuRunOptions has padding, so it is not compared as bytes.
  [[nodiscard]] bool operator==(const BatchedExpansionOptions& other) const noexcept;
};

enum class BatchedExpansionOptionsError : std::uint8_t {
  none,
  invalid_run_options,
  invalid_planner_policy,
  missing_execution_configuration_fingerprint,
  unspecified_schedule,
  invalid_schedule,
  invalid_fixed_ring_size,
  unexpected_fixed_ring_size,
  invalid_hybrid_small_expansion_count,
  unexpected_hybrid_small_expansion_count,
  invalid_terminal_policy,
  invalid_compact_paths_flag,
};

[[nodiscard]] BatchedExpansionOptionsError validate_batched_expansion_options(
    const BatchedExpansionOptions& options) noexcept;

enum class ExpansionWorkEvidence : std::uint8_t {
  unavailable,
  measured,
};

struct ExpansionBatchContext {
  std::uint64_t execution_ordinal{};
  std::uint32_t planning_pass{};
  std::uint32_t batch_index{};
  bool retry_pass{};
  constexpr bool operator==(const ExpansionBatchContext&) const noexcept = default;
};

struct ExpansionBatchExecution {
  GpuSsspResult result{};
  ExpansionWorkEvidence work_evidence{ExpansionWorkEvidence::unavailable};
  // Edge records read by the device (shared by the lanes of a tile visit)
  // and per-lane admitted relaxations; meaningful when measured.
  std::uint64_t shared_edge_work{};
  std::uint64_t logical_lane_edge_work{};
  // Vertex-major, lane-contiguous labels; empty unless the runner downloads
  // them (graph.vertex_count() * batch.lane_width entries otherwise).
  std::vector<float> final_distances;
  // Compact mode: one payload per valid lane of a cleanly converged batch,
  // sorted by QueryId.
  std::vector<CompactPathPayload> compact_paths;
  CompactTransferAccounting compact_transfer{};
  std::uint64_t compact_status_bytes{};
  std::uint64_t compact_error_bytes{};
  std::uint64_t compact_total_device_to_host_bytes{};
  std::uint64_t compact_controller_poll_count{};
  std::uint64_t compact_controller_poll_bytes{};
  std::uint64_t compact_overall_device_to_host_bytes{};
  CompactPathExecutionMetrics compact_execution{};
};

using ExpansionBatchRunner = std::function<ExpansionBatchExecution(
    std::span<const RouteQuery> queries, std::span<const BatchQueryFeatures> features,
    const BatchPlanEntry& batch, const ExpansionBatchContext& context)>;

enum class ExpansionQueryDisposition : std::uint8_t {
  reached,
  unreachable_in_full_region,
  expansion_limit,
  region_stalled,
  identity_or_count_overflow,
  engine_failure,
};

struct ExpansionQueryOutcome {
  RouteQuery final_query;  // the last region tried, with its generation
  ExpansionQueryDisposition disposition{ExpansionQueryDisposition::engine_failure};
  std::uint32_t attempts{};
  std::uint32_t scheduled_expansions{};
  std::uint32_t total_expansions{};  // scheduled + the fallback, if used
  bool used_full_region_fallback{};
  std::uint64_t selected_vertex_count{};
  std::uint64_t selected_edge_count{};
  std::uint32_t terminal_stop_reason{};
  std::uint32_t terminal_error_bits{};
  std::vector<float> final_distances;  // when the runner returned labels
  // Compact mode: the reached generation's paths, or a full-region miss's
  // classified targets.
  std::optional<CompactPathPayload> compact_paths;
  [[nodiscard]] constexpr bool reached() const noexcept {
    return disposition == ExpansionQueryDisposition::reached;
  }
};

struct ExpansionBatchTrace {
  ExpansionBatchContext context{};
  std::uint32_t lane_width{};
  LaneMask valid_lane_mask{};
  LaneMask reached_lane_mask{};
  LaneMask miss_lane_mask{};
  std::vector<QueryId> query_ids_by_lane;
  std::vector<std::uint32_t> expansion_generations_by_lane;
  std::uint64_t union_vertex_count{};
  std::uint64_t union_edge_estimate{};
  std::uint64_t selected_lane_vertex_count{};
  std::uint64_t selected_lane_edge_estimate{};
  std::uint64_t shared_edge_work{};
  std::uint64_t logical_lane_edge_work{};
  ExpansionWorkEvidence work_evidence{ExpansionWorkEvidence::unavailable};
  std::uint32_t stop_reason{};
  std::uint32_t error_bits{};
  bool operator==(const ExpansionBatchTrace&) const = default;
};

struct BatchedExpansionMetrics {
  std::uint64_t input_queries{};
  std::uint64_t initial_reached_queries{};
  std::uint64_t reached_queries{};
  std::uint64_t unreachable_full_region_queries{};
  std::uint64_t expansion_limit_queries{};
  std::uint64_t stalled_region_queries{};
  std::uint64_t identity_or_count_overflow_queries{};
  std::uint64_t engine_failure_queries{};

  std::uint64_t planning_passes{};
  std::uint64_t batches_executed{};
  std::uint64_t initial_batches_executed{};
  std::uint64_t retry_batches_executed{};
  std::uint64_t failed_lane_observations{};
  std::uint64_t failed_origin_valid_lane_observations{};
  std::uint64_t retry_valid_lane_observations{};
  std::uint64_t retry_lane_capacity{};

  std::uint64_t scheduled_expansions{};
  std::uint64_t full_region_fallbacks{};
  std::uint64_t repeated_selected_edge_estimate{};  // edges of regions discarded by restarts
  std::uint64_t attempted_selected_vertex_count{};
  std::uint64_t attempted_selected_edge_estimate{};
  std::uint64_t final_selected_tile_count{};
  std::uint64_t final_selected_vertex_count{};
  std::uint64_t final_selected_edge_count{};

  // Measured only when every batch measured it.  Retry work: batches of
  // planning passes after the first.  Failed-batch work: whole batches with
  // at least one miss (shared work of a mixed batch cannot be split between
  // its reached and missed lanes).
  ExpansionWorkEvidence work_evidence{ExpansionWorkEvidence::measured};
  std::uint64_t work_measured_batches{};
  std::uint64_t shared_edge_work{};
  std::uint64_t logical_lane_edge_work{};
  std::uint64_t retry_work_measured_batches{};
  std::uint64_t retry_shared_edge_work{};
  std::uint64_t retry_logical_lane_edge_work{};
  std::uint64_t failed_batch_work_measured_batches{};
  std::uint64_t failed_batch_shared_edge_work{};
  std::uint64_t failed_batch_logical_lane_edge_work{};

  // Sum of the batches' counters; expansion_count = scheduled + fallbacks.
  DeviceWorkStatistics device_work{};

  // Stage timing is measured only when every batch measured it.
  CompactStageTimingEvidence compact_device_timing{CompactStageTimingEvidence::unavailable};
  CompactStageTimingEvidence compact_host_timing{CompactStageTimingEvidence::unavailable};
  std::uint64_t compact_device_timing_measured_batches{};
  std::uint64_t compact_host_timing_measured_batches{};
  std::uint64_t sssp_device_nanoseconds{};
  std::uint64_t reconstruction_device_nanoseconds{};
  std::uint64_t result_transfer_device_nanoseconds{};
  std::uint64_t sssp_host_nanoseconds{};
  std::uint64_t reconstruction_host_nanoseconds{};
  std::uint64_t result_transfer_host_nanoseconds{};
  std::uint64_t compact_end_to_end_host_nanoseconds{};
  std::uint64_t geometric_expansion_host_nanoseconds{};
  CompactTransferAccounting compact_transfer{};
  std::uint64_t compact_status_bytes{};
  std::uint64_t compact_error_bytes{};
  std::uint64_t compact_total_device_to_host_bytes{};
  std::uint64_t compact_controller_poll_count{};
  std::uint64_t compact_controller_poll_bytes{};
  std::uint64_t compact_overall_device_to_host_bytes{};

  // Host wall time (not part of deterministic comparisons).
  std::uint64_t initial_planning_nanoseconds{};
  std::uint64_t replanning_nanoseconds{};
  std::uint64_t execution_nanoseconds{};
  std::uint64_t total_nanoseconds{};
  // floor(queries per second * 1000); zero when total_nanoseconds is zero.
  std::uint64_t all_query_throughput_milliqueries_per_second{};

  // FNV-1a identity of graph, initial queries, engine/control, planner,
  // limit, terminal policy and compact mode -- not of the schedule, so runs
  // of different schedules on the same work compare.
  std::uint64_t schedule_comparison_fingerprint{};
  // [i] = queries that ended after i expansions (fallback included).
  std::vector<std::uint64_t> expansion_count_histogram;
};

struct BatchedExpansionRunResult {
  std::vector<ExpansionQueryOutcome> queries;  // QueryId order
  std::vector<ExpansionBatchTrace> trace;
  BatchedExpansionMetrics metrics{};
};

// Throws std::invalid_argument for invalid options, a missing runner, a
// directory or tile runs that are not the graph's, or invalid/duplicate
// queries; std::logic_error when the runner returns a result that does not
// describe the batch it was given.  The runner must start every batch from
// the queries' sources; a batch that stops for any reason other than clean
// convergence classifies its lanes as engine_failure and never expands.
// Graph checks are memoised with the graph; tile runs are compared with the
// graph's element by element unless they are graph_tile_runs(graph) itself.
[[nodiscard]] BatchedExpansionRunResult run_batched_expansion(
    const WeightedGraph& graph, const TileDirectory& directory, const TileRunLayout64& tile_runs,
    std::span<const RouteQuery> queries, const BatchedExpansionOptions& options,
    const ExpansionBatchRunner& runner);

// One schedule's run, reduced to what ranks schedules against each other
// (reference: include/bfnew/batched_expansion.hpp:352-394).
struct ExpansionScheduleEvidence {
  ExpansionSchedulePolicy schedule{};
  std::uint64_t input_queries{};
  std::uint64_t reached_queries{};
  std::uint64_t terminal_failures{};
  std::uint64_t batches_executed{};
  std::uint64_t total_expansions{};
  std::uint64_t attempted_selected_edge_estimate{};
  ExpansionWorkEvidence work_evidence{ExpansionWorkEvidence::unavailable};
  std::uint64_t shared_edge_work{};
  std::uint64_t logical_lane_edge_work{};
  std::uint64_t comparison_fingerprint{};
  bool campaign_valid{};  // no engine/identity failure, totals representable
  constexpr bool operator==(const ExpansionScheduleEvidence&) const noexcept = default;
};

[[nodiscard]] ExpansionScheduleEvidence make_expansion_schedule_evidence(
    const ExpansionSchedulePolicy& schedule, const BatchedExpansionRunResult& result) noexcept;

// The unique best of exactly one valid record per schedule kind over the same
// comparison fingerprint: more reached queries, then fewer failures, then
// (when every record measured work) less logical then shared work, then less
// attempted selected-edge estimate, fewer expansions, fewer batches.  No
// answer (nullopt) on a tie or on records that are incomplete or unlike.
[[nodiscard]] std::optional<ExpansionSchedulePolicy> select_expansion_schedule_from_evidence(
    std::span<const ExpansionScheduleEvidence> evidence) noexcept;

}  // namespace bfnew
