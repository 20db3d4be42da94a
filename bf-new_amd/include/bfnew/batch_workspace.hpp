// batch_workspace.hpp -- the reference's byte model of a lane-batched
// workspace, admission by device budget, reusable reservations and the
// recorded strategy decision (reference: include/bfnew/batch_workspace.hpp:13-180,
// src/batch_workspace.cpp).
//
// The model prices the reference's lane-matrix layouts: distance labels for
// lane_width lanes per vertex, either for every graph vertex
// (full_graph_vertex_major) or for the batch's union tiles only
// (compact_union_tiles, plus one u32 bias per tile), times one or two
// distance slots; tile-run admission as one mask per run (retained) or as
// descriptors of the nonzero runs with per-union-vertex offsets; and the
// batch metadata the device reads.  The engines of this library store labels
// per (query, selected tile) slot instead; bfn_batch_estimate
// (include/bfnew_amd.h) prices that layout.  A caller sizing batches for the
// reference's engines keeps this API and its numbers.
#pragma once

#include "bfnew/batch_layout.hpp"
#include "bfnew/batch_plan.hpp"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bfnew {

enum class BatchVertexStorageStrategy : std::uint8_t {
  full_graph_vertex_major,
  compact_union_tiles,
};

struct BatchWorkspaceBudget {
  std::uint64_t device_capacity_bytes{};
  std::uint64_t resident_graph_bytes{};
  std::uint64_t explicit_reserve_bytes{};
  constexpr bool operator==(const BatchWorkspaceBudget&) const noexcept = default;
};

struct BatchWorkspaceStrategyTiming {
  std::uint64_t mapping_build_nanoseconds{};
  std::uint64_t run_build_nanoseconds{};
  constexpr bool operator==(const BatchWorkspaceStrategyTiming&) const noexcept = default;
};

struct BatchWorkspaceBuildMeasurements {
  std::uint64_t compact_mapping_build_nanoseconds{};
  std::uint64_t retained_run_build_nanoseconds{};
  std::uint64_t descriptor_run_build_nanoseconds{};
  constexpr bool operator==(const BatchWorkspaceBuildMeasurements&) const noexcept = default;
};

struct BatchWorkspaceEstimate {
  BatchVertexStorageStrategy vertex_storage{BatchVertexStorageStrategy::full_graph_vertex_major};
  BatchRunRepresentation run_representation{BatchRunRepresentation::retained_per_run_masks};
  std::uint32_t lane_width{};
  std::uint32_t distance_slot_count{};

  std::uint64_t storage_vertex_count{};
  std::uint64_t union_vertex_count{};
  std::uint64_t selected_lane_vertex_count{};
  std::uint64_t allocated_lane_vertex_count{};
  std::uint64_t wasted_lane_vertex_count{};

  std::uint64_t distance_bytes{};
  std::uint64_t tile_mapping_bytes{};
  std::uint64_t run_storage_bytes{};
  std::uint64_t descriptor_offset_bytes{};
  std::uint64_t batch_metadata_bytes{};
  std::uint64_t total_workspace_bytes{};

  // Preparation writes implied by the description's counters.
  std::uint64_t distance_reset_bytes{};
  std::uint64_t tile_mapping_write_bytes{};
  std::uint64_t run_preparation_write_bytes{};
  std::uint64_t total_preparation_write_bytes{};

  std::uint64_t mapping_build_nanoseconds{};
  std::uint64_t run_build_nanoseconds{};

  std::uint64_t active_csr_runs{};
  std::uint64_t active_csc_runs{};
  std::uint64_t zero_csr_runs{};
  std::uint64_t zero_csc_runs{};
  std::uint64_t maximum_concurrent_workspaces{};
  bool reusable_allocation{};
  constexpr bool operator==(const BatchWorkspaceEstimate&) const noexcept = default;
};

// storage_vertex_count x lane_width x slots x 4 bytes, for widths
// 1 / 8 / 16 / 32 and one or two slots (std::invalid_argument otherwise;
// std::overflow_error past 64 bits).
[[nodiscard]] std::uint64_t estimate_vertex_major_distance_bytes(std::uint64_t storage_vertex_count,
                                                                 std::uint32_t lane_width,
                                                                 std::uint32_t distance_slot_count);

// The byte model of one strategy row for a prepared description
// (std::invalid_argument when the description was not prepared for that
// row, or its counters are inconsistent with the tile runs).  Run storage
// and run preparation writes price the larger orientation; a
// device-materialised row scans the selected CSC columns to count admitted
// runs and their initialisation writes.
[[nodiscard]] BatchWorkspaceEstimate estimate_batch_workspace(
    const WeightedGraph& graph, const TileRunLayout64& tile_runs, const BatchPlanEntry& batch,
    const BatchDeviceDescription& description, BatchVertexStorageStrategy vertex_storage,
    BatchRunRepresentation run_representation, std::uint32_t distance_slot_count,
    const BatchWorkspaceBudget& budget, const BatchWorkspaceStrategyTiming& timing = {});

// The 2 x 2 matrix (full / compact vertex storage x retained / descriptor
// runs), in that order.
[[nodiscard]] std::vector<BatchWorkspaceEstimate> compare_batch_workspace_strategies(
    const WeightedGraph& graph, const TileRunLayout64& tile_runs, const BatchPlanEntry& batch,
    const BatchDeviceDescription& retained_description,
    const BatchDeviceDescription& descriptor_description, std::uint32_t distance_slot_count,
    const BatchWorkspaceBudget& budget, const BatchWorkspaceBuildMeasurements& measurements = {});

struct BatchWorkspaceCapacity {
  std::uint64_t distance_bytes{};
  std::uint64_t tile_mapping_bytes{};
  std::uint64_t run_storage_bytes{};
  std::uint64_t descriptor_offset_bytes{};
  std::uint64_t batch_metadata_bytes{};
  [[nodiscard]] std::uint64_t total_bytes() const;
};

struct BatchWorkspaceReservationResult {
  bool capacity_grew{};
  std::uint64_t generation{};
};

// Per-component capacities that only grow (doubling from the first request
// until a request fits); every reserve is a new generation.
class ReusableBatchWorkspaceReservation final {
 public:
  [[nodiscard]] BatchWorkspaceReservationResult reserve(const BatchWorkspaceEstimate& estimate);
  [[nodiscard]] const BatchWorkspaceCapacity& capacity() const noexcept { return capacity_; }
  [[nodiscard]] std::uint64_t growth_events() const noexcept { return growth_events_; }

 private:
  BatchWorkspaceCapacity capacity_{};
  std::uint64_t generation_{};
  std::uint64_t growth_events_{};
};

enum class BatchMeasurementScope : std::uint8_t {
  bounded_synthetic,
  real_query_corpus,
};

struct BatchWorkspaceDecision {
  std::uint64_t plan_fingerprint{};
  BatchMeasurementScope measurement_scope{BatchMeasurementScope::bounded_synthetic};
  BatchWorkspaceBudget budget{};
  std::uint64_t graph_vertex_count{};
  std::uint64_t graph_tile_count{};
  std::uint32_t lane_width{};
  std::uint32_t distance_slot_count{};
  std::vector<BatchWorkspaceEstimate> compared_strategies;
  BatchVertexStorageStrategy selected_vertex_storage{BatchVertexStorageStrategy::full_graph_vertex_major};
  BatchRunRepresentation selected_run_representation{BatchRunRepresentation::retained_per_run_masks};
  std::uint64_t selected_mapping_build_nanoseconds{};
  std::uint64_t selected_run_build_nanoseconds{};
  std::string quantitative_reason;
};

// Every row's formulas recomputed from its own counts, the four rows once
// each and consistent with one another, and the selected row present,
// admissible (at least one workspace fits) and carrying its timings.
[[nodiscard]] bool validate_batch_workspace_decision(const BatchWorkspaceDecision& decision) noexcept;

// Text record of a valid decision (std::invalid_argument otherwise): the
// reference's bfnew_phase13_workspace_decision_v1 key=value header and one
// tab-separated row per strategy.
[[nodiscard]] std::string serialize_batch_workspace_decision(const BatchWorkspaceDecision& decision);

}  // namespace bfnew
