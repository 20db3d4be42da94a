// batch_layout.hpp -- the host description of one batch of route queries
// (reference: include/bfnew/batch_layout.hpp:23-120, the fields the
// relaxation path reads).
//
// A batch has lane_width lanes; lane l is valid when bit l of
// valid_lane_mask is set and then carries one query: its canonical sources
// sources[source_offsets[l] .. source_offsets[l+1]), its canonical targets
// likewise, and the tiles T with bit l set in tile_lane_masks[T].  Invalid
// lanes have empty slices and no tile bits.  The engines of this library
// store labels per (lane, selected tile) and never per (vertex, lane) of the
// whole graph, so the run-mask and compact-mapping images of the reference
// description have no counterpart here.
#pragma once

#include "bfnew/device_layout.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"

#include <cstdint>
#include <span>
#include <vector>

namespace bfnew {

struct BatchVertexRange {
  std::uint32_t begin{};
  std::uint32_t end{};
  LaneMask lane_mask{};
  constexpr bool operator==(const BatchVertexRange&) const noexcept = default;
};
static_assert(sizeof(BatchVertexRange) == 12U);

// How a lane-batched description records which lanes each tile run admits
// (reference: include/bfnew/batch_layout.hpp:13-21).  The engines of this
// library derive admission from per-slot neighbour tables and use none of
// these images; they are kept for the reference's workspace byte model.
enum class BatchRunRepresentation : std::uint8_t {
  retained_per_run_masks,
  compact_nonzero_descriptors,
  device_materialized_run_masks,
};

struct RunLaneMaskDescriptor {
  std::uint32_t run_id{};
  LaneMask lane_mask{};
  constexpr bool operator==(const RunLaneMaskDescriptor&) const noexcept = default;
};
static_assert(sizeof(RunLaneMaskDescriptor) == 8U);

// Host preparation work of the run images: runs visited and admitted per
// orientation, and the retained-mask / descriptor entries initialised,
// cleared (the previous batch's ledger) and written.
struct BatchRunPreparationReport {
  std::uint64_t csr_runs_visited{};
  std::uint64_t csc_runs_visited{};
  std::uint64_t active_csr_runs{};
  std::uint64_t active_csc_runs{};
  std::uint64_t csr_lane_edge_pairs{};
  std::uint64_t csc_lane_edge_pairs{};
  std::uint64_t retained_entries_initialized{};
  std::uint64_t retained_entries_cleared{};
  std::uint64_t retained_entries_written{};
  std::uint64_t descriptor_entries_written{};
  std::uint64_t csr_retained_entries_initialized{};
  std::uint64_t csc_retained_entries_initialized{};
  std::uint64_t csr_retained_entries_cleared{};
  std::uint64_t csc_retained_entries_cleared{};
  std::uint64_t csr_retained_entries_written{};
  std::uint64_t csc_retained_entries_written{};
  std::uint64_t csr_descriptor_entries_written{};
  std::uint64_t csc_descriptor_entries_written{};
  constexpr bool operator==(const BatchRunPreparationReport&) const noexcept = default;
};

struct BatchCompactMappingReport {
  std::uint64_t entries_initialized{};
  std::uint64_t entries_cleared{};
  std::uint64_t entries_written{};
  constexpr bool operator==(const BatchCompactMappingReport&) const noexcept = default;
};

struct BatchDeviceDescription {
  std::uint32_t lane_width{};
  LaneMask valid_lane_mask{};
  LaneMask reached_lane_mask{};
  LaneMask miss_lane_mask{};

  std::vector<std::uint32_t> query_ids_by_lane;
  std::vector<std::uint32_t> expansion_generations_by_lane;
  std::vector<std::uint32_t> source_offsets;  // [lane_width + 1]
  std::vector<std::uint32_t> sources;
  std::vector<std::uint32_t> target_offsets;  // [lane_width + 1]
  std::vector<std::uint32_t> targets;
  std::vector<std::uint64_t> selected_vertex_counts_by_lane;
  std::vector<std::uint64_t> selected_edge_estimates_by_lane;

  std::vector<std::uint32_t> union_tiles;          // sorted
  std::vector<LaneMask> tile_lane_masks;           // [tile_count]
  std::vector<BatchVertexRange> selected_vertex_ranges;  // one per union tile

  // Optional compact-label mapping (prepare_compact_vertex_mapping): for a
  // union tile T and global vertex v of T, the packed index is
  // v - compact_vertex_biases_by_tile[T].  touched_compact_tiles is the
  // ledger of written entries, cleared by the next mapping build.
  std::vector<std::uint32_t> compact_vertex_biases_by_tile;
  std::vector<std::uint32_t> touched_compact_tiles;
  bool compact_vertex_mapping_valid{};
  BatchCompactMappingReport compact_mapping_report{};

  // Optional run images (the run-representation preparation overload).
  BatchRunRepresentation run_representation{BatchRunRepresentation::retained_per_run_masks};
  bool run_representation_initialized{};
  std::vector<LaneMask> csr_run_lane_masks;   // retained: one per run
  std::vector<LaneMask> csc_run_lane_masks;
  std::vector<std::uint32_t> touched_csr_runs;  // retained: nonzero entries
  std::vector<std::uint32_t> touched_csc_runs;
  std::vector<RunLaneMaskDescriptor> csr_run_descriptors;  // descriptors: nonzero runs
  std::vector<RunLaneMaskDescriptor> csc_run_descriptors;
  std::vector<std::uint32_t> csr_descriptor_offsets_by_union_vertex;  // [union V + 1]
  std::vector<std::uint32_t> csc_descriptor_offsets_by_union_vertex;
  BatchRunPreparationReport run_report{};
};

// Lays out `queries` in lanes 0 .. queries.size()-1 of a batch of
// lane_width lanes (1 <= queries.size() <= lane_width <= 32).  Throws
// std::invalid_argument for an invalid query or shape.
void prepare_batch_device_description(const WeightedGraph& graph,
                                      std::span<const RouteQuery> queries,
                                      std::uint32_t lane_width, BatchDeviceDescription& output);

enum class BatchLayoutValidationErrorCode : std::uint8_t {
  none,
  invalid_batch_shape,
  invalid_lane_identity,
  invalid_source_offsets,
  invalid_target_offsets,
  invalid_terminal_payload,
  invalid_union_tiles,
  invalid_tile_lane_masks,
  invalid_vertex_ranges,
  nonzero_initial_result_mask,
  invalid_compact_vertex_mapping,
};

struct BatchLayoutValidationResult {
  static constexpr std::size_t no_position = static_cast<std::size_t>(-1);
  BatchLayoutValidationErrorCode code{BatchLayoutValidationErrorCode::none};
  std::size_t position{no_position};
  [[nodiscard]] constexpr bool ok() const noexcept {
    return code == BatchLayoutValidationErrorCode::none;
  }
};

// Structural check of a description against the graph: lane shape, canonical
// terminal slices inside valid lanes only, sorted union tiles agreeing with
// the dense masks, terminals owned by a selected tile of their lane, and one
// vertex range per union tile.
[[nodiscard]] BatchLayoutValidationResult validate_batch_device_description(
    const WeightedGraph& graph, const BatchDeviceDescription& description) noexcept;

struct BatchPlanEntry;
struct BatchQueryFeatures;

// The description of one planned batch with its run images
// (reference: src/batch_layout.cpp:1056-1296): the lane fields of the
// plan-entry overload (batch_plan.hpp), then, per selected union vertex, the
// lanes each CSR / CSC tile run admits (the vertex range's lanes that also
// selected the run's remote tile).  Retained masks keep one mask per run and
// reuse the output's storage across calls: the previous batch's nonzero
// entries (its ledger) are cleared, not the whole image.  Descriptors list
// the nonzero runs with per-union-vertex offsets.  device_materialized
// builds no host run image.  Throws std::invalid_argument for inputs that do
// not describe one batch of those queries, for a representation change on a
// reused description, or when the admitted run edges disagree with the
// plan's union / per-lane edge estimates.
void prepare_batch_device_description(const WeightedGraph& graph, const TileRunLayout64& tile_runs,
                                      std::span<const RouteQuery> queries,
                                      std::span<const BatchQueryFeatures> features,
                                      const BatchPlanEntry& batch,
                                      BatchRunRepresentation run_representation,
                                      BatchDeviceDescription& output);

// Builds compact_vertex_biases_by_tile for the prepared batch (reference:
// src/batch_layout.cpp:1298-1390): union tiles packed in order, one bias per
// tile; entries of the previous ledger are cleared, a first build
// initialises the whole table.
void prepare_compact_vertex_mapping(const WeightedGraph& graph, const BatchPlanEntry& batch,
                                    BatchDeviceDescription& output);

// Deep check of the compact mapping: exactly the union tiles carry a bias,
// in packed order, and every other entry is zero.
[[nodiscard]] BatchLayoutValidationResult validate_compact_vertex_mapping(
    const WeightedGraph& graph, const BatchPlanEntry& batch,
    const BatchDeviceDescription& description) noexcept;

}  // namespace bfnew
