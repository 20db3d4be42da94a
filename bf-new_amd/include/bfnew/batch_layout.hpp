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

}  // namespace bfnew
