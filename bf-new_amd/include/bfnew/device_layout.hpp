// device_layout.hpp -- tile runs and the 32-bit device image
// (reference: include/bfnew/device_layout.hpp:15-233).
#pragma once

#include "bfnew/graph.hpp"

#include <cstdint>
#include <span>
#include <vector>

namespace bfnew {

// Maximal runs of one remote tile inside each CSR row / CSC column; a run
// never crosses a row or column boundary.
struct TileRunLayout64 {
  std::vector<EdgeOffset> csr_row_run_offsets;
  std::vector<EdgeOffset> csr_run_edge_offsets;
  std::vector<TileId> csr_run_destination_tiles;
  std::vector<EdgeOffset> csc_column_run_offsets;
  std::vector<EdgeOffset> csc_run_edge_offsets;
  std::vector<TileId> csc_run_source_tiles;
  bool operator==(const TileRunLayout64&) const = default;
};

[[nodiscard]] TileRunLayout64 build_tile_run_layout(const WeightedGraph& graph);

// build_tile_run_layout once per graph, kept in its memo.
[[nodiscard]] const TileRunLayout64& graph_tile_runs(const WeightedGraph& graph);

// Host staging image of the immutable device graph.
struct DeviceGraphLayout32 {
  std::uint32_t vertex_count{};
  std::uint32_t edge_count{};
  std::uint32_t tile_count{};
  std::vector<std::uint32_t> owner_tiles;
  std::vector<std::uint32_t> csr_row_offsets;
  std::vector<std::uint32_t> csr_destinations;
  std::vector<float> csr_weights;
  std::vector<std::uint32_t> csr_row_run_offsets;
  std::vector<std::uint32_t> csr_run_edge_offsets;
  std::vector<std::uint32_t> csr_run_destination_tiles;
  std::vector<std::uint32_t> csc_column_offsets;
  std::vector<std::uint32_t> csc_sources;
  std::vector<float> csc_weights;
  std::vector<std::uint32_t> csc_edge_ids;
  std::vector<std::uint32_t> csc_column_run_offsets;
  std::vector<std::uint32_t> csc_run_edge_offsets;
  std::vector<std::uint32_t> csc_run_source_tiles;
};

struct DeviceGraphMemoryReport {
  std::uint64_t owner_tiles_bytes{};
  std::uint64_t csr_row_offsets_bytes{};
  std::uint64_t csr_destinations_bytes{};
  std::uint64_t csr_weights_bytes{};
  std::uint64_t csr_row_run_offsets_bytes{};
  std::uint64_t csr_run_edge_offsets_bytes{};
  std::uint64_t csr_run_destination_tiles_bytes{};
  std::uint64_t csc_column_offsets_bytes{};
  std::uint64_t csc_sources_bytes{};
  std::uint64_t csc_weights_bytes{};
  std::uint64_t csc_edge_ids_bytes{};
  std::uint64_t csc_column_run_offsets_bytes{};
  std::uint64_t csc_run_edge_offsets_bytes{};
  std::uint64_t csc_run_source_tiles_bytes{};
  std::uint64_t total_bytes{};
  constexpr bool operator==(const DeviceGraphMemoryReport&) const noexcept = default;
};

// Content identity of the image (identity guard, not a cryptographic hash).
struct DeviceGraphFingerprint {
  std::uint64_t first{};
  std::uint64_t second{};
  constexpr bool operator==(const DeviceGraphFingerprint&) const noexcept = default;
};

[[nodiscard]] std::uint32_t checked_device_offset32(EdgeOffset offset);
[[nodiscard]] DeviceGraphLayout32 build_device_graph_layout32(const WeightedGraph& graph,
                                                              const TileRunLayout64& tile_runs);
[[nodiscard]] bool device_graph_layouts_deep_equal(const DeviceGraphLayout32& left,
                                                   const DeviceGraphLayout32& right);
[[nodiscard]] DeviceGraphMemoryReport report_device_graph_memory(
    const DeviceGraphLayout32& layout) noexcept;
[[nodiscard]] DeviceGraphFingerprint fingerprint_device_graph_layout32(
    const DeviceGraphLayout32& layout) noexcept;

// fingerprint_device_graph_layout32 of the layout built from the graph and
// its tile runs (reference: src/device_layout.cpp:711-760).  Throws
// std::invalid_argument when tile_runs are not the graph's.  The value is
// kept with the graph (WeightedGraph::memo), so a graph pays for it once;
// passing graph_tile_runs(graph) also skips the element-wise comparison.
[[nodiscard]] DeviceGraphFingerprint fingerprint_device_graph_source32(const WeightedGraph& graph,
                                                                       const TileRunLayout64& tile_runs);

struct TileRunLaneMasks {
  std::vector<std::uint32_t> csr_run_masks;
  std::vector<std::uint32_t> csc_run_masks;
  bool operator==(const TileRunLaneMasks&) const = default;
};

// run mask = lane mask of the run's own tile & lane mask of its remote tile.
void compute_tile_run_lane_masks(const WeightedGraph& graph, const TileRunLayout64& tile_runs,
                                 std::span<const std::uint32_t> tile_lane_masks,
                                 TileRunLaneMasks& output);

}  // namespace bfnew
