// spatial.hpp -- spatial order, tiles and the tile directory
// (reference: include/bfnew/spatial.hpp:19-193).
#pragma once

#include "bfnew/types.hpp"

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <span>
#include <vector>

namespace bfnew {

class InputGraph;
class WeightedGraph;
struct PartitionedGraph;
struct WeightedGraphMemo;

struct SpatialOrderConfig {
  std::int32_t origin_x{};
  std::int32_t origin_y{};
  std::uint32_t tile_width{};
  std::uint32_t tile_height{};
  constexpr bool operator==(const SpatialOrderConfig&) const noexcept = default;
};

struct TileCoordinate {
  std::int64_t tile_x{};
  std::int64_t tile_y{};
  bool has_location{};
  [[nodiscard]] static constexpr TileCoordinate located(std::int64_t x, std::int64_t y) noexcept {
    return TileCoordinate{x, y, true};
  }
  [[nodiscard]] static constexpr TileCoordinate spill() noexcept { return TileCoordinate{0, 0, false}; }
  constexpr bool operator==(const TileCoordinate&) const noexcept = default;
};

using LocalityKeyPolicy = std::function<std::uint64_t(std::uint32_t local_x, std::uint32_t local_y)>;

// Bit-interleaved (Z-order) key of the local coordinate: x on even bits.
struct MortonLocalityPolicy {
  [[nodiscard]] std::uint64_t operator()(std::uint32_t local_x, std::uint32_t local_y) const noexcept;
};

[[nodiscard]] WeightedGraph build_spatially_ordered_graph(
    const InputGraph& input, const SpatialOrderConfig& config,
    const LocalityKeyPolicy& locality_policy = MortonLocalityPolicy{});

// Per-tile neighbourhood (reference: tiled_graph.cpp:214-291): Chebyshev-1
// neighbour tiles, the tile's internal edges, outgoing/incoming cross edges
// (by stable EdgeId) and the halo of remote source vertices.
class TileDirectory {
 public:
  [[nodiscard]] std::size_t tile_count() const noexcept {
    return neighbor_tile_offsets_.empty() ? 0U : neighbor_tile_offsets_.size() - 1U;
  }
  [[nodiscard]] TileId spill_tile() const noexcept { return spill_tile_; }
  [[nodiscard]] std::span<const EdgeOffset> neighbor_tile_offsets() const noexcept { return neighbor_tile_offsets_; }
  [[nodiscard]] std::span<const TileId> neighbor_tiles() const noexcept { return neighbor_tiles_; }
  [[nodiscard]] std::span<const EdgeOffset> internal_edge_offsets() const noexcept { return internal_edge_offsets_; }
  [[nodiscard]] std::span<const EdgeId> internal_edge_ids() const noexcept { return internal_edge_ids_; }
  [[nodiscard]] std::span<const EdgeOffset> outgoing_cross_edge_offsets() const noexcept { return outgoing_cross_edge_offsets_; }
  [[nodiscard]] std::span<const EdgeId> outgoing_cross_edge_ids() const noexcept { return outgoing_cross_edge_ids_; }
  [[nodiscard]] std::span<const EdgeOffset> incoming_cross_edge_offsets() const noexcept { return incoming_cross_edge_offsets_; }
  [[nodiscard]] std::span<const EdgeId> incoming_cross_edge_ids() const noexcept { return incoming_cross_edge_ids_; }
  [[nodiscard]] std::span<const EdgeOffset> halo_vertex_offsets() const noexcept { return halo_vertex_offsets_; }
  [[nodiscard]] std::span<const VertexId> halo_vertices() const noexcept { return halo_vertices_; }
  // True when this directory was built from `graph` or from a copy of it
  // (graphs are immutable; copies share one memo).
  [[nodiscard]] bool built_from(const WeightedGraph& graph) const noexcept;

 private:
  std::weak_ptr<const WeightedGraphMemo> built_from_;
  TileId spill_tile_{};
  std::vector<EdgeOffset> neighbor_tile_offsets_;
  std::vector<TileId> neighbor_tiles_;
  std::vector<EdgeOffset> internal_edge_offsets_;
  std::vector<EdgeId> internal_edge_ids_;
  std::vector<EdgeOffset> outgoing_cross_edge_offsets_;
  std::vector<EdgeId> outgoing_cross_edge_ids_;
  std::vector<EdgeOffset> incoming_cross_edge_offsets_;
  std::vector<EdgeId> incoming_cross_edge_ids_;
  std::vector<EdgeOffset> halo_vertex_offsets_;
  std::vector<VertexId> halo_vertices_;
  friend TileDirectory build_tile_directory(const WeightedGraph& graph);
};

[[nodiscard]] TileDirectory build_tile_directory(const WeightedGraph& graph);

class SpatialPartitioner {
 public:
  virtual ~SpatialPartitioner() = default;
  [[nodiscard]] virtual PartitionedGraph partition(const InputGraph& input) const = 0;
};

class UniformGridPartitioner final : public SpatialPartitioner {
 public:
  explicit UniformGridPartitioner(SpatialOrderConfig config,
                                  LocalityKeyPolicy locality_policy = MortonLocalityPolicy{});
  [[nodiscard]] const SpatialOrderConfig& config() const noexcept { return config_; }
  [[nodiscard]] PartitionedGraph partition(const InputGraph& input) const override;

 private:
  SpatialOrderConfig config_;
  LocalityKeyPolicy locality_policy_;
};

}  // namespace bfnew
