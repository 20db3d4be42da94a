// graph.hpp -- input and weighted graphs (reference: include/bfnew/graph.hpp:18-267).
#pragma once

#include "bfnew/spatial.hpp"
#include "bfnew/types.hpp"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <mutex>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace bfnew {

struct VertexMetadata {
  std::int32_t x{};
  std::int32_t y{};
  bool has_location{};
  ResourceClassId resource_class{};
  [[nodiscard]] static constexpr VertexMetadata located(std::int32_t x, std::int32_t y,
                                                        ResourceClassId c) noexcept {
    return VertexMetadata{x, y, true, c};
  }
  [[nodiscard]] static constexpr VertexMetadata unlocated(ResourceClassId c) noexcept {
    return VertexMetadata{0, 0, false, c};
  }
  constexpr bool operator==(const VertexMetadata&) const noexcept = default;
};

struct EdgeInputRecord {
  VertexId source{};
  VertexId destination{};
  float weight{};
  PhysicalProvenance provenance{};
  constexpr bool operator==(const EdgeInputRecord&) const noexcept = default;
};

struct OutgoingCsrView {
  std::span<const EdgeOffset> row_offsets;
  std::span<const VertexId> destinations;
  std::span<const float> weights;
  std::span<const EdgeId> edge_ids;
};

struct IncomingCscView {
  std::span<const EdgeOffset> column_offsets;
  std::span<const VertexId> sources;
  std::span<const float> weights;
  std::span<const EdgeId> edge_ids;
};

// -0 is stored as +0 (graph_builder.cpp:91-93).
[[nodiscard]] float canonicalize_weight(float weight) noexcept;
[[nodiscard]] bool is_valid_weight(float weight) noexcept;
[[nodiscard]] std::uint32_t canonical_weight_bits(float weight) noexcept;
[[nodiscard]] bool is_valid_vertex_id(VertexId id, std::uint32_t vertex_count) noexcept;

// Canonicalizes weights and validates endpoints/weights on construction;
// throws std::invalid_argument / std::out_of_range like the reference.
class InputGraph {
 public:
  InputGraph(std::vector<VertexMetadata> vertices, std::vector<EdgeInputRecord> edges);
  [[nodiscard]] std::uint32_t vertex_count() const noexcept { return vertex_count_; }
  [[nodiscard]] std::uint64_t edge_count() const noexcept { return edges_.size(); }
  [[nodiscard]] std::span<const VertexMetadata> vertices() const noexcept { return vertices_; }
  [[nodiscard]] std::span<const EdgeInputRecord> edges() const noexcept { return edges_; }

 private:
  std::uint32_t vertex_count_{};
  std::vector<VertexMetadata> vertices_;
  std::vector<EdgeInputRecord> edges_;
};

struct SyntheticTileGridConfig;
class SyntheticTileGraph;

// Values derived from an immutable WeightedGraph, computed at most once per
// graph and shared by its copies (which hold the same data).
struct WeightedGraphMemo {
  std::once_flag device_fingerprint_once;
  std::uint64_t device_fingerprint[2]{};
  std::once_flag validation_once;
  std::string validation_message;  // empty when the graph is valid
  // Derived layouts built once per graph by the libraries that own their
  // types (tile runs, selected-region index); typed where they are used.
  std::once_flag tile_runs_once;
  std::shared_ptr<const void> tile_runs;
  std::once_flag region_index_once;
  std::shared_ptr<const void> region_index;
};

class WeightedGraph {
 public:
  [[nodiscard]] std::uint32_t vertex_count() const noexcept { return vertex_count_; }
  [[nodiscard]] EdgeCount edge_count() const noexcept { return edge_count_; }
  [[nodiscard]] std::span<const VertexMetadata> vertices() const noexcept { return vertices_; }
  [[nodiscard]] std::span<const PhysicalProvenance> edge_provenance() const noexcept {
    return edge_provenance_;
  }
  [[nodiscard]] OutgoingCsrView outgoing() const noexcept {
    return {out_offsets_, out_destinations_, out_weights_, out_edge_ids_};
  }
  [[nodiscard]] IncomingCscView incoming() const noexcept {
    return {in_offsets_, in_sources_, in_weights_, in_edge_ids_};
  }
  [[nodiscard]] std::span<const VertexId> old_to_new() const noexcept { return old_to_new_; }
  [[nodiscard]] std::span<const VertexId> new_to_old() const noexcept { return new_to_old_; }
  [[nodiscard]] std::span<const VertexId> original_vertex_ids() const noexcept { return new_to_old_; }
  [[nodiscard]] bool has_spatial_ordering() const noexcept { return spatial_; }
  [[nodiscard]] std::span<const TileId> owner_tiles() const noexcept { return owner_tiles_; }
  [[nodiscard]] std::span<const TileCoordinate> tile_coordinates() const noexcept { return tile_coordinates_; }
  [[nodiscard]] std::span<const EdgeOffset> tile_vertex_offsets() const noexcept { return tile_vertex_offsets_; }
  [[nodiscard]] const SpatialOrderConfig& spatial_order_config() const noexcept { return config_; }
  // Throws std::logic_error on a moved-from graph.
  [[nodiscard]] WeightedGraphMemo& memo() const {
    if (!memo_) throw std::logic_error{"WeightedGraph: moved-from graph"};
    return *memo_;
  }
  [[nodiscard]] std::weak_ptr<const WeightedGraphMemo> memo_handle() const noexcept { return memo_; }

 private:
  WeightedGraph() = default;
  std::uint32_t vertex_count_{};
  EdgeCount edge_count_{};
  std::vector<VertexMetadata> vertices_;
  std::vector<PhysicalProvenance> edge_provenance_;
  std::vector<EdgeOffset> out_offsets_;
  std::vector<VertexId> out_destinations_;
  std::vector<float> out_weights_;
  std::vector<EdgeId> out_edge_ids_;
  std::vector<EdgeOffset> in_offsets_;
  std::vector<VertexId> in_sources_;
  std::vector<float> in_weights_;
  std::vector<EdgeId> in_edge_ids_;
  std::vector<VertexId> old_to_new_;
  std::vector<VertexId> new_to_old_;
  bool spatial_{};
  SpatialOrderConfig config_{};
  std::vector<TileId> owner_tiles_;
  std::vector<TileCoordinate> tile_coordinates_;
  std::vector<EdgeOffset> tile_vertex_offsets_;
  std::shared_ptr<WeightedGraphMemo> memo_ = std::make_shared<WeightedGraphMemo>();

  friend WeightedGraph build_weighted_graph(const InputGraph& input);
  friend WeightedGraph build_spatially_ordered_graph(const InputGraph& input,
                                                     const SpatialOrderConfig& config,
                                                     const LocalityKeyPolicy& locality_policy);
  friend class SyntheticTileGraph;
};

struct PartitionedGraph {
  WeightedGraph graph;
  TileDirectory tiles;
};

// Canonical EdgeIds are ranks in (source, destination, weight bits,
// provenance) order; CSR rows are ordered by (destination, EdgeId) and CSC
// columns by (source, EdgeId) (graph_builder.cpp:182-283).
[[nodiscard]] WeightedGraph build_weighted_graph(const InputGraph& input);

// Structural checks of a built graph: offsets, ranges, canonical orders,
// transpose agreement, permutation round trip and tile spans.  Returns an
// empty string when valid, else a description of the first violation.
[[nodiscard]] const char* validate_weighted_graph_message(const WeightedGraph& graph);

// validate_weighted_graph_message, checked once per graph (graphs are
// immutable) and kept in its memo.
[[nodiscard]] const std::string& weighted_graph_validation(const WeightedGraph& graph);

}  // namespace bfnew
