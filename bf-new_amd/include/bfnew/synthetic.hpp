// synthetic.hpp -- the synthetic tile-grid workloads (no reference
// counterpart: the reference runs FPGA graphs read from artifacts).
//
// Graph: tiles_x * tiles_y tiles of nodes_per_tile nodes, numbered tile-major
// (tile = ty * tiles_x + tx), every node located at its tile coordinate with
// resource class 1, so the reference reorder is the identity and one empty
// spill tile follows the grid.  Per node: intra_edges edges to uniform other
// nodes of the same tile, then one edge per offset {+1,-1,+2,-2,+6,-6} along a
// random axis to a uniform node of the clamped destination tile.  Every edge
// into v weighs cost[v], cost uniform in [cost_min, cost_max] from
//   cost = cost_min + (float)(r >> 40) * ((cost_max - cost_min) / 2^24)
// evaluated in fp32 without contraction.  Provenance record = 10 * v + j.
//
// Randomness: SplitMix64.  Node v consumes stream positions 64 v + 1 ...;
// query q of a query set consumes positions 64 q + 1 ... of a stream seeded
// with the set's seed.  Identical on every host.
#pragma once

#include "bfnew/device_layout.hpp"
#include "bfnew/graph.hpp"

#include <cstdint>
#include <span>
#include <vector>

namespace bfnew {

struct SyntheticTileGridConfig {
  std::uint32_t tiles_x{16};
  std::uint32_t tiles_y{16};
  std::uint32_t nodes_per_tile{64};
  std::uint32_t intra_edges{4};
  std::uint64_t seed{1};
  float cost_min{1.0F};
  float cost_max{4.0F};
};

// SplitMix64 output at stream position `index` of the stream seeded `seed`.
[[nodiscard]] std::uint64_t splitmix64_at(std::uint64_t seed, std::uint64_t index) noexcept;

class SyntheticTileGraph {
 public:
  // threads == 0 uses the hardware concurrency.
  [[nodiscard]] static SyntheticTileGraph generate(const SyntheticTileGridConfig& config,
                                                   unsigned threads = 0);

  [[nodiscard]] const SyntheticTileGridConfig& config() const noexcept { return config_; }
  [[nodiscard]] const DeviceGraphLayout32& layout() const noexcept { return layout_; }
  [[nodiscard]] std::span<const std::uint32_t> tile_vertex_offsets() const noexcept { return tile_offsets_; }
  [[nodiscard]] std::span<const float> node_costs() const noexcept { return costs_; }
  [[nodiscard]] std::uint32_t tile_count() const noexcept { return layout_.tile_count; }

  // The same graph as reference input records (for builder cross-checks).
  [[nodiscard]] InputGraph input_graph() const;
  // The same graph as the spatially ordered WeightedGraph the reference
  // builder produces for it, built directly from the layout.
  [[nodiscard]] WeightedGraph weighted_graph() const;

 private:
  SyntheticTileGridConfig config_{};
  DeviceGraphLayout32 layout_{};
  std::vector<std::uint32_t> tile_offsets_;
  std::vector<float> costs_;
};

struct SyntheticQuery {
  std::uint32_t query_id{};
  std::uint32_t source{};
  std::uint32_t target{};
  std::uint32_t box_x{};
  std::uint32_t box_y{};
  std::uint32_t box_side{};
  bool unbounded{};
};

enum class SyntheticBoxMix : std::uint32_t {
  log_uniform = 0,   // side log-uniform in [side_min, side_max]
  serving = 1,       // 70% 4-8, 25% 16-32, 5% 64-or-unbounded
};

struct SyntheticQueryConfig {
  std::uint32_t count{64};
  std::uint64_t seed{2};
  std::uint32_t side_min{4};
  std::uint32_t side_max{32};
  double unbounded_fraction{0.0};
  SyntheticBoxMix mix{SyntheticBoxMix::log_uniform};
};

[[nodiscard]] std::vector<SyntheticQuery> generate_synthetic_queries(
    const SyntheticTileGridConfig& grid, const SyntheticQueryConfig& config);

// Sorted selected tiles: the box's tiles, or every tile (spill included) for
// an unbounded query.
[[nodiscard]] std::vector<TileId> synthetic_query_tiles(const SyntheticTileGridConfig& grid,
                                                        const SyntheticQuery& query);

}  // namespace bfnew
