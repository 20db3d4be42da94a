// batch_plan.hpp -- greedy overlapping planner of <= 32-lane query batches
// (reference: include/bfnew/batch_plan.hpp:12-163, src/batch_plan.cpp).
//
// The relaxation engines of this library batch any number of queries in one
// launch; this planner groups queries into the reference's fixed-width lane
// batches for callers that use the lane-batched C++ engines
// (BatchDeviceDescription), with the reference's selection rule so that the
// same queries give the same batches, lane order and serialised plan.
//
// Rule: anchors in descending (selected edges, selected vertices, selected
// tiles), then ascending (expansion generation, query id).  A batch starts
// from the first unassigned anchor and repeatedly admits the unassigned query
// that ranks first by: larger tile Jaccard with the batch union, larger
// terminal-tile overlap, smaller projected union inflation, then smaller
// (projected union edges, projected union vertices, new tiles, |generation
// difference|, |source count difference|, |target count difference|, query
// id) -- among queries whose Jaccard is at least the policy minimum and whose
// projected inflation is at most the policy maximum -- until the batch is
// full or no query qualifies.
#pragma once

#include "bfnew/batch_layout.hpp"
#include "bfnew/device_layout.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <vector>

namespace bfnew {

// Vertex and induced-edge counts of tile sets, from per-tile counts and the
// tile-to-tile edge counts of the CSR tile runs (no per-vertex scan).
class SelectedRegionIndex {
 public:
  SelectedRegionIndex(const WeightedGraph& graph, const TileRunLayout64& tile_runs);

  [[nodiscard]] std::uint32_t tile_count() const noexcept {
    return static_cast<std::uint32_t>(vertex_counts_.size());
  }
  // True for the graph the index was built from and its copies (which share
  // its memo and hold the same data).
  [[nodiscard]] bool matches_graph(const WeightedGraph& graph) const noexcept {
    const auto other = graph.memo_handle();
    return !graph_memo_.owner_before(other) && !other.owner_before(graph_memo_);
  }
  // Throws std::out_of_range for a tile outside the graph; tiles must be
  // sorted and unique.
  [[nodiscard]] std::uint64_t selected_vertex_count(std::span<const TileId> tiles) const;
  [[nodiscard]] std::uint64_t selected_edge_count(std::span<const TileId> tiles) const;

  // Per-tile counts for incremental region arithmetic (tile < tile_count()).
  struct TilePairs {
    std::span<const std::uint32_t> tiles;  // the other tile, ascending
    std::span<const std::uint64_t> edges;  // edges between the two
  };
  [[nodiscard]] std::uint64_t tile_vertex_count(std::uint32_t tile) const noexcept {
    return vertex_counts_[tile];
  }
  [[nodiscard]] TilePairs pairs_from(std::uint32_t tile) const noexcept;
  [[nodiscard]] TilePairs pairs_into(std::uint32_t tile) const noexcept;

 private:
  std::weak_ptr<const WeightedGraphMemo> graph_memo_;
  std::vector<std::uint64_t> vertex_counts_;
  std::vector<std::uint64_t> pair_offsets_;   // [tile_count + 1]
  std::vector<std::uint32_t> pair_tiles_;     // destination tile, ascending per source tile
  std::vector<std::uint64_t> pair_edges_;     // edges from the source tile into it
  std::vector<std::uint64_t> in_pair_offsets_;  // the same pairs grouped by destination tile
  std::vector<std::uint32_t> in_pair_tiles_;    // source tile, ascending per destination tile
  std::vector<std::uint64_t> in_pair_edges_;
};

// The index of the graph and its own tile runs, built once per graph and
// kept in its memo.
[[nodiscard]] const SelectedRegionIndex& graph_region_index(const WeightedGraph& graph);

inline constexpr std::uint32_t invalid_batch_query_index = 0xFFFFFFFFu;
inline constexpr QueryId invalid_batch_query_id{0xFFFFFFFFu};

struct BatchFraction {
  std::uint64_t numerator{};
  std::uint64_t denominator{1U};
  constexpr bool operator==(const BatchFraction&) const noexcept = default;
};

struct BatchPlannerPolicy {
  std::uint32_t lane_width{32U};
  std::uint32_t minimum_jaccard_numerator{1U};
  std::uint32_t minimum_jaccard_denominator{8U};
  std::uint32_t maximum_union_inflation_numerator{2U};
  std::uint32_t maximum_union_inflation_denominator{1U};
  constexpr bool operator==(const BatchPlannerPolicy&) const noexcept = default;
};

enum class BatchPlannerPolicyError : std::uint8_t {
  none,
  unsupported_lane_width,
  invalid_jaccard_fraction,
  invalid_union_inflation_fraction,
};

[[nodiscard]] BatchPlannerPolicyError validate_batch_planner_policy(
    const BatchPlannerPolicy& policy) noexcept;

struct BatchQueryFeatures {
  QueryId query_id{};
  std::uint32_t expansion_generation{};
  std::uint32_t source_count{};
  std::uint32_t target_count{};
  std::vector<TileId> source_tiles;
  std::vector<TileId> target_tiles;
  std::vector<TileId> selected_tiles;
  std::uint64_t selected_vertex_count{};
  std::uint64_t selected_edge_estimate{};
  bool operator==(const BatchQueryFeatures&) const = default;
};

// Features of valid route queries, sorted by query id (ids must be unique).
[[nodiscard]] std::vector<BatchQueryFeatures> make_batch_query_features(
    const WeightedGraph& graph, const SelectedRegionIndex& selected_regions,
    std::span<const RouteQuery> queries);

// |A n B| / |A u B| of sorted unique tile sets (1/1 for two empty sets).
[[nodiscard]] BatchFraction tile_set_jaccard(std::span<const TileId> left,
                                             std::span<const TileId> right);

// Union tile/lane positions over the lanes' own selected tile positions.
[[nodiscard]] BatchFraction batch_union_tile_inflation(
    std::span<const BatchQueryFeatures> queries, std::span<const std::uint32_t> query_indices,
    std::span<const TileId> union_tiles);

struct BatchPlanEntry {
  std::uint32_t lane_width{};
  LaneMask valid_lane_mask{};
  std::vector<std::uint32_t> query_indices_by_lane;   // into the features span
  std::vector<QueryId> query_ids_by_lane;
  std::vector<std::uint32_t> expansion_generations_by_lane;
  std::vector<TileId> union_tiles;
  std::vector<LaneMask> union_tile_lane_masks;
  std::uint64_t union_vertex_count{};
  std::uint64_t union_edge_estimate{};
  std::uint64_t selected_lane_vertex_count{};
  std::uint64_t selected_lane_edge_estimate{};
  bool operator==(const BatchPlanEntry&) const = default;
};

struct BatchPlan {
  BatchPlannerPolicy policy{};
  std::uint32_t input_query_count{};
  std::vector<BatchPlanEntry> batches;
  bool operator==(const BatchPlan&) const = default;
};

// Widths 32, 16 and 8 of the same policy.
struct BatchPlanFamily {
  std::array<BatchPlan, 3U> plans;
  bool operator==(const BatchPlanFamily&) const = default;
};

// Throw std::invalid_argument for an invalid policy or invalid features.
[[nodiscard]] BatchPlan make_overlapping_batch_plan(const SelectedRegionIndex& selected_regions,
                                                    std::span<const BatchQueryFeatures> queries,
                                                    const BatchPlannerPolicy& policy);
[[nodiscard]] BatchPlanFamily make_standard_batch_plan_family(
    const SelectedRegionIndex& selected_regions, std::span<const BatchQueryFeatures> queries,
    BatchPlannerPolicy policy = {});

enum class BatchPlanValidationErrorCode : std::uint8_t {
  none,
  invalid_policy,
  query_count_overflow,
  empty_query_set,
  duplicate_query_id,
  invalid_query_feature,
  empty_plan,
  invalid_batch_width,
  invalid_batch_shape,
  invalid_valid_lane_mask,
  invalid_padding,
  query_index_out_of_range,
  query_identity_mismatch,
  duplicate_assignment,
  omitted_query,
  union_tile_mismatch,
  tile_lane_mask_mismatch,
  estimate_mismatch,
};

struct BatchPlanValidationResult {
  static constexpr std::size_t no_position = static_cast<std::size_t>(-1);
  BatchPlanValidationErrorCode code{BatchPlanValidationErrorCode::none};
  std::size_t batch{no_position};
  std::size_t lane{no_position};
  [[nodiscard]] constexpr bool ok() const noexcept {
    return code == BatchPlanValidationErrorCode::none;
  }
};

// Structural checks, then equality with the plan the rule builds.
[[nodiscard]] BatchPlanValidationResult validate_batch_plan(
    const SelectedRegionIndex& selected_regions, std::span<const BatchQueryFeatures> queries,
    const BatchPlan& plan) noexcept;

// The bfnew.batch-plan.v1 TSV image and its 64-bit FNV-1a fingerprint.
[[nodiscard]] std::string serialize_batch_plan_tsv(const BatchPlan& plan);
[[nodiscard]] std::uint64_t fingerprint_batch_plan(const BatchPlan& plan);

// Lays out one planned batch for the lane-batched engines: lane l carries
// queries[batch.query_indices_by_lane[l]] (queries and features index alike,
// as make_batch_query_features orders them).  Throws std::invalid_argument
// when the entry does not describe those queries
// (reference: include/bfnew/batch_layout.hpp:128-135).
void prepare_batch_device_description(const WeightedGraph& graph,
                                      std::span<const RouteQuery> queries,
                                      std::span<const BatchQueryFeatures> features,
                                      const BatchPlanEntry& batch,
                                      BatchDeviceDescription& output);

}  // namespace bfnew
