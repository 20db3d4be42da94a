// batch_plan.cpp -- greedy overlapping batch planner
// (reference rule: src/batch_plan.cpp:271-543; TSV image :870-929).
#include "bfnew/batch_plan.hpp"

#include "bfnew/device_layout.hpp"

#include <algorithm>
#include <bit>
#include <iterator>
#include <stdexcept>
#include <string>
#include <tuple>

namespace bfnew {
namespace {

using Tiles = std::span<const TileId>;

bool supported_width(std::uint32_t w) { return w == 1U || w == 8U || w == 16U || w == 32U; }

bool strictly_increasing(Tiles tiles) {
  return std::adjacent_find(tiles.begin(), tiles.end(),
                            [](TileId a, TileId b) { return !(a < b); }) == tiles.end();
}

bool canonical(Tiles tiles, std::uint32_t tile_count) {
  return strictly_increasing(tiles) && (tiles.empty() || tiles.back().value() < tile_count);
}

std::uint64_t common_count(Tiles a, Tiles b) {
  std::uint64_t n = 0;
  for (std::size_t i = 0, j = 0; i < a.size() && j < b.size();) {
    if (a[i] < b[j]) {
      ++i;
    } else if (b[j] < a[i]) {
      ++j;
    } else {
      ++n, ++i, ++j;
    }
  }
  return n;
}

std::vector<TileId> merged(Tiles a, Tiles b) {
  std::vector<TileId> out;
  out.reserve(a.size() + b.size());
  std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
  return out;
}

// Exact order of two non-negative fractions (denominators > 0).
int order(BatchFraction x, BatchFraction y) {
  const unsigned __int128 l = static_cast<unsigned __int128>(x.numerator) * y.denominator;
  const unsigned __int128 r = static_cast<unsigned __int128>(y.numerator) * x.denominator;
  return l < r ? -1 : (l > r ? 1 : 0);
}

LaneMask low_lanes(std::uint32_t n) { return n >= 32U ? ~LaneMask{0} : (LaneMask{1} << n) - 1U; }

std::uint32_t gap(std::uint32_t a, std::uint32_t b) { return a > b ? a - b : b - a; }

std::vector<TileId> owner_tiles_of(const WeightedGraph& graph, std::span<const VertexId> vertices) {
  std::vector<TileId> out;
  for (VertexId v : vertices) out.push_back(graph.owner_tiles()[v.value()]);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

bool features_valid(const SelectedRegionIndex& index, const BatchQueryFeatures& f) noexcept {
  try {
    const std::uint32_t T = index.tile_count();
    const bool shape = f.source_count > 0U && f.target_count > 0U && !f.source_tiles.empty() &&
                       !f.target_tiles.empty() && !f.selected_tiles.empty() &&
                       f.source_tiles.size() <= f.source_count &&
                       f.target_tiles.size() <= f.target_count &&
                       canonical(f.source_tiles, T) && canonical(f.target_tiles, T) &&
                       canonical(f.selected_tiles, T) &&
                       std::includes(f.selected_tiles.begin(), f.selected_tiles.end(),
                                     f.source_tiles.begin(), f.source_tiles.end()) &&
                       std::includes(f.selected_tiles.begin(), f.selected_tiles.end(),
                                     f.target_tiles.begin(), f.target_tiles.end());
    if (!shape) return false;
    const std::uint64_t v = index.selected_vertex_count(f.selected_tiles);
    return v != 0U && v == f.selected_vertex_count &&
           index.selected_edge_count(f.selected_tiles) == f.selected_edge_estimate;
  } catch (...) {
    return false;
  }
}

bool ids_unique(std::span<const BatchQueryFeatures> queries) {
  std::vector<std::uint32_t> ids;
  for (const auto& q : queries) ids.push_back(q.query_id.value());
  std::sort(ids.begin(), ids.end());
  return std::adjacent_find(ids.begin(), ids.end()) == ids.end();
}

void require_plannable(const SelectedRegionIndex& index, std::span<const BatchQueryFeatures> queries) {
  if (queries.empty()) throw std::invalid_argument{"batch planning needs at least one query"};
  if (queries.size() > 0xFFFFFFFFull) throw std::invalid_argument{"too many queries"};
  for (const auto& q : queries)
    if (!features_valid(index, q)) throw std::invalid_argument{"batch planner query features are invalid"};
  if (!ids_unique(queries)) throw std::invalid_argument{"batch planner query ids must be unique"};
}

// The open batch: lanes, union tiles (sorted, and as a membership map), the
// lanes' terminal tiles, and the union's running vertex and edge counts.
struct OpenBatch {
  std::vector<std::uint32_t> lanes;
  std::vector<TileId> tiles, source_tiles, target_tiles;
  std::uint64_t lane_tiles = 0;  // selected tile positions of the lanes
  std::uint64_t vertices = 0, edges = 0;
};

// One query's offer to join the open batch, ranked by (larger Jaccard, more
// shared terminal tiles, smaller inflation, fewer union edges, fewer union
// vertices, fewer new tiles, closer generation / source / target counts to
// the anchor, smaller id).  The union's edge and vertex counts are costly, so
// they are computed only when a comparison reaches them.
struct Offer {
  std::uint32_t query = 0;
  BatchFraction jaccard{};
  std::uint64_t shared_terminals = 0;
  BatchFraction inflation{};
  std::uint64_t new_tiles = 0;
  std::uint32_t generation_gap = 0, source_gap = 0, target_gap = 0, id = 0;
  bool counted = false;
  std::uint64_t union_edges = 0, union_vertices = 0;
};

// Tile-set arithmetic against the open batch: `in_union` marks the union's
// tiles and `common[q]` holds |union n selected(q)| for every query.
class Planner {
 public:
  Planner(const SelectedRegionIndex& index, std::span<const BatchQueryFeatures> queries)
      : index_{index}, queries_{queries}, in_union_(index.tile_count(), 0U), in_new_(index.tile_count(), 0U),
        common_(queries.size(), 0U), holders_offsets_(index.tile_count() + 1U, 0U) {
    for (const auto& f : queries)
      for (TileId t : f.selected_tiles) ++holders_offsets_[t.value() + 1U];
    for (std::size_t t = 0; t < index.tile_count(); ++t) holders_offsets_[t + 1] += holders_offsets_[t];
    holders_.resize(holders_offsets_.back());
    std::vector<std::uint64_t> fill(holders_offsets_.begin(), holders_offsets_.end() - 1);
    for (std::uint32_t q = 0; q < queries.size(); ++q)
      for (TileId t : queries[q].selected_tiles) holders_[fill[t.value()]++] = q;
  }

  void open(OpenBatch& b, std::uint32_t anchor) {
    const BatchQueryFeatures& f = queries_[anchor];
    b.lanes = {anchor};
    b.tiles = f.selected_tiles;
    b.source_tiles = f.source_tiles;
    b.target_tiles = f.target_tiles;
    b.lane_tiles = f.selected_tiles.size();
    b.vertices = f.selected_vertex_count;
    b.edges = f.selected_edge_estimate;
    for (TileId t : b.tiles) mark(t.value());
  }

  void close(const OpenBatch& b) {
    for (TileId t : b.tiles) {
      in_union_[t.value()] = 0U;
      for (auto i = holders_offsets_[t.value()]; i < holders_offsets_[t.value() + 1]; ++i) common_[holders_[i]] = 0U;
    }
  }

  Offer offer(const OpenBatch& b, std::uint32_t q, const BatchQueryFeatures& anchor) const {
    const BatchQueryFeatures& f = queries_[q];
    Offer o;
    o.query = q;
    const std::uint64_t both = common_[q];
    const std::uint64_t either = b.tiles.size() + f.selected_tiles.size() - both;
    o.jaccard = BatchFraction{both, either};
    o.shared_terminals = common_count(b.source_tiles, f.source_tiles) + common_count(b.target_tiles, f.target_tiles);
    o.inflation = BatchFraction{(b.lanes.size() + 1U) * either, b.lane_tiles + f.selected_tiles.size()};
    o.new_tiles = f.selected_tiles.size() - both;
    o.generation_gap = gap(f.expansion_generation, anchor.expansion_generation);
    o.source_gap = gap(f.source_count, anchor.source_count);
    o.target_gap = gap(f.target_count, anchor.target_count);
    o.id = f.query_id.value();
    return o;
  }

  // Union edges and vertices after admitting the offer's query: the union's
  // own plus those of its new tiles N (edges N->N, N->union, union->N).
  void count(const OpenBatch& b, Offer& o) {
    if (o.counted) return;
    const auto& sel = queries_[o.query].selected_tiles;
    std::uint64_t vertices = b.vertices, edges = b.edges;
    for (TileId t : sel)
      if (!in_union_[t.value()]) in_new_[t.value()] = 1U;
    for (TileId t : sel) {
      const std::uint32_t n = t.value();
      if (!in_new_[n]) continue;
      vertices += index_.tile_vertex_count(n);
      const auto out = index_.pairs_from(n);
      for (std::size_t i = 0; i < out.tiles.size(); ++i)
        if (in_union_[out.tiles[i]] || in_new_[out.tiles[i]]) edges += out.edges[i];
      const auto in = index_.pairs_into(n);
      for (std::size_t i = 0; i < in.tiles.size(); ++i)
        if (in_union_[in.tiles[i]]) edges += in.edges[i];
    }
    for (TileId t : sel) in_new_[t.value()] = 0U;
    o.union_vertices = vertices;
    o.union_edges = edges;
    o.counted = true;
  }

  bool before(const OpenBatch& b, Offer& x, Offer& y) {
    if (int c = order(x.jaccard, y.jaccard)) return c > 0;
    if (x.shared_terminals != y.shared_terminals) return x.shared_terminals > y.shared_terminals;
    if (int c = order(x.inflation, y.inflation)) return c < 0;
    count(b, x);
    count(b, y);
    return std::tie(x.union_edges, x.union_vertices, x.new_tiles, x.generation_gap, x.source_gap, x.target_gap,
                    x.id) < std::tie(y.union_edges, y.union_vertices, y.new_tiles, y.generation_gap,
                                     y.source_gap, y.target_gap, y.id);
  }

  void admit(OpenBatch& b, Offer& o) {
    count(b, o);
    const BatchQueryFeatures& f = queries_[o.query];
    for (TileId t : f.selected_tiles)
      if (!in_union_[t.value()]) mark(t.value());
    b.lanes.push_back(o.query);
    b.lane_tiles += f.selected_tiles.size();
    b.tiles = merged(b.tiles, f.selected_tiles);
    b.source_tiles = merged(b.source_tiles, f.source_tiles);
    b.target_tiles = merged(b.target_tiles, f.target_tiles);
    b.vertices = o.union_vertices;
    b.edges = o.union_edges;
  }

 private:
  void mark(std::uint32_t t) {
    in_union_[t] = 1U;
    for (auto i = holders_offsets_[t]; i < holders_offsets_[t + 1]; ++i) ++common_[holders_[i]];
  }

  const SelectedRegionIndex& index_;
  std::span<const BatchQueryFeatures> queries_;
  std::vector<std::uint8_t> in_union_, in_new_;
  std::vector<std::uint32_t> common_;
  std::vector<std::uint64_t> holders_offsets_;  // tile -> queries selecting it
  std::vector<std::uint32_t> holders_;
};

BatchPlanEntry close_batch(const SelectedRegionIndex& index,
                           std::span<const BatchQueryFeatures> queries, std::uint32_t width,
                           OpenBatch&& batch) {
  BatchPlanEntry e;
  e.lane_width = width;
  e.valid_lane_mask = low_lanes(static_cast<std::uint32_t>(batch.lanes.size()));
  e.query_indices_by_lane.assign(width, invalid_batch_query_index);
  e.query_ids_by_lane.assign(width, invalid_batch_query_id);
  e.expansion_generations_by_lane.assign(width, 0U);
  e.union_tiles = std::move(batch.tiles);
  e.union_tile_lane_masks.assign(e.union_tiles.size(), 0U);
  for (std::uint32_t lane = 0; lane < batch.lanes.size(); ++lane) {
    const BatchQueryFeatures& f = queries[batch.lanes[lane]];
    e.query_indices_by_lane[lane] = batch.lanes[lane];
    e.query_ids_by_lane[lane] = f.query_id;
    e.expansion_generations_by_lane[lane] = f.expansion_generation;
    e.selected_lane_vertex_count += f.selected_vertex_count;
    e.selected_lane_edge_estimate += f.selected_edge_estimate;
    // selected tiles are a subset of the union: walk both in order
    std::size_t u = 0;
    for (TileId t : f.selected_tiles) {
      while (e.union_tiles[u] < t) ++u;
      e.union_tile_lane_masks[u] |= LaneMask{1} << lane;
    }
  }
  e.union_vertex_count = index.selected_vertex_count(e.union_tiles);
  e.union_edge_estimate = index.selected_edge_count(e.union_tiles);
  return e;
}

BatchPlan plan_batches(const SelectedRegionIndex& index, std::span<const BatchQueryFeatures> queries,
                       const BatchPlannerPolicy& policy) {
  BatchPlan plan;
  plan.policy = policy;
  plan.input_query_count = static_cast<std::uint32_t>(queries.size());
  std::vector<std::uint32_t> anchors(queries.size());
  for (std::uint32_t i = 0; i < anchors.size(); ++i) anchors[i] = i;
  std::sort(anchors.begin(), anchors.end(), [&](std::uint32_t a, std::uint32_t b) {
    const BatchQueryFeatures& x = queries[a];
    const BatchQueryFeatures& y = queries[b];
    // larger work first, then earlier generation, then smaller id
    return std::make_tuple(y.selected_edge_estimate, y.selected_vertex_count,
                           y.selected_tiles.size(), x.expansion_generation, x.query_id) <
           std::make_tuple(x.selected_edge_estimate, x.selected_vertex_count,
                           x.selected_tiles.size(), y.expansion_generation, y.query_id);
  });
  const BatchFraction min_jaccard{policy.minimum_jaccard_numerator,
                                  policy.minimum_jaccard_denominator};
  const BatchFraction max_inflation{policy.maximum_union_inflation_numerator,
                                    policy.maximum_union_inflation_denominator};
  std::vector<bool> taken(queries.size(), false);
  Planner planner{index, queries};
  for (std::uint32_t anchor_index : anchors) {
    if (taken[anchor_index]) continue;
    const BatchQueryFeatures& anchor = queries[anchor_index];
    taken[anchor_index] = true;
    OpenBatch batch;
    planner.open(batch, anchor_index);
    while (batch.lanes.size() < policy.lane_width) {
      bool found = false;
      Offer best;
      for (std::uint32_t q = 0; q < queries.size(); ++q) {
        if (taken[q]) continue;
        Offer o = planner.offer(batch, q, anchor);
        if (order(o.jaccard, min_jaccard) < 0 || order(o.inflation, max_inflation) > 0) continue;
        if (!found || planner.before(batch, o, best)) {
          best = o;
          found = true;
        }
      }
      if (!found) break;
      taken[best.query] = true;
      planner.admit(batch, best);
    }
    planner.close(batch);
    plan.batches.push_back(close_batch(index, queries, policy.lane_width, std::move(batch)));
  }
  return plan;
}

BatchPlanValidationResult fault(BatchPlanValidationErrorCode code,
                                std::size_t batch = BatchPlanValidationResult::no_position,
                                std::size_t lane = BatchPlanValidationResult::no_position) {
  return BatchPlanValidationResult{code, batch, lane};
}

// One TSV row of the 17-column bfnew.batch-plan.v1 image; empty cells are '-'.
class Row {
 public:
  explicit Row(const char* record) { cells_[0] = record; }
  Row& set(int column, std::uint64_t value) {
    cells_[column] = std::to_string(value);
    return *this;
  }
  Row& set(int column, std::string value) {
    cells_[column] = std::move(value);
    return *this;
  }
  void append_to(std::string& out) const {
    for (int c = 0; c < 17; ++c) {
      if (c) out += '\t';
      out += cells_[c].empty() ? std::string("-") : cells_[c];
    }
    out += '\n';
  }

 private:
  std::array<std::string, 17> cells_;
};

// Column numbers of the TSV image.
enum Column : int {
  kBatch = 1, kPosition, kLaneWidth, kValidMask, kQueryIndex, kQueryId, kGeneration, kTileId,
  kTileMask, kUnionVertices, kUnionEdges, kLaneVertices, kLaneEdges, kInputQueries,
  kMinJaccard, kMaxInflation,
};

}  // namespace

// ---------------------------------------------------------------------------
SelectedRegionIndex::SelectedRegionIndex(const WeightedGraph& graph,
                                         const TileRunLayout64& tile_runs)
    : graph_memo_{graph.memo_handle()} {
  const auto offsets = graph.tile_vertex_offsets();
  const auto owners = graph.owner_tiles();
  const std::size_t V = graph.vertex_count();
  if (!graph.has_spatial_ordering() || offsets.size() != graph.tile_coordinates().size() + 1U ||
      tile_runs.csr_row_run_offsets.size() != V + 1U ||
      tile_runs.csr_run_edge_offsets.size() != tile_runs.csr_run_destination_tiles.size() + 1U ||
      tile_runs.csr_row_run_offsets[V] != tile_runs.csr_run_destination_tiles.size())
    throw std::invalid_argument{"selected-region index needs a spatial graph and its tile runs"};
  const std::size_t T = graph.tile_coordinates().size();
  if (T > 0xFFFFFFFFull) throw std::overflow_error{"tile count exceeds TileId"};
  vertex_counts_.resize(T);
  for (std::size_t t = 0; t < T; ++t) vertex_counts_[t] = offsets[t + 1] - offsets[t];
  // (source tile, destination tile) edge counts, coalesced per pair.  Rows
  // are tile-major, so each source tile's runs are contiguous: accumulate
  // them into a dense per-destination slot and emit the touched slots.
  std::vector<std::uint64_t> acc(T, 0U);
  std::vector<std::uint32_t> touched;
  pair_offsets_.assign(T + 1U, 0U);
  for (std::size_t t = 0; t < T; ++t) {
    for (std::size_t v = offsets[t]; v < offsets[t + 1]; ++v) {
      if (owners[v].value() != t) throw std::invalid_argument{"selected-region index needs tile-major vertices"};
      for (auto r = tile_runs.csr_row_run_offsets[v]; r < tile_runs.csr_row_run_offsets[v + 1]; ++r) {
        const std::uint32_t to = tile_runs.csr_run_destination_tiles[r].value();
        if (to >= T) throw std::invalid_argument{"tile run names a tile outside the graph"};
        if (acc[to] == 0U) touched.push_back(to);
        acc[to] += tile_runs.csr_run_edge_offsets[r + 1] - tile_runs.csr_run_edge_offsets[r];
      }
    }
    std::sort(touched.begin(), touched.end());
    for (std::uint32_t to : touched) {
      pair_tiles_.push_back(to);
      pair_edges_.push_back(acc[to]);
      acc[to] = 0U;
    }
    pair_offsets_[t + 1] = pair_tiles_.size();
    touched.clear();
  }
  // The transpose: pairs grouped by destination, sources ascending.
  in_pair_offsets_.assign(T + 1U, 0U);
  for (std::uint32_t to : pair_tiles_) ++in_pair_offsets_[to + 1U];
  for (std::size_t t = 0; t < T; ++t) in_pair_offsets_[t + 1] += in_pair_offsets_[t];
  in_pair_tiles_.resize(pair_tiles_.size());
  in_pair_edges_.resize(pair_tiles_.size());
  std::vector<std::uint64_t> fill(in_pair_offsets_.begin(), in_pair_offsets_.end() - 1);
  for (std::size_t from = 0; from < T; ++from)
    for (auto i = pair_offsets_[from]; i < pair_offsets_[from + 1]; ++i) {
      const std::uint64_t at = fill[pair_tiles_[i]]++;
      in_pair_tiles_[at] = static_cast<std::uint32_t>(from);
      in_pair_edges_[at] = pair_edges_[i];
    }
}

const SelectedRegionIndex& graph_region_index(const WeightedGraph& graph) {
  WeightedGraphMemo& memo = graph.memo();
  std::call_once(memo.region_index_once, [&] {
    memo.region_index = std::make_shared<const SelectedRegionIndex>(graph, graph_tile_runs(graph));
  });
  return *static_cast<const SelectedRegionIndex*>(memo.region_index.get());
}

SelectedRegionIndex::TilePairs SelectedRegionIndex::pairs_from(std::uint32_t tile) const noexcept {
  const std::size_t a = pair_offsets_[tile], b = pair_offsets_[tile + 1U];
  return {std::span<const std::uint32_t>(pair_tiles_).subspan(a, b - a),
          std::span<const std::uint64_t>(pair_edges_).subspan(a, b - a)};
}

SelectedRegionIndex::TilePairs SelectedRegionIndex::pairs_into(std::uint32_t tile) const noexcept {
  const std::size_t a = in_pair_offsets_[tile], b = in_pair_offsets_[tile + 1U];
  return {std::span<const std::uint32_t>(in_pair_tiles_).subspan(a, b - a),
          std::span<const std::uint64_t>(in_pair_edges_).subspan(a, b - a)};
}

std::uint64_t SelectedRegionIndex::selected_vertex_count(std::span<const TileId> tiles) const {
  if (!strictly_increasing(tiles)) throw std::invalid_argument{"selected tiles must be sorted and unique"};
  std::uint64_t n = 0;
  for (TileId t : tiles) {
    if (t.value() >= tile_count()) throw std::out_of_range{"selected tile outside the graph"};
    n += vertex_counts_[t.value()];
  }
  return n;
}

std::uint64_t SelectedRegionIndex::selected_edge_count(std::span<const TileId> tiles) const {
  if (!strictly_increasing(tiles)) throw std::invalid_argument{"selected tiles must be sorted and unique"};
  std::uint64_t n = 0;
  for (TileId t : tiles) {
    if (t.value() >= tile_count()) throw std::out_of_range{"selected tile outside the graph"};
    for (auto i = pair_offsets_[t.value()]; i < pair_offsets_[t.value() + 1]; ++i)
      if (std::binary_search(tiles.begin(), tiles.end(), TileId{pair_tiles_[i]})) n += pair_edges_[i];
  }
  return n;
}

BatchPlannerPolicyError validate_batch_planner_policy(const BatchPlannerPolicy& p) noexcept {
  if (!supported_width(p.lane_width)) return BatchPlannerPolicyError::unsupported_lane_width;
  if (p.minimum_jaccard_denominator == 0U ||
      p.minimum_jaccard_numerator > p.minimum_jaccard_denominator)
    return BatchPlannerPolicyError::invalid_jaccard_fraction;
  if (p.maximum_union_inflation_denominator == 0U ||
      p.maximum_union_inflation_numerator < p.maximum_union_inflation_denominator)
    return BatchPlannerPolicyError::invalid_union_inflation_fraction;
  return BatchPlannerPolicyError::none;
}

std::vector<BatchQueryFeatures> make_batch_query_features(const WeightedGraph& graph,
                                                          const SelectedRegionIndex& index,
                                                          std::span<const RouteQuery> queries) {
  if (!graph.has_spatial_ordering() || !index.matches_graph(graph))
    throw std::invalid_argument{"batch query features need the index of this spatial graph"};
  std::vector<BatchQueryFeatures> out;
  for (const RouteQuery& q : queries) {
    if (!validate_route_query(graph, q).ok())
      throw std::invalid_argument{"batch query features need valid route queries"};
    BatchQueryFeatures f;
    f.query_id = q.query_id;
    f.expansion_generation = q.expansion_generation;
    f.source_count = static_cast<std::uint32_t>(q.sources.size());
    f.target_count = static_cast<std::uint32_t>(q.targets.size());
    f.source_tiles = owner_tiles_of(graph, q.sources);
    f.target_tiles = owner_tiles_of(graph, q.targets);
    f.selected_tiles = q.selected_tiles;
    f.selected_vertex_count = index.selected_vertex_count(f.selected_tiles);
    f.selected_edge_estimate = index.selected_edge_count(f.selected_tiles);
    out.push_back(std::move(f));
  }
  if (!ids_unique(out)) throw std::invalid_argument{"batch query ids must be unique"};
  std::sort(out.begin(), out.end(),
            [](const BatchQueryFeatures& a, const BatchQueryFeatures& b) { return a.query_id < b.query_id; });
  return out;
}

BatchFraction tile_set_jaccard(std::span<const TileId> left, std::span<const TileId> right) {
  if (!strictly_increasing(left) || !strictly_increasing(right))
    throw std::invalid_argument{"Jaccard tile sets must be sorted and unique"};
  const std::uint64_t both = common_count(left, right);
  const std::uint64_t either = left.size() + right.size() - both;
  return either == 0U ? BatchFraction{1U, 1U} : BatchFraction{both, either};
}

BatchFraction batch_union_tile_inflation(std::span<const BatchQueryFeatures> queries,
                                         std::span<const std::uint32_t> query_indices,
                                         std::span<const TileId> union_tiles) {
  if (query_indices.empty() || !strictly_increasing(union_tiles))
    throw std::invalid_argument{"batch inflation needs lanes and a canonical union"};
  std::vector<std::uint32_t> seen(query_indices.begin(), query_indices.end());
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    throw std::invalid_argument{"batch inflation lanes must be unique"};
  std::uint64_t lane_tiles = 0;
  for (std::uint32_t i : query_indices) {
    if (i >= queries.size()) throw std::invalid_argument{"batch inflation lane out of range"};
    const auto& tiles = queries[i].selected_tiles;
    if (tiles.empty() || !strictly_increasing(tiles) ||
        !std::includes(union_tiles.begin(), union_tiles.end(), tiles.begin(), tiles.end()))
      throw std::invalid_argument{"batch inflation lane is not covered by the union"};
    lane_tiles += tiles.size();
  }
  return BatchFraction{query_indices.size() * union_tiles.size(), lane_tiles};
}

BatchPlan make_overlapping_batch_plan(const SelectedRegionIndex& index,
                                      std::span<const BatchQueryFeatures> queries,
                                      const BatchPlannerPolicy& policy) {
  if (validate_batch_planner_policy(policy) != BatchPlannerPolicyError::none)
    throw std::invalid_argument{"batch planner policy is invalid"};
  require_plannable(index, queries);
  return plan_batches(index, queries, policy);
}

BatchPlanFamily make_standard_batch_plan_family(const SelectedRegionIndex& index,
                                                std::span<const BatchQueryFeatures> queries,
                                                BatchPlannerPolicy policy) {
  if (validate_batch_planner_policy(policy) != BatchPlannerPolicyError::none)
    throw std::invalid_argument{"batch planner policy is invalid"};
  require_plannable(index, queries);
  BatchPlanFamily family;
  const std::uint32_t widths[3] = {32U, 16U, 8U};
  for (int i = 0; i < 3; ++i) {
    policy.lane_width = widths[i];
    family.plans[i] = plan_batches(index, queries, policy);
  }
  return family;
}

BatchPlanValidationResult validate_batch_plan(const SelectedRegionIndex& index,
                                              std::span<const BatchQueryFeatures> queries,
                                              const BatchPlan& plan) noexcept {
  using E = BatchPlanValidationErrorCode;
  try {
    if (validate_batch_planner_policy(plan.policy) != BatchPlannerPolicyError::none)
      return fault(E::invalid_policy);
    if (queries.size() > 0xFFFFFFFFull) return fault(E::query_count_overflow);
    if (queries.empty()) return fault(E::empty_query_set);
    for (const auto& q : queries)
      if (!features_valid(index, q)) return fault(E::invalid_query_feature);
    if (!ids_unique(queries)) return fault(E::duplicate_query_id);
    if (plan.input_query_count != queries.size()) return fault(E::invalid_batch_shape);
    if (plan.batches.empty()) return fault(E::empty_plan);
    std::vector<bool> assigned(queries.size(), false);
    for (std::size_t bi = 0; bi < plan.batches.size(); ++bi) {
      const BatchPlanEntry& b = plan.batches[bi];
      const std::uint32_t W = b.lane_width;
      if (W != plan.policy.lane_width || !supported_width(W)) return fault(E::invalid_batch_width, bi);
      if (b.query_indices_by_lane.size() != W || b.query_ids_by_lane.size() != W ||
          b.expansion_generations_by_lane.size() != W || b.union_tiles.empty() ||
          b.union_tile_lane_masks.size() != b.union_tiles.size())
        return fault(E::invalid_batch_shape, bi);
      const auto valid = static_cast<std::uint32_t>(std::popcount(b.valid_lane_mask));
      if (valid == 0U || valid > W || b.valid_lane_mask != low_lanes(valid))
        return fault(E::invalid_valid_lane_mask, bi);
      std::vector<TileId> expect_union;
      std::uint64_t lane_vertices = 0, lane_edges = 0;
      for (std::uint32_t lane = 0; lane < W; ++lane) {
        const std::uint32_t qi = b.query_indices_by_lane[lane];
        if (lane >= valid) {
          if (qi != invalid_batch_query_index || b.query_ids_by_lane[lane] != invalid_batch_query_id ||
              b.expansion_generations_by_lane[lane] != 0U)
            return fault(E::invalid_padding, bi, lane);
          continue;
        }
        if (qi >= queries.size()) return fault(E::query_index_out_of_range, bi, lane);
        const BatchQueryFeatures& f = queries[qi];
        if (b.query_ids_by_lane[lane] != f.query_id ||
            b.expansion_generations_by_lane[lane] != f.expansion_generation)
          return fault(E::query_identity_mismatch, bi, lane);
        if (assigned[qi]) return fault(E::duplicate_assignment, bi, lane);
        assigned[qi] = true;
        expect_union = merged(expect_union, f.selected_tiles);
        lane_vertices += f.selected_vertex_count;
        lane_edges += f.selected_edge_estimate;
      }
      if (b.union_tiles != expect_union) return fault(E::union_tile_mismatch, bi);
      for (std::size_t t = 0; t < b.union_tiles.size(); ++t) {
        LaneMask m = 0;
        for (std::uint32_t lane = 0; lane < valid; ++lane) {
          const auto& sel = queries[b.query_indices_by_lane[lane]].selected_tiles;
          if (std::binary_search(sel.begin(), sel.end(), b.union_tiles[t])) m |= LaneMask{1} << lane;
        }
        if (b.union_tile_lane_masks[t] != m) return fault(E::tile_lane_mask_mismatch, bi, t);
      }
      if (b.union_vertex_count != index.selected_vertex_count(b.union_tiles) ||
          b.union_edge_estimate != index.selected_edge_count(b.union_tiles) ||
          b.selected_lane_vertex_count != lane_vertices || b.selected_lane_edge_estimate != lane_edges)
        return fault(E::estimate_mismatch, bi);
    }
    if (std::find(assigned.begin(), assigned.end(), false) != assigned.end())
      return fault(E::omitted_query);
    // A consistent but differently chosen plan is still not this planner's.
    if (plan_batches(index, queries, plan.policy) != plan) return fault(E::invalid_batch_shape);
    return {};
  } catch (...) {
    return fault(E::invalid_query_feature);
  }
}

std::string serialize_batch_plan_tsv(const BatchPlan& plan) {
  if (validate_batch_planner_policy(plan.policy) != BatchPlannerPolicyError::none ||
      plan.input_query_count == 0U || plan.batches.empty())
    throw std::invalid_argument{"cannot serialise an invalid batch plan"};
  const auto fraction = [](std::uint32_t n, std::uint32_t d) {
    return std::to_string(n) + "/" + std::to_string(d);
  };
  std::string out = "schema\tbfnew.batch-plan.v1\n";
  out += "record\tbatch\tposition\tlane_width\tvalid_lane_mask\tquery_index\tquery_id\t"
         "expansion_generation\ttile_id\ttile_lane_mask\tunion_vertex_count\t"
         "union_edge_estimate\tselected_lane_vertex_count\tselected_lane_edge_estimate\t"
         "input_query_count\tminimum_jaccard\tmaximum_union_inflation\n";
  const BatchPlannerPolicy& p = plan.policy;
  Row("plan")
      .set(kLaneWidth, p.lane_width)
      .set(kInputQueries, plan.input_query_count)
      .set(kMinJaccard, fraction(p.minimum_jaccard_numerator, p.minimum_jaccard_denominator))
      .set(kMaxInflation,
           fraction(p.maximum_union_inflation_numerator, p.maximum_union_inflation_denominator))
      .append_to(out);
  for (std::size_t bi = 0; bi < plan.batches.size(); ++bi) {
    const BatchPlanEntry& b = plan.batches[bi];
    const auto valid = static_cast<std::uint32_t>(std::popcount(b.valid_lane_mask));
    if (b.lane_width != p.lane_width || b.query_indices_by_lane.size() != b.lane_width ||
        b.query_ids_by_lane.size() != b.lane_width ||
        b.expansion_generations_by_lane.size() != b.lane_width || valid == 0U ||
        valid > b.lane_width || b.valid_lane_mask != low_lanes(valid) ||
        b.union_tiles.size() != b.union_tile_lane_masks.size())
      throw std::invalid_argument{"cannot serialise a malformed batch entry"};
    Row("batch")
        .set(kBatch, bi)
        .set(kLaneWidth, b.lane_width)
        .set(kValidMask, b.valid_lane_mask)
        .set(kUnionVertices, b.union_vertex_count)
        .set(kUnionEdges, b.union_edge_estimate)
        .set(kLaneVertices, b.selected_lane_vertex_count)
        .set(kLaneEdges, b.selected_lane_edge_estimate)
        .append_to(out);
    for (std::uint32_t lane = 0; lane < b.lane_width; ++lane)
      Row("lane")
          .set(kBatch, bi)
          .set(kPosition, lane)
          .set(kLaneWidth, b.lane_width)
          .set(kValidMask, b.valid_lane_mask)
          .set(kQueryIndex, b.query_indices_by_lane[lane])
          .set(kQueryId, b.query_ids_by_lane[lane].value())
          .set(kGeneration, b.expansion_generations_by_lane[lane])
          .append_to(out);
    for (std::size_t t = 0; t < b.union_tiles.size(); ++t)
      Row("tile")
          .set(kBatch, bi)
          .set(kPosition, t)
          .set(kLaneWidth, b.lane_width)
          .set(kValidMask, b.valid_lane_mask)
          .set(kTileId, b.union_tiles[t].value())
          .set(kTileMask, b.union_tile_lane_masks[t])
          .append_to(out);
  }
  return out;
}

std::uint64_t fingerprint_batch_plan(const BatchPlan& plan) {
  // 64-bit FNV-1a (standard offset basis and prime) over the TSV bytes.
  std::uint64_t h = 14695981039346656037ULL;
  for (unsigned char c : serialize_batch_plan_tsv(plan)) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

void prepare_batch_device_description(const WeightedGraph& graph,
                                      std::span<const RouteQuery> queries,
                                      std::span<const BatchQueryFeatures> features,
                                      const BatchPlanEntry& batch,
                                      BatchDeviceDescription& output) {
  const auto valid = static_cast<std::uint32_t>(std::popcount(batch.valid_lane_mask));
  if (queries.size() != features.size() || !supported_width(batch.lane_width) || valid == 0U ||
      valid > batch.lane_width || batch.valid_lane_mask != low_lanes(valid) ||
      batch.query_indices_by_lane.size() != batch.lane_width ||
      batch.query_ids_by_lane.size() != batch.lane_width)
    throw std::invalid_argument{"batch description: plan entry does not fit the queries"};
  std::vector<RouteQuery> lanes;
  lanes.reserve(valid);
  for (std::uint32_t lane = 0; lane < valid; ++lane) {
    const std::uint32_t qi = batch.query_indices_by_lane[lane];
    if (qi >= queries.size() || queries[qi].query_id != features[qi].query_id ||
        batch.query_ids_by_lane[lane] != queries[qi].query_id ||
        queries[qi].selected_tiles != features[qi].selected_tiles)
      throw std::invalid_argument{"batch description: lane " + std::to_string(lane) +
                                  " names another query"};
    lanes.push_back(queries[qi]);
  }
  prepare_batch_device_description(graph, lanes, batch.lane_width, output);
  if (output.union_tiles.size() != batch.union_tiles.size() ||
      !std::equal(output.union_tiles.begin(), output.union_tiles.end(), batch.union_tiles.begin(),
                  [](std::uint32_t t, TileId u) { return t == u.value(); }))
    throw std::invalid_argument{"batch description: plan union differs from its lanes"};
}

}  // namespace bfnew
