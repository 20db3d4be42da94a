// graph_builder.cpp -- InputGraph, canonical EdgeIds, the spatial reorder and
// the tile directory.
//
// Reference behaviour restated (not translated):
//   EdgeId      = rank of the edge in (source, destination, canonical weight
//                 bits, provenance) order (graph_builder.cpp:182-214);
//   reorder     = vertices sorted by (located first, tile_y, tile_x,
//                 resource class, Morton key of the in-tile coordinate,
//                 original id); unlocated vertices form a trailing spill tile
//                 ordered by (class, original id) (reorder.cpp:151-273);
//   CSR / CSC   = rows by (source, destination tile, destination, EdgeId),
//                 columns by (destination, source tile, source, EdgeId)
//                 (reorder.cpp:293-342).
// Implementation: one index sort per key (stable, so equal keys keep input
// order), then counting-sort placement into the offset arrays.
#include "bfnew/graph.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>

namespace bfnew {

float canonicalize_weight(float weight) noexcept { return weight == 0.0F ? 0.0F : weight; }
bool is_valid_weight(float weight) noexcept { return std::isfinite(weight) && !(weight < 0.0F); }
std::uint32_t canonical_weight_bits(float weight) noexcept {
  return std::bit_cast<std::uint32_t>(canonicalize_weight(weight));
}
bool is_valid_vertex_id(VertexId id, std::uint32_t vertex_count) noexcept {
  return id.value() < vertex_count;
}

InputGraph::InputGraph(std::vector<VertexMetadata> vertices, std::vector<EdgeInputRecord> edges)
    : vertices_{std::move(vertices)}, edges_{std::move(edges)} {
  if (!std::in_range<std::uint32_t>(vertices_.size()))
    throw std::length_error{"vertex count exceeds the 32-bit representation"};
  vertex_count_ = static_cast<std::uint32_t>(vertices_.size());
  for (VertexMetadata& v : vertices_)
    if (!v.has_location) v.x = v.y = 0;
  for (std::size_t i = 0; i < edges_.size(); ++i) {
    EdgeInputRecord& e = edges_[i];
    e.weight = canonicalize_weight(e.weight);
    if (!std::isfinite(e.weight))
      throw std::invalid_argument{"edge weight must be finite at edge input index " + std::to_string(i)};
    if (e.weight < 0.0F)
      throw std::invalid_argument{"edge weight must be nonnegative at edge input index " + std::to_string(i)};
    if (e.source.value() >= vertex_count_)
      throw std::out_of_range{"edge source is outside the vertex range"};
    if (e.destination.value() >= vertex_count_)
      throw std::out_of_range{"edge destination is outside the vertex range"};
  }
}

std::uint64_t MortonLocalityPolicy::operator()(std::uint32_t local_x,
                                               std::uint32_t local_y) const noexcept {
  // Interleave: bit i of x -> bit 2i, bit i of y -> bit 2i+1.
  std::uint64_t key = 0;
  for (unsigned bit = 0; bit < 32; ++bit) {
    key |= static_cast<std::uint64_t>((local_x >> bit) & 1U) << (2 * bit);
    key |= static_cast<std::uint64_t>((local_y >> bit) & 1U) << (2 * bit + 1);
  }
  return key;
}

namespace {

// EdgeId of every input edge (its rank in canonical order).
std::vector<std::uint64_t> canonical_edge_ids(const InputGraph& input) {
  const auto edges = input.edges();
  std::vector<std::uint64_t> order(edges.size());
  std::iota(order.begin(), order.end(), 0ULL);
  std::stable_sort(order.begin(), order.end(), [&](std::uint64_t a, std::uint64_t b) {
    const EdgeInputRecord& l = edges[a];
    const EdgeInputRecord& r = edges[b];
    return std::tuple{l.source.value(), l.destination.value(), canonical_weight_bits(l.weight),
                      l.provenance.domain, l.provenance.kind_and_flags,
                      l.provenance.source_record} <
           std::tuple{r.source.value(), r.destination.value(), canonical_weight_bits(r.weight),
                      r.provenance.domain, r.provenance.kind_and_flags,
                      r.provenance.source_record};
  });
  std::vector<std::uint64_t> id(edges.size());
  for (std::uint64_t rank = 0; rank < order.size(); ++rank) id[order[rank]] = rank;
  return id;
}

struct Placed {
  std::uint32_t major, minor, minor_tile;
  float weight;
  std::uint64_t edge_id;
};

// Fills offsets/endpoints/weights/ids of one orientation sorted by
// (major, minor tile, minor, EdgeId).
void place(std::uint32_t vertex_count, std::vector<Placed>& items, std::vector<EdgeOffset>& offsets,
           std::vector<VertexId>& endpoints, std::vector<float>& weights, std::vector<EdgeId>& ids) {
  std::sort(items.begin(), items.end(), [](const Placed& a, const Placed& b) {
    return std::tuple{a.major, a.minor_tile, a.minor, a.edge_id} <
           std::tuple{b.major, b.minor_tile, b.minor, b.edge_id};
  });
  offsets.assign(static_cast<std::size_t>(vertex_count) + 1, 0);
  endpoints.resize(items.size());
  weights.resize(items.size());
  ids.resize(items.size());
  for (std::size_t i = 0; i < items.size(); ++i) {
    ++offsets[items[i].major + 1];
    endpoints[i] = VertexId{items[i].minor};
    weights[i] = items[i].weight;
    ids[i] = EdgeId{items[i].edge_id};
  }
  for (std::size_t v = 1; v < offsets.size(); ++v) offsets[v] += offsets[v - 1];
}

std::int64_t floor_div(std::int64_t n, std::int64_t d) {
  std::int64_t q = n / d;
  if ((n % d) != 0 && ((n < 0) != (d < 0))) --q;
  return q;
}

}  // namespace

WeightedGraph build_weighted_graph(const InputGraph& input) {
  WeightedGraph g;
  const std::uint32_t V = input.vertex_count();
  g.vertex_count_ = V;
  g.edge_count_ = input.edge_count();
  g.vertices_.assign(input.vertices().begin(), input.vertices().end());
  g.old_to_new_.resize(V);
  g.new_to_old_.resize(V);
  for (std::uint32_t v = 0; v < V; ++v) g.old_to_new_[v] = g.new_to_old_[v] = VertexId{v};
  const std::vector<std::uint64_t> ids = canonical_edge_ids(input);
  const auto edges = input.edges();
  g.edge_provenance_.resize(edges.size());
  std::vector<Placed> items(edges.size());
  for (std::size_t i = 0; i < edges.size(); ++i) {
    g.edge_provenance_[ids[i]] = edges[i].provenance;
    items[i] = Placed{edges[i].source.value(), edges[i].destination.value(), 0, edges[i].weight, ids[i]};
  }
  place(V, items, g.out_offsets_, g.out_destinations_, g.out_weights_, g.out_edge_ids_);
  for (std::size_t i = 0; i < edges.size(); ++i)
    items[i] = Placed{edges[i].destination.value(), edges[i].source.value(), 0, edges[i].weight, ids[i]};
  place(V, items, g.in_offsets_, g.in_sources_, g.in_weights_, g.in_edge_ids_);
  return g;
}

WeightedGraph build_spatially_ordered_graph(const InputGraph& input,
                                            const SpatialOrderConfig& config,
                                            const LocalityKeyPolicy& locality_policy) {
  if (config.tile_width == 0U || config.tile_height == 0U)
    throw std::invalid_argument{"tile width and height must both be positive"};
  if (!locality_policy) throw std::invalid_argument{"locality-key policy must be callable"};
  const std::uint32_t V = input.vertex_count();
  struct Key {
    bool located;
    std::int64_t ty, tx;
    std::uint16_t cls;
    std::uint64_t locality;
    std::uint32_t original;
  };
  std::vector<Key> keys(V);
  std::vector<std::pair<std::int64_t, std::int64_t>> tiles;  // (y, x)
  for (std::uint32_t v = 0; v < V; ++v) {
    const VertexMetadata& m = input.vertices()[v];
    Key k{m.has_location, 0, 0, m.resource_class.value(), 0, v};
    if (m.has_location) {
      const std::int64_t dx = static_cast<std::int64_t>(m.x) - config.origin_x;
      const std::int64_t dy = static_cast<std::int64_t>(m.y) - config.origin_y;
      k.tx = floor_div(dx, config.tile_width);
      k.ty = floor_div(dy, config.tile_height);
      const auto lx = static_cast<std::uint32_t>(dx - k.tx * config.tile_width);
      const auto ly = static_cast<std::uint32_t>(dy - k.ty * config.tile_height);
      k.locality = locality_policy(lx, ly);
      tiles.emplace_back(k.ty, k.tx);
    }
    keys[v] = k;
  }
  std::sort(tiles.begin(), tiles.end());
  tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
  std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
    if (a.located != b.located) return a.located;
    if (!a.located) return std::tuple{a.cls, a.original} < std::tuple{b.cls, b.original};
    return std::tuple{a.ty, a.tx, a.cls, a.locality, a.original} <
           std::tuple{b.ty, b.tx, b.cls, b.locality, b.original};
  });

  WeightedGraph g;
  g.vertex_count_ = V;
  g.edge_count_ = input.edge_count();
  g.spatial_ = true;
  g.config_ = config;
  const auto spill = static_cast<std::uint32_t>(tiles.size());
  g.tile_coordinates_.reserve(tiles.size() + 1);
  for (const auto& [y, x] : tiles) g.tile_coordinates_.push_back(TileCoordinate::located(x, y));
  g.tile_coordinates_.push_back(TileCoordinate::spill());
  g.old_to_new_.resize(V);
  g.new_to_old_.resize(V);
  g.owner_tiles_.resize(V);
  g.vertices_.resize(V);
  g.tile_vertex_offsets_.assign(tiles.size() + 2, 0);
  for (std::uint32_t n = 0; n < V; ++n) {
    const Key& k = keys[n];
    std::uint32_t owner = spill;
    if (k.located) {
      owner = static_cast<std::uint32_t>(
          std::lower_bound(tiles.begin(), tiles.end(), std::pair{k.ty, k.tx}) - tiles.begin());
    }
    g.old_to_new_[k.original] = VertexId{n};
    g.new_to_old_[n] = VertexId{k.original};
    g.owner_tiles_[n] = TileId{owner};
    g.vertices_[n] = input.vertices()[k.original];
    ++g.tile_vertex_offsets_[owner + 1];
  }
  for (std::size_t t = 1; t < g.tile_vertex_offsets_.size(); ++t)
    g.tile_vertex_offsets_[t] += g.tile_vertex_offsets_[t - 1];

  const std::vector<std::uint64_t> ids = canonical_edge_ids(input);
  const auto edges = input.edges();
  g.edge_provenance_.resize(edges.size());
  std::vector<Placed> items(edges.size());
  for (std::size_t i = 0; i < edges.size(); ++i) {
    g.edge_provenance_[ids[i]] = edges[i].provenance;
    const std::uint32_t s = g.old_to_new_[edges[i].source.value()].value();
    const std::uint32_t d = g.old_to_new_[edges[i].destination.value()].value();
    items[i] = Placed{s, d, g.owner_tiles_[d].value(), edges[i].weight, ids[i]};
  }
  place(V, items, g.out_offsets_, g.out_destinations_, g.out_weights_, g.out_edge_ids_);
  for (std::size_t i = 0; i < edges.size(); ++i) {
    const std::uint32_t s = g.old_to_new_[edges[i].source.value()].value();
    const std::uint32_t d = g.old_to_new_[edges[i].destination.value()].value();
    items[i] = Placed{d, s, g.owner_tiles_[s].value(), edges[i].weight, ids[i]};
  }
  place(V, items, g.in_offsets_, g.in_sources_, g.in_weights_, g.in_edge_ids_);
  return g;
}

const char* validate_weighted_graph_message(const WeightedGraph& g) {
  const std::uint32_t V = g.vertex_count();
  const auto out = g.outgoing();
  const auto in = g.incoming();
  if (out.row_offsets.size() != V + 1ULL || in.column_offsets.size() != V + 1ULL)
    return "offset array sizes";
  if (out.row_offsets.front() != 0 || out.row_offsets.back() != g.edge_count() ||
      in.column_offsets.front() != 0 || in.column_offsets.back() != g.edge_count())
    return "offset bounds";
  std::vector<std::uint8_t> seen(g.edge_count(), 0);
  for (std::uint32_t v = 0; v < V; ++v) {
    if (out.row_offsets[v + 1] < out.row_offsets[v] || in.column_offsets[v + 1] < in.column_offsets[v])
      return "offsets not monotone";
    for (auto p = out.row_offsets[v]; p < out.row_offsets[v + 1]; ++p) {
      if (out.destinations[p].value() >= V) return "destination out of range";
      if (!is_valid_weight(out.weights[p])) return "invalid stored weight";
      if (std::bit_cast<std::uint32_t>(out.weights[p]) == 0x80000000U) return "noncanonical zero";
      if (out.edge_ids[p].value() >= g.edge_count() || seen[out.edge_ids[p].value()]++)
        return "edge id out of range or duplicated";
    }
  }
  if (g.has_spatial_ordering()) {
    const auto offs = g.tile_vertex_offsets();
    if (offs.size() != g.tile_coordinates().size() + 1 || offs.front() != 0 || offs.back() != V)
      return "tile vertex offsets";
    for (std::size_t t = 0; t + 1 < offs.size(); ++t)
      for (auto v = offs[t]; v < offs[t + 1]; ++v)
        if (g.owner_tiles()[v].value() != t) return "noncontiguous tile span";
  }
  for (std::uint32_t v = 0; v < V; ++v)
    if (g.new_to_old()[g.old_to_new()[v].value()].value() != v) return "permutation round trip";
  return "";
}

TileDirectory build_tile_directory(const WeightedGraph& g) {
  if (!g.has_spatial_ordering()) throw std::invalid_argument{"tile directory requires a spatially ordered graph"};
  const std::size_t T = g.tile_coordinates().size();
  const auto owner = g.owner_tiles();
  const auto out = g.outgoing();
  const auto in = g.incoming();
  std::vector<std::vector<TileId>> nbr(T);
  std::vector<std::vector<EdgeId>> internal(T), outgoing(T), incoming(T);
  std::vector<std::vector<VertexId>> halo(T);
  // Chebyshev-1 located neighbours.
  std::vector<std::pair<std::pair<std::int64_t, std::int64_t>, std::uint32_t>> coords;
  for (std::size_t t = 0; t + 1 < T; ++t)
    coords.push_back({{g.tile_coordinates()[t].tile_x, g.tile_coordinates()[t].tile_y},
                      static_cast<std::uint32_t>(t)});
  std::sort(coords.begin(), coords.end());
  for (std::size_t t = 0; t + 1 < T; ++t) {
    const auto c = g.tile_coordinates()[t];
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (!dx && !dy) continue;
        const std::pair<std::int64_t, std::int64_t> want{c.tile_x + dx, c.tile_y + dy};
        auto it = std::lower_bound(coords.begin(), coords.end(),
                                   std::pair{want, std::uint32_t{0}});
        if (it != coords.end() && it->first == want) nbr[t].push_back(TileId{it->second});
      }
  }
  const TileId spill{static_cast<std::uint32_t>(T - 1)};
  for (std::uint32_t s = 0; s < g.vertex_count(); ++s) {
    const TileId st = owner[s];
    for (auto p = out.row_offsets[s]; p < out.row_offsets[s + 1]; ++p) {
      const VertexId d = out.destinations[p];
      const TileId dt = owner[d.value()];
      if (st == dt) {
        internal[st.value()].push_back(out.edge_ids[p]);
        continue;
      }
      outgoing[st.value()].push_back(out.edge_ids[p]);
      halo[st.value()].push_back(d);
      halo[dt.value()].push_back(VertexId{s});
      if (st == spill || dt == spill) {
        nbr[st.value()].push_back(dt);
        nbr[dt.value()].push_back(st);
      }
    }
  }
  for (std::uint32_t d = 0; d < g.vertex_count(); ++d)
    for (auto p = in.column_offsets[d]; p < in.column_offsets[d + 1]; ++p)
      if (owner[in.sources[p].value()] != owner[d]) incoming[owner[d].value()].push_back(in.edge_ids[p]);
  TileDirectory dir;
  dir.built_from_ = g.memo_handle();
  dir.spill_tile_ = spill;
  auto flatten = [](auto& groups, auto& offsets, auto& values) {
    offsets.assign(1, 0);
    for (auto& grp : groups) {
      values.insert(values.end(), grp.begin(), grp.end());
      offsets.push_back(values.size());
    }
  };
  for (auto& grp : nbr) {
    std::sort(grp.begin(), grp.end());
    grp.erase(std::unique(grp.begin(), grp.end()), grp.end());
  }
  for (auto& grp : halo) {
    std::sort(grp.begin(), grp.end(), [&](VertexId a, VertexId b) {
      return std::pair{owner[a.value()], a} < std::pair{owner[b.value()], b};
    });
    grp.erase(std::unique(grp.begin(), grp.end()), grp.end());
  }
  flatten(nbr, dir.neighbor_tile_offsets_, dir.neighbor_tiles_);
  flatten(internal, dir.internal_edge_offsets_, dir.internal_edge_ids_);
  flatten(outgoing, dir.outgoing_cross_edge_offsets_, dir.outgoing_cross_edge_ids_);
  flatten(incoming, dir.incoming_cross_edge_offsets_, dir.incoming_cross_edge_ids_);
  flatten(halo, dir.halo_vertex_offsets_, dir.halo_vertices_);
  return dir;
}

UniformGridPartitioner::UniformGridPartitioner(SpatialOrderConfig config,
                                               LocalityKeyPolicy locality_policy)
    : config_{config}, locality_policy_{std::move(locality_policy)} {
  if (config_.tile_width == 0U || config_.tile_height == 0U)
    throw std::invalid_argument{"tile width and height must both be positive"};
  if (!locality_policy_) throw std::invalid_argument{"locality-key policy must be callable"};
}

PartitionedGraph UniformGridPartitioner::partition(const InputGraph& input) const {
  WeightedGraph graph = build_spatially_ordered_graph(input, config_, locality_policy_);
  TileDirectory directory = build_tile_directory(graph);
  return PartitionedGraph{std::move(graph), std::move(directory)};
}

bool TileDirectory::built_from(const WeightedGraph& graph) const noexcept {
  const auto other = graph.memo_handle();
  return !built_from_.expired() && !built_from_.owner_before(other) && !other.owner_before(built_from_);
}

const std::string& weighted_graph_validation(const WeightedGraph& graph) {
  WeightedGraphMemo& memo = graph.memo();
  std::call_once(memo.validation_once, [&] {
    const char* why = validate_weighted_graph_message(graph);
    memo.validation_message = why ? why : "";
  });
  return memo.validation_message;
}

}  // namespace bfnew
