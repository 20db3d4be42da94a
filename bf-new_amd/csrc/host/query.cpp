// query.cpp -- route queries (reference behaviour: query.cpp:28-230).
#include "bfnew/query.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>

namespace bfnew {

namespace {

void canonicalize(std::span<const VertexId> terminals, std::vector<VertexId>& canonical,
                  std::vector<std::uint32_t>& map) {
  canonical.assign(terminals.begin(), terminals.end());
  std::sort(canonical.begin(), canonical.end());
  canonical.erase(std::unique(canonical.begin(), canonical.end()), canonical.end());
  map.clear();
  for (const VertexId t : terminals)
    map.push_back(static_cast<std::uint32_t>(
        std::lower_bound(canonical.begin(), canonical.end(), t) - canonical.begin()));
}

std::int64_t sat_sub(std::int64_t v, std::uint32_t a) {
  return v < std::numeric_limits<std::int64_t>::min() + static_cast<std::int64_t>(a)
             ? std::numeric_limits<std::int64_t>::min()
             : v - a;
}
std::int64_t sat_add(std::int64_t v, std::uint32_t a) {
  return v > std::numeric_limits<std::int64_t>::max() - static_cast<std::int64_t>(a)
             ? std::numeric_limits<std::int64_t>::max()
             : v + a;
}

template <typename T>
bool strictly_increasing(const std::vector<T>& v) {
  for (std::size_t i = 1; i < v.size(); ++i)
    if (!(v[i - 1] < v[i])) return false;
  return true;
}

}  // namespace

std::vector<TileId> select_route_tiles(const WeightedGraph& graph,
                                       std::span<const VertexId> terminals,
                                       std::uint32_t padding) {
  bool located = false, spill = false;
  std::int64_t x0 = std::numeric_limits<std::int64_t>::max(), y0 = x0;
  std::int64_t x1 = std::numeric_limits<std::int64_t>::min(), y1 = x1;
  const auto coords = graph.tile_coordinates();
  for (const VertexId t : terminals) {
    if (t.value() >= graph.vertex_count()) throw std::out_of_range{"query terminal is outside the graph"};
    const TileCoordinate c = coords[graph.owner_tiles()[t.value()].value()];
    if (!c.has_location) {
      spill = true;
      continue;
    }
    located = true;
    x0 = std::min(x0, c.tile_x);
    y0 = std::min(y0, c.tile_y);
    x1 = std::max(x1, c.tile_x);
    y1 = std::max(y1, c.tile_y);
  }
  std::vector<TileId> out;
  if (located) {
    const std::int64_t lx = sat_sub(x0, padding), ly = sat_sub(y0, padding);
    const std::int64_t ux = sat_add(x1, padding), uy = sat_add(y1, padding);
    for (std::size_t t = 0; t < coords.size(); ++t) {
      const TileCoordinate& c = coords[t];
      if (c.has_location && c.tile_x >= lx && c.tile_x <= ux && c.tile_y >= ly && c.tile_y <= uy)
        out.push_back(TileId{static_cast<std::uint32_t>(t)});
    }
  }
  if (spill) out.push_back(TileId{static_cast<std::uint32_t>(coords.size() - 1)});
  return out;
}

RouteQuery make_route_query(QueryId query_id, const WeightedGraph& graph,
                            std::span<const VertexId> source_terminals,
                            std::span<const VertexId> target_terminals,
                            std::uint32_t tile_padding, std::uint32_t expansion_generation) {
  if (!graph.has_spatial_ordering())
    throw std::invalid_argument{"route queries require a spatially ordered graph"};
  if (source_terminals.empty() || target_terminals.empty())
    throw std::invalid_argument{"route queries require sources and targets"};
  RouteQuery q;
  q.query_id = query_id;
  q.source_terminals.assign(source_terminals.begin(), source_terminals.end());
  q.target_terminals.assign(target_terminals.begin(), target_terminals.end());
  canonicalize(source_terminals, q.sources, q.source_terminal_to_source);
  canonicalize(target_terminals, q.targets, q.target_terminal_to_target);
  std::vector<VertexId> all(q.source_terminals);
  all.insert(all.end(), q.target_terminals.begin(), q.target_terminals.end());
  q.selected_tiles = select_route_tiles(graph, all, tile_padding);
  q.expansion_generation = expansion_generation;
  if (!validate_route_query(graph, q).ok())
    throw std::logic_error{"constructed route query failed deep validation"};
  return q;
}

RouteQueryValidationResult validate_route_query(const WeightedGraph& graph,
                                                const RouteQuery& q) noexcept {
  using C = RouteQueryValidationErrorCode;
  auto err = [](C c, std::size_t p = RouteQueryValidationResult::no_position) {
    return RouteQueryValidationResult{c, p};
  };
  if (!graph.has_spatial_ordering()) return err(C::graph_is_not_spatially_ordered);
  if (q.sources.empty() || q.source_terminals.empty()) return err(C::empty_sources);
  if (q.targets.empty() || q.target_terminals.empty()) return err(C::empty_targets);
  for (std::size_t i = 0; i < q.sources.size(); ++i)
    if (q.sources[i].value() >= graph.vertex_count()) return err(C::source_out_of_range, i);
  for (std::size_t i = 0; i < q.targets.size(); ++i)
    if (q.targets[i].value() >= graph.vertex_count()) return err(C::target_out_of_range, i);
  if (!strictly_increasing(q.sources)) return err(C::noncanonical_sources);
  if (!strictly_increasing(q.targets)) return err(C::noncanonical_targets);
  if (q.source_terminal_to_source.size() != q.source_terminals.size()) return err(C::source_map_size_mismatch);
  if (q.target_terminal_to_target.size() != q.target_terminals.size()) return err(C::target_map_size_mismatch);
  for (std::size_t i = 0; i < q.source_terminals.size(); ++i) {
    const std::uint32_t m = q.source_terminal_to_source[i];
    if (m >= q.sources.size() || q.sources[m] != q.source_terminals[i]) return err(C::source_map_mismatch, i);
  }
  for (std::size_t i = 0; i < q.target_terminals.size(); ++i) {
    const std::uint32_t m = q.target_terminal_to_target[i];
    if (m >= q.targets.size() || q.targets[m] != q.target_terminals[i]) return err(C::target_map_mismatch, i);
  }
  if (q.selected_tiles.empty()) return err(C::empty_selected_tiles);
  if (!strictly_increasing(q.selected_tiles)) return err(C::noncanonical_selected_tiles);
  for (std::size_t i = 0; i < q.selected_tiles.size(); ++i)
    if (q.selected_tiles[i].value() >= graph.tile_coordinates().size())
      return err(C::selected_tile_out_of_range, i);
  auto owns = [&](VertexId t) {
    return t.value() < graph.vertex_count() &&
           std::binary_search(q.selected_tiles.begin(), q.selected_tiles.end(),
                              graph.owner_tiles()[t.value()]);
  };
  for (std::size_t i = 0; i < q.source_terminals.size(); ++i)
    if (!owns(q.source_terminals[i])) return err(C::terminal_owner_tile_not_selected, i);
  for (std::size_t i = 0; i < q.target_terminals.size(); ++i)
    if (!owns(q.target_terminals[i])) return err(C::terminal_owner_tile_not_selected, i);
  return {};
}

std::uint64_t estimate_selected_vertex_count(const WeightedGraph& graph,
                                             std::span<const TileId> tiles) {
  std::uint64_t n = 0;
  const auto off = graph.tile_vertex_offsets();
  for (const TileId t : tiles) {
    if (t.value() >= graph.tile_coordinates().size()) throw std::out_of_range{"selected tile is outside the graph"};
    n += off[t.value() + 1] - off[t.value()];
  }
  return n;
}

std::uint64_t estimate_selected_edge_count(const WeightedGraph& graph,
                                           std::span<const TileId> tiles) {
  std::vector<bool> sel(graph.tile_coordinates().size(), false);
  for (const TileId t : tiles) {
    if (t.value() >= sel.size()) throw std::out_of_range{"selected tile is outside the graph"};
    sel[t.value()] = true;
  }
  std::uint64_t n = 0;
  const auto out = graph.outgoing();
  const auto owner = graph.owner_tiles();
  for (std::uint32_t s = 0; s < graph.vertex_count(); ++s) {
    if (!sel[owner[s].value()]) continue;
    for (auto p = out.row_offsets[s]; p < out.row_offsets[s + 1]; ++p)
      n += sel[owner[out.destinations[p].value()].value()];
  }
  return n;
}

}  // namespace bfnew
