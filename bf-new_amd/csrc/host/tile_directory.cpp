// tile_directory.cpp -- deep check of a TileDirectory against its graph
// (reference behaviour: src/tiled_graph.cpp:293-484, error codes
// include/bfnew/spatial.hpp:126-151).
//
// Expectations are derived from the graph in one pass over the CSR: per
// stable edge id its source, destination and CSR rank; per CSC position the
// edge id gives its CSC rank.  The directory's groups are then checked
// against those tables, so each error names the first offending position.
#include "bfnew/graph.hpp"
#include "bfnew/spatial.hpp"

#include <algorithm>
#include <stdexcept>
#include <tuple>

namespace bfnew {
namespace {

using Code = TileDirectoryValidationErrorCode;

TileDirectoryValidationResult error(Code c, EdgeOffset at = TileDirectoryValidationResult::no_position) {
  return TileDirectoryValidationResult{c, at};
}

bool offsets_ok(std::span<const EdgeOffset> off, std::size_t groups, std::size_t values) {
  if (off.size() != groups + 1U || off.front() != 0U || off.back() != values) return false;
  for (std::size_t i = 1; i < off.size(); ++i)
    if (off[i] < off[i - 1]) return false;
  return true;
}

struct EdgeTable {
  std::vector<std::uint32_t> src, dst;
  std::vector<std::uint64_t> csr_rank, csc_rank;
};

EdgeTable edge_table(const WeightedGraph& g) {
  const std::size_t E = g.edge_count();
  EdgeTable t;
  t.src.resize(E);
  t.dst.resize(E);
  t.csr_rank.resize(E);
  t.csc_rank.resize(E);
  const OutgoingCsrView out = g.outgoing();
  for (std::uint32_t s = 0; s < g.vertex_count(); ++s)
    for (auto p = out.row_offsets[s]; p < out.row_offsets[s + 1]; ++p) {
      const auto id = out.edge_ids[p].value();
      t.src[id] = s;
      t.dst[id] = out.destinations[p].value();
      t.csr_rank[id] = p;
    }
  const IncomingCscView in = g.incoming();
  for (std::uint64_t p = 0; p < E; ++p) t.csc_rank[in.edge_ids[p].value()] = p;
  return t;
}

}  // namespace

TileDirectoryValidationResult validate_tile_directory(const WeightedGraph& g, const TileDirectory& dir) {
  if (!weighted_graph_validation(g).empty()) return error(Code::graph_validation_failed);
  if (!g.has_spatial_ordering()) return error(Code::graph_is_not_spatially_ordered);
  const std::size_t T = g.tile_coordinates().size();
  const std::size_t E = g.edge_count();
  if (dir.tile_count() != T) return error(Code::tile_count_mismatch);
  if (dir.spill_tile().value() != T - 1U) return error(Code::spill_tile_mismatch);
  if (!offsets_ok(dir.neighbor_tile_offsets(), T, dir.neighbor_tiles().size()))
    return error(Code::invalid_neighbor_offsets);
  if (!offsets_ok(dir.internal_edge_offsets(), T, dir.internal_edge_ids().size()))
    return error(Code::invalid_internal_edge_offsets);
  if (!offsets_ok(dir.outgoing_cross_edge_offsets(), T, dir.outgoing_cross_edge_ids().size()))
    return error(Code::invalid_outgoing_cross_edge_offsets);
  if (!offsets_ok(dir.incoming_cross_edge_offsets(), T, dir.incoming_cross_edge_ids().size()))
    return error(Code::invalid_incoming_cross_edge_offsets);
  if (!offsets_ok(dir.halo_vertex_offsets(), T, dir.halo_vertices().size()))
    return error(Code::invalid_halo_offsets);

  const auto owner = g.owner_tiles();
  const EdgeTable et = edge_table(g);
  auto tile_of = [&](std::uint32_t v) { return owner[v].value(); };

  // Expected neighbour groups: located Chebyshev-1 tiles, plus both ends of
  // every edge that touches the spill tile.
  std::vector<std::vector<std::uint32_t>> want_nbr(T);
  {
    std::vector<std::tuple<std::int64_t, std::int64_t, std::uint32_t>> located;
    for (std::uint32_t t = 0; t + 1 < T; ++t)
      located.emplace_back(g.tile_coordinates()[t].tile_x, g.tile_coordinates()[t].tile_y, t);
    std::sort(located.begin(), located.end());
    for (std::uint32_t t = 0; t + 1 < T; ++t) {
      const TileCoordinate c = g.tile_coordinates()[t];
      for (std::int64_t dy = -1; dy <= 1; ++dy)
        for (std::int64_t dx = -1; dx <= 1; ++dx) {
          if (!dx && !dy) continue;
          const auto it = std::lower_bound(located.begin(), located.end(),
                                           std::tuple{c.tile_x + dx, c.tile_y + dy, std::uint32_t{0}});
          if (it != located.end() && std::get<0>(*it) == c.tile_x + dx && std::get<1>(*it) == c.tile_y + dy)
            want_nbr[t].push_back(std::get<2>(*it));
        }
    }
    const std::uint32_t spill = static_cast<std::uint32_t>(T - 1U);
    for (std::size_t e = 0; e < E; ++e) {
      const std::uint32_t a = tile_of(et.src[e]), b = tile_of(et.dst[e]);
      if (a != b && (a == spill || b == spill)) {
        want_nbr[a].push_back(b);
        want_nbr[b].push_back(a);
      }
    }
    for (auto& grp : want_nbr) {
      std::sort(grp.begin(), grp.end());
      grp.erase(std::unique(grp.begin(), grp.end()), grp.end());
    }
  }
  const auto nbr_off = dir.neighbor_tile_offsets();
  const auto nbr = dir.neighbor_tiles();
  for (std::size_t t = 0; t < T; ++t) {
    for (auto p = nbr_off[t]; p < nbr_off[t + 1]; ++p) {
      const std::uint32_t n = nbr[p].value();
      if (n >= T || n == t) return error(Code::neighbor_out_of_range, p);
      if (p > nbr_off[t] && !(nbr[p - 1] < nbr[p])) return error(Code::neighbor_order_violation, p);
      if (!std::binary_search(nbr.begin() + nbr_off[n], nbr.begin() + nbr_off[n + 1],
                              TileId{static_cast<std::uint32_t>(t)}))
        return error(Code::neighbor_asymmetry, p);
    }
    const auto& want = want_nbr[t];
    if (nbr_off[t + 1] - nbr_off[t] != want.size() ||
        !std::equal(want.begin(), want.end(), nbr.begin() + nbr_off[t],
                    [](std::uint32_t a, TileId b) { return a == b.value(); }))
      return error(Code::unexpected_neighbor, t);
  }

  // Edge groups: internal and outgoing by source tile in CSR order,
  // incoming by destination tile in CSC order; each edge exactly once.
  std::vector<std::uint8_t> by_source(E, 0U), by_destination(E, 0U);
  auto source_group = [&](std::span<const EdgeOffset> off, std::span<const EdgeId> ids,
                          bool internal) -> TileDirectoryValidationResult {
    for (std::size_t t = 0; t < T; ++t)
      for (auto p = off[t]; p < off[t + 1]; ++p) {
        const auto id = ids[p].value();
        if (id >= E) return error(Code::edge_id_out_of_range, p);
        const std::uint32_t a = tile_of(et.src[id]), b = tile_of(et.dst[id]);
        if (a != t || (a == b) != internal) return error(Code::edge_classification_mismatch, p);
        if (by_source[id]) return error(Code::duplicate_source_edge_classification, p);
        if (p > off[t] && et.csr_rank[id] <= et.csr_rank[ids[p - 1].value()])
          return error(Code::edge_metadata_order_violation, p);
        by_source[id] = 1U;
      }
    return {};
  };
  if (auto r = source_group(dir.internal_edge_offsets(), dir.internal_edge_ids(), true); !r.ok()) return r;
  if (auto r = source_group(dir.outgoing_cross_edge_offsets(), dir.outgoing_cross_edge_ids(), false); !r.ok())
    return r;
  {
    const auto off = dir.incoming_cross_edge_offsets();
    const auto ids = dir.incoming_cross_edge_ids();
    for (std::size_t t = 0; t < T; ++t)
      for (auto p = off[t]; p < off[t + 1]; ++p) {
        const auto id = ids[p].value();
        if (id >= E) return error(Code::edge_id_out_of_range, p);
        const std::uint32_t a = tile_of(et.src[id]), b = tile_of(et.dst[id]);
        if (b != t || a == b) return error(Code::edge_classification_mismatch, p);
        if (by_destination[id]) return error(Code::duplicate_incoming_cross_edge, p);
        if (p > off[t] && et.csc_rank[id] <= et.csc_rank[ids[p - 1].value()])
          return error(Code::edge_metadata_order_violation, p);
        by_destination[id] = 1U;
      }
  }
  for (std::size_t e = 0; e < E; ++e) {
    if (!by_source[e]) return error(Code::missing_source_edge_classification, e);
    const bool cross = tile_of(et.src[e]) != tile_of(et.dst[e]);
    if (cross && !by_destination[e]) return error(Code::missing_incoming_cross_edge, e);
  }

  // Halos: remote endpoints of each tile's cross edges, by (owner, vertex).
  std::vector<std::vector<std::uint32_t>> want_halo(T);
  for (std::size_t e = 0; e < E; ++e) {
    const std::uint32_t a = tile_of(et.src[e]), b = tile_of(et.dst[e]);
    if (a == b) continue;
    want_halo[a].push_back(et.dst[e]);
    want_halo[b].push_back(et.src[e]);
  }
  auto halo_less = [&](std::uint32_t x, std::uint32_t y) {
    return std::pair{tile_of(x), x} < std::pair{tile_of(y), y};
  };
  const auto h_off = dir.halo_vertex_offsets();
  const auto halo = dir.halo_vertices();
  for (std::size_t t = 0; t < T; ++t) {
    auto& want = want_halo[t];
    std::sort(want.begin(), want.end(), halo_less);
    want.erase(std::unique(want.begin(), want.end()), want.end());
    for (auto p = h_off[t]; p < h_off[t + 1]; ++p) {
      const std::uint32_t v = halo[p].value();
      if (v >= g.vertex_count()) return error(Code::halo_vertex_out_of_range, p);
      if (tile_of(v) == t) return error(Code::halo_owner_mismatch, p);
      if (p > h_off[t] && !halo_less(halo[p - 1].value(), v)) return error(Code::halo_order_violation, p);
    }
    if (h_off[t + 1] - h_off[t] != want.size() ||
        !std::equal(want.begin(), want.end(), halo.begin() + h_off[t],
                    [](std::uint32_t a, VertexId b) { return a == b.value(); }))
      return error(Code::halo_set_mismatch, t);
  }
  return {};
}

TileDirectory tile_directory_with_swapped_entries(const TileDirectory& directory, int which,
                                                  std::size_t a, std::size_t b) {
  TileDirectory copy = directory;
  auto swap_in = [&](auto& values) {
    if (a >= values.size() || b >= values.size()) throw std::out_of_range{"directory entry"};
    std::swap(values[a], values[b]);
  };
  switch (which) {
    case 0: swap_in(copy.neighbor_tiles_); break;
    case 1: swap_in(copy.internal_edge_ids_); break;
    case 2: swap_in(copy.outgoing_cross_edge_ids_); break;
    case 3: swap_in(copy.incoming_cross_edge_ids_); break;
    case 4: swap_in(copy.halo_vertices_); break;
    default: throw std::invalid_argument{"directory group"};
  }
  return copy;
}

}  // namespace bfnew
