// synthetic.cpp -- direct O(E) construction of the synthetic tile grids.
//
// The builder never sorts: with tile-major ids the reference's orders are
// known in advance.  Within a source every edge into v weighs cost[v], so the
// canonical key (source, destination, weight, provenance) orders a row by
// (destination, input index j); that is also the CSR order (source,
// destination tile, destination, EdgeId) because the owner tile grows with
// the vertex id.  EdgeIds are therefore CSR positions.  The CSC is a
// counting sort by destination that keeps (source, EdgeId) order.
#include "bfnew/synthetic.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace bfnew {

namespace {

constexpr std::uint64_t kGamma = 0x9E3779B97F4A7C15ULL;
constexpr std::array<int, 6> kOffsets{+1, -1, +2, -2, +6, -6};

template <typename F>
void parallel_for(std::uint64_t n, unsigned threads, F&& body) {
  if (threads <= 1 || n < 4096) {
    body(0ULL, n);
    return;
  }
  std::vector<std::thread> pool;
  const std::uint64_t chunk = (n + threads - 1) / threads;
  for (unsigned t = 0; t < threads; ++t) {
    const std::uint64_t lo = std::min<std::uint64_t>(n, t * chunk);
    const std::uint64_t hi = std::min<std::uint64_t>(n, lo + chunk);
    if (lo < hi) pool.emplace_back([&body, lo, hi] { body(lo, hi); });
  }
  for (auto& th : pool) th.join();
}

}  // namespace

std::uint64_t splitmix64_at(std::uint64_t seed, std::uint64_t index) noexcept {
  std::uint64_t z = seed + index * kGamma;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

SyntheticTileGraph SyntheticTileGraph::generate(const SyntheticTileGridConfig& c, unsigned threads) {
  if (c.tiles_x == 0 || c.tiles_y == 0 || c.nodes_per_tile == 0)
    throw std::invalid_argument{"synthetic grid dimensions must be positive"};
  if (!(c.cost_min >= 0.0F) || !(c.cost_max >= c.cost_min) || !std::isfinite(c.cost_max))
    throw std::invalid_argument{"synthetic costs must be finite, nonnegative and ordered"};
  const std::uint64_t V64 = static_cast<std::uint64_t>(c.tiles_x) * c.tiles_y * c.nodes_per_tile;
  const std::uint32_t degree = c.intra_edges + static_cast<std::uint32_t>(kOffsets.size());
  if (V64 * degree >= 0xFFFFFFFFULL) throw std::overflow_error{"synthetic graph exceeds 32-bit edges"};
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  const auto V = static_cast<std::uint32_t>(V64);
  const std::uint32_t E = V * degree;
  const std::uint32_t npt = c.nodes_per_tile;
  const std::uint32_t grid_tiles = c.tiles_x * c.tiles_y;

  SyntheticTileGraph g;
  g.config_ = c;
  DeviceGraphLayout32& L = g.layout_;
  L.vertex_count = V;
  L.edge_count = E;
  L.tile_count = grid_tiles + 1;
  g.tile_offsets_.resize(L.tile_count + 1);
  for (std::uint32_t t = 0; t <= grid_tiles; ++t) g.tile_offsets_[t] = t * npt;
  g.tile_offsets_[L.tile_count] = V;  // empty spill tile
  g.costs_.resize(V);
  L.owner_tiles.resize(V);
  const float span = (c.cost_max - c.cost_min) / 16777216.0F;
  parallel_for(V, threads, [&](std::uint64_t lo, std::uint64_t hi) {
    for (std::uint64_t v = lo; v < hi; ++v) {
      const std::uint64_t r = splitmix64_at(c.seed, 64 * v + 1);
      const float u24 = static_cast<float>(r >> 40);
      const float scaled = u24 * span;
      g.costs_[v] = c.cost_min + scaled;
      L.owner_tiles[v] = static_cast<std::uint32_t>(v / npt);
    }
  });

  // CSR: fixed degree, each row sorted by (destination, input index).
  L.csr_row_offsets.resize(V + 1ULL);
  L.csr_destinations.resize(E);
  L.csr_weights.resize(E);
  parallel_for(V, threads, [&](std::uint64_t lo, std::uint64_t hi) {
    std::vector<std::pair<std::uint32_t, std::uint32_t>> row(degree);
    for (std::uint64_t v = lo; v < hi; ++v) {
      const std::uint32_t tile = static_cast<std::uint32_t>(v / npt);
      const std::uint32_t local = static_cast<std::uint32_t>(v % npt);
      const std::uint32_t tx = tile % c.tiles_x, ty = tile / c.tiles_x;
      std::uint32_t j = 0;
      for (std::uint32_t i = 0; i < c.intra_edges; ++i, ++j) {
        const std::uint64_t r = splitmix64_at(c.seed, 64 * v + 2 + j);
        std::uint32_t d = local;
        if (npt > 1) {
          d = static_cast<std::uint32_t>(r % (npt - 1));
          if (d >= local) ++d;
        }
        row[j] = {tile * npt + d, j};
      }
      for (std::size_t o = 0; o < kOffsets.size(); ++o, ++j) {
        const std::uint64_t r = splitmix64_at(c.seed, 64 * v + 2 + j);
        const bool along_y = (r >> 63) != 0;
        const std::uint32_t d = static_cast<std::uint32_t>(r & 0xFFFFFFFFULL) % npt;
        std::int64_t nx = tx, ny = ty;
        if (along_y) ny = std::clamp<std::int64_t>(ny + kOffsets[o], 0, c.tiles_y - 1);
        else nx = std::clamp<std::int64_t>(nx + kOffsets[o], 0, c.tiles_x - 1);
        row[j] = {static_cast<std::uint32_t>(ny * c.tiles_x + nx) * npt + d, j};
      }
      std::sort(row.begin(), row.end());
      const std::uint64_t base = v * degree;
      L.csr_row_offsets[v] = static_cast<std::uint32_t>(base);
      for (std::uint32_t k = 0; k < degree; ++k) {
        L.csr_destinations[base + k] = row[k].first;
        L.csr_weights[base + k] = g.costs_[row[k].first];
      }
    }
  });
  L.csr_row_offsets[V] = E;

  // CSC by counting sort; destination ranges per thread keep (source, id) order.
  std::vector<std::uint32_t> in_degree(V + 1ULL, 0);
  for (std::uint32_t e = 0; e < E; ++e) ++in_degree[L.csr_destinations[e] + 1ULL];
  L.csc_column_offsets.resize(V + 1ULL);
  L.csc_column_offsets[0] = 0;
  for (std::uint32_t v = 0; v < V; ++v) L.csc_column_offsets[v + 1] = L.csc_column_offsets[v] + in_degree[v + 1];
  L.csc_sources.resize(E);
  L.csc_weights.resize(E);
  L.csc_edge_ids.resize(E);
  parallel_for(V, threads, [&](std::uint64_t lo, std::uint64_t hi) {
    std::vector<std::uint32_t> cursor(hi - lo);
    for (std::uint64_t v = lo; v < hi; ++v) cursor[v - lo] = L.csc_column_offsets[v];
    for (std::uint32_t s = 0; s < V; ++s) {
      for (std::uint32_t e = L.csr_row_offsets[s]; e < L.csr_row_offsets[s + 1]; ++e) {
        const std::uint32_t d = L.csr_destinations[e];
        if (d < lo || d >= hi) continue;
        const std::uint32_t p = cursor[d - lo]++;
        L.csc_sources[p] = s;
        L.csc_weights[p] = L.csr_weights[e];
        L.csc_edge_ids[p] = e;
      }
    }
  });

  // Tile runs per row / column.
  auto runs = [&](const std::vector<std::uint32_t>& offsets, const std::vector<std::uint32_t>& ends,
                  std::vector<std::uint32_t>& row_runs, std::vector<std::uint32_t>& run_edges,
                  std::vector<std::uint32_t>& run_tiles) {
    std::vector<std::uint32_t> count(V + 1ULL, 0);
    parallel_for(V, threads, [&](std::uint64_t lo, std::uint64_t hi) {
      for (std::uint64_t v = lo; v < hi; ++v) {
        std::uint32_t n = 0;
        for (std::uint32_t p = offsets[v]; p < offsets[v + 1]; ++p)
          if (p == offsets[v] || ends[p] / npt != ends[p - 1] / npt) ++n;
        count[v + 1] = n;
      }
    });
    row_runs.resize(V + 1ULL);
    row_runs[0] = 0;
    for (std::uint32_t v = 0; v < V; ++v) row_runs[v + 1] = row_runs[v] + count[v + 1];
    const std::uint32_t R = row_runs[V];
    run_edges.resize(R + 1ULL);
    run_tiles.resize(R);
    parallel_for(V, threads, [&](std::uint64_t lo, std::uint64_t hi) {
      for (std::uint64_t v = lo; v < hi; ++v) {
        std::uint32_t r = row_runs[v];
        for (std::uint32_t p = offsets[v]; p < offsets[v + 1]; ++p) {
          if (p == offsets[v] || ends[p] / npt != ends[p - 1] / npt) {
            if (r > row_runs[v]) run_edges[r] = p;
            run_tiles[r] = ends[p] / npt;
            ++r;
          }
        }
        if (r > row_runs[v]) run_edges[r] = offsets[v + 1];
      }
    });
    run_edges[0] = 0;
  };
  runs(L.csr_row_offsets, L.csr_destinations, L.csr_row_run_offsets, L.csr_run_edge_offsets,
       L.csr_run_destination_tiles);
  runs(L.csc_column_offsets, L.csc_sources, L.csc_column_run_offsets, L.csc_run_edge_offsets,
       L.csc_run_source_tiles);
  return g;
}

InputGraph SyntheticTileGraph::input_graph() const {
  const auto& c = config_;
  const std::uint32_t V = layout_.vertex_count;
  std::vector<VertexMetadata> vertices(V);
  for (std::uint32_t v = 0; v < V; ++v) {
    const std::uint32_t tile = v / c.nodes_per_tile;
    vertices[v] = VertexMetadata::located(static_cast<std::int32_t>(tile % c.tiles_x),
                                          static_cast<std::int32_t>(tile / c.tiles_x),
                                          ResourceClassId{1});
  }
  // Input order: node v's edges j = 0..degree-1 as drawn (record 10 v + j).
  const std::uint32_t degree = c.intra_edges + 6;
  std::vector<EdgeInputRecord> edges(layout_.edge_count);
  const std::uint32_t npt = c.nodes_per_tile;
  for (std::uint32_t v = 0; v < V; ++v) {
    const std::uint32_t tile = v / npt, local = v % npt;
    const std::uint32_t tx = tile % c.tiles_x, ty = tile / c.tiles_x;
    for (std::uint32_t j = 0; j < degree; ++j) {
      const std::uint64_t r = splitmix64_at(c.seed, 64ULL * v + 2 + j);
      std::uint32_t d;
      if (j < c.intra_edges) {
        d = local;
        if (npt > 1) {
          d = static_cast<std::uint32_t>(r % (npt - 1));
          if (d >= local) ++d;
        }
        d += tile * npt;
      } else {
        const int off = kOffsets[j - c.intra_edges];
        std::int64_t nx = tx, ny = ty;
        if ((r >> 63) != 0) ny = std::clamp<std::int64_t>(ny + off, 0, c.tiles_y - 1);
        else nx = std::clamp<std::int64_t>(nx + off, 0, c.tiles_x - 1);
        d = static_cast<std::uint32_t>(ny * c.tiles_x + nx) * npt +
            static_cast<std::uint32_t>(r & 0xFFFFFFFFULL) % npt;
      }
      edges[static_cast<std::size_t>(v) * degree + j] = EdgeInputRecord{
          VertexId{v}, VertexId{d}, costs_[d],
          PhysicalProvenance{provenance_domain::synthetic, provenance_kind::synthetic_edge,
                             10ULL * v + j}};
    }
  }
  return InputGraph{std::move(vertices), std::move(edges)};
}

WeightedGraph SyntheticTileGraph::weighted_graph() const {
  const auto& L = layout_;
  const auto& c = config_;
  WeightedGraph g;
  g.vertex_count_ = L.vertex_count;
  g.edge_count_ = L.edge_count;
  g.spatial_ = true;
  g.config_ = SpatialOrderConfig{0, 0, 1, 1};
  g.vertices_.resize(L.vertex_count);
  g.old_to_new_.resize(L.vertex_count);
  g.new_to_old_.resize(L.vertex_count);
  g.owner_tiles_.resize(L.vertex_count);
  for (std::uint32_t v = 0; v < L.vertex_count; ++v) {
    const std::uint32_t tile = v / c.nodes_per_tile;
    g.vertices_[v] = VertexMetadata::located(static_cast<std::int32_t>(tile % c.tiles_x),
                                             static_cast<std::int32_t>(tile / c.tiles_x),
                                             ResourceClassId{1});
    g.old_to_new_[v] = g.new_to_old_[v] = VertexId{v};
    g.owner_tiles_[v] = TileId{L.owner_tiles[v]};
  }
  for (std::uint32_t ty = 0; ty < c.tiles_y; ++ty)
    for (std::uint32_t tx = 0; tx < c.tiles_x; ++tx)
      g.tile_coordinates_.push_back(TileCoordinate::located(tx, ty));
  g.tile_coordinates_.push_back(TileCoordinate::spill());
  g.tile_vertex_offsets_.assign(tile_offsets_.begin(), tile_offsets_.end());
  g.out_offsets_.assign(L.csr_row_offsets.begin(), L.csr_row_offsets.end());
  g.out_destinations_.resize(L.edge_count);
  g.out_edge_ids_.resize(L.edge_count);
  g.edge_provenance_.resize(L.edge_count);
  const std::uint32_t degree = c.intra_edges + 6;
  for (std::uint32_t e = 0; e < L.edge_count; ++e) {
    g.out_destinations_[e] = VertexId{L.csr_destinations[e]};
    g.out_edge_ids_[e] = EdgeId{e};
  }
  // Provenance by EdgeId: recover each row's input index order.
  for (std::uint32_t v = 0; v < L.vertex_count; ++v) {
    std::vector<std::pair<std::uint32_t, std::uint32_t>> row;
    const std::uint32_t tile = v / c.nodes_per_tile, local = v % c.nodes_per_tile;
    const std::uint32_t tx = tile % c.tiles_x, ty = tile / c.tiles_x;
    for (std::uint32_t j = 0; j < degree; ++j) {
      const std::uint64_t r = splitmix64_at(c.seed, 64ULL * v + 2 + j);
      std::uint32_t d;
      if (j < c.intra_edges) {
        d = local;
        if (c.nodes_per_tile > 1) {
          d = static_cast<std::uint32_t>(r % (c.nodes_per_tile - 1));
          if (d >= local) ++d;
        }
        d += tile * c.nodes_per_tile;
      } else {
        const int off = kOffsets[j - c.intra_edges];
        std::int64_t nx = tx, ny = ty;
        if ((r >> 63) != 0) ny = std::clamp<std::int64_t>(ny + off, 0, c.tiles_y - 1);
        else nx = std::clamp<std::int64_t>(nx + off, 0, c.tiles_x - 1);
        d = static_cast<std::uint32_t>(ny * c.tiles_x + nx) * c.nodes_per_tile +
            static_cast<std::uint32_t>(r & 0xFFFFFFFFULL) % c.nodes_per_tile;
      }
      row.emplace_back(d, j);
    }
    std::sort(row.begin(), row.end());
    for (std::uint32_t k = 0; k < degree; ++k)
      g.edge_provenance_[static_cast<std::size_t>(v) * degree + k] = PhysicalProvenance{
          provenance_domain::synthetic, provenance_kind::synthetic_edge, 10ULL * v + row[k].second};
  }
  g.out_weights_.assign(L.csr_weights.begin(), L.csr_weights.end());
  g.in_offsets_.assign(L.csc_column_offsets.begin(), L.csc_column_offsets.end());
  g.in_sources_.resize(L.edge_count);
  g.in_edge_ids_.resize(L.edge_count);
  for (std::uint32_t e = 0; e < L.edge_count; ++e) {
    g.in_sources_[e] = VertexId{L.csc_sources[e]};
    g.in_edge_ids_[e] = EdgeId{L.csc_edge_ids[e]};
  }
  g.in_weights_.assign(L.csc_weights.begin(), L.csc_weights.end());
  return g;
}

std::vector<SyntheticQuery> generate_synthetic_queries(const SyntheticTileGridConfig& grid,
                                                       const SyntheticQueryConfig& qc) {
  if (qc.side_min == 0 || qc.side_max < qc.side_min)
    throw std::invalid_argument{"invalid query box side range"};
  const std::uint32_t limit = std::min(grid.tiles_x, grid.tiles_y);
  const std::uint32_t npt = grid.nodes_per_tile;
  std::vector<SyntheticQuery> out(qc.count);
  auto unit = [](std::uint64_t r) { return static_cast<double>(r >> 11) * 0x1.0p-53; };
  for (std::uint32_t q = 0; q < qc.count; ++q) {
    auto draw = [&](std::uint32_t i) { return splitmix64_at(qc.seed, 64ULL * q + 1 + i); };
    SyntheticQuery& s = out[q];
    s.query_id = q;
    std::uint32_t lo = qc.side_min, hi = qc.side_max;
    double unbounded_p = qc.unbounded_fraction;
    if (qc.mix == SyntheticBoxMix::serving) {
      const double cls = unit(draw(0));
      if (cls < 0.70) { lo = 4; hi = 8; unbounded_p = 0.0; }
      else if (cls < 0.95) { lo = 16; hi = 32; unbounded_p = 0.0; }
      else { lo = 64; hi = 64; unbounded_p = 0.5; }
    }
    s.unbounded = unit(draw(1)) < unbounded_p;
    const double u = unit(draw(2));
    const double side_f = static_cast<double>(lo) * std::pow(static_cast<double>(hi) / lo, u);
    std::uint32_t side = static_cast<std::uint32_t>(std::floor(side_f + 0.5));
    side = std::clamp(side, lo, hi);
    side = std::min(side, limit);
    s.box_side = side;
    s.box_x = static_cast<std::uint32_t>(draw(3) % (grid.tiles_x - side + 1));
    s.box_y = static_cast<std::uint32_t>(draw(4) % (grid.tiles_y - side + 1));
    auto pick = [&](std::uint32_t i) {
      const std::uint32_t tx = s.box_x + static_cast<std::uint32_t>(draw(i) % side);
      const std::uint32_t ty = s.box_y + static_cast<std::uint32_t>(draw(i + 1) % side);
      const std::uint32_t local = static_cast<std::uint32_t>(draw(i + 2) % npt);
      return (ty * grid.tiles_x + tx) * npt + local;
    };
    s.source = pick(5);
    s.target = pick(8);
    for (std::uint32_t retry = 0; s.target == s.source && retry < 16; ++retry)
      s.target = pick(11 + 3 * retry);
  }
  return out;
}

std::vector<TileId> synthetic_query_tiles(const SyntheticTileGridConfig& grid,
                                          const SyntheticQuery& q) {
  std::vector<TileId> tiles;
  if (q.unbounded) {
    const std::uint32_t T = grid.tiles_x * grid.tiles_y + 1;
    tiles.reserve(T);
    for (std::uint32_t t = 0; t < T; ++t) tiles.push_back(TileId{t});
    return tiles;
  }
  tiles.reserve(static_cast<std::size_t>(q.box_side) * q.box_side);
  for (std::uint32_t y = q.box_y; y < q.box_y + q.box_side; ++y)
    for (std::uint32_t x = q.box_x; x < q.box_x + q.box_side; ++x)
      tiles.push_back(TileId{y * grid.tiles_x + x});
  return tiles;
}

}  // namespace bfnew
