// device_layout.cpp -- tile runs, the checked 32-bit image and its identity.
//
// Runs (device_layout.cpp:186-251 in the reference): scan each CSR row and
// CSC column once and cut a run wherever the remote endpoint's owner tile
// changes.  The fingerprint reproduces the reference's component identity
// (device_layout.cpp:68-104, 684-709): components in a fixed tag order, each
// introduced by (u32 tag, u64 element count) and followed by its u32 values,
// every integer fed little-endian byte by byte into an FNV-1a-style lane and a
// second multiplicative-mix lane.  It is kept bit-compatible so a resident
// image can be checked against a reference-built layout.
#include "bfnew/device_layout.hpp"

#include <bit>
#include <limits>
#include <stdexcept>

namespace bfnew {

namespace {

template <typename Endpoint>
void build_runs(std::uint32_t V, std::span<const EdgeOffset> offsets,
                std::span<const Endpoint> endpoints, std::span<const TileId> owner,
                std::vector<EdgeOffset>& row_runs, std::vector<EdgeOffset>& run_edges,
                std::vector<TileId>& run_tiles) {
  row_runs.assign(1, 0);
  run_edges.assign(1, 0);
  run_tiles.clear();
  row_runs.reserve(V + 1ULL);
  for (std::uint32_t v = 0; v < V; ++v) {
    EdgeOffset p = offsets[v];
    const EdgeOffset end = offsets[v + 1];
    while (p < end) {
      const TileId tile = owner[endpoints[p].value()];
      ++p;
      while (p < end && owner[endpoints[p].value()] == tile) ++p;
      run_tiles.push_back(tile);
      run_edges.push_back(p);
    }
    row_runs.push_back(run_tiles.size());
  }
}

class Fingerprint {
 public:
  void component(std::uint32_t tag, std::uint64_t count) {
    u32(tag);
    for (unsigned s = 0; s < 64; s += 8) byte(static_cast<std::uint8_t>(count >> s));
  }
  void u32(std::uint32_t v) {
    for (unsigned s = 0; s < 32; s += 8) byte(static_cast<std::uint8_t>(v >> s));
  }
  template <typename T>
  void values(std::uint32_t tag, const std::vector<T>& xs) {
    component(tag, xs.size());
    for (const T& x : xs) {
      if constexpr (std::is_same_v<T, float>) u32(std::bit_cast<std::uint32_t>(x));
      else u32(x);
    }
  }
  DeviceGraphFingerprint finish() const { return {a_, b_}; }

 private:
  void byte(std::uint8_t v) {
    a_ = (a_ ^ v) * 0x00000100000001B3ULL;
    b_ ^= static_cast<std::uint64_t>(v) + 0x9E3779B97F4A7C15ULL + (b_ << 6) + (b_ >> 2);
    b_ *= 0xD6E8FEB86659FD93ULL;
  }
  // The reference seeds this lane with 1469598103934665603, which is not the
  // standard FNV-1a basis (14695981039346656037); kept for identity parity.
  std::uint64_t a_ = 1469598103934665603ULL;
  std::uint64_t b_ = 0xA0761D6478BD642FULL;
};

template <typename T>
std::uint64_t bytes_of(const std::vector<T>& v) {
  return static_cast<std::uint64_t>(v.size()) * sizeof(T);
}

}  // namespace

TileRunLayout64 build_tile_run_layout(const WeightedGraph& graph) {
  if (!graph.has_spatial_ordering()) throw std::invalid_argument{"tile runs require a spatially ordered graph"};
  TileRunLayout64 r;
  const auto out = graph.outgoing();
  const auto in = graph.incoming();
  build_runs(graph.vertex_count(), out.row_offsets, out.destinations, graph.owner_tiles(),
             r.csr_row_run_offsets, r.csr_run_edge_offsets, r.csr_run_destination_tiles);
  build_runs(graph.vertex_count(), in.column_offsets, in.sources, graph.owner_tiles(),
             r.csc_column_run_offsets, r.csc_run_edge_offsets, r.csc_run_source_tiles);
  return r;
}

std::uint32_t checked_device_offset32(EdgeOffset offset) {
  if (offset > std::numeric_limits<std::uint32_t>::max())
    throw std::overflow_error{"64-bit graph offset does not fit the device layout"};
  return static_cast<std::uint32_t>(offset);
}

DeviceGraphLayout32 build_device_graph_layout32(const WeightedGraph& graph,
                                                const TileRunLayout64& runs) {
  constexpr std::uint64_t max32 = std::numeric_limits<std::uint32_t>::max();
  if (graph.edge_count() > max32 || graph.tile_coordinates().size() > max32 ||
      runs.csr_run_destination_tiles.size() > max32 || runs.csc_run_source_tiles.size() > max32)
    throw std::overflow_error{"graph counts do not fit the 32-bit device layout"};
  if (runs.csr_row_run_offsets.size() != graph.vertex_count() + 1ULL ||
      runs.csc_column_run_offsets.size() != graph.vertex_count() + 1ULL)
    throw std::invalid_argument{"device layout requires valid tile-run metadata"};
  DeviceGraphLayout32 L;
  L.vertex_count = graph.vertex_count();
  L.edge_count = static_cast<std::uint32_t>(graph.edge_count());
  L.tile_count = static_cast<std::uint32_t>(graph.tile_coordinates().size());
  auto narrow = [](auto span, std::vector<std::uint32_t>& dst) {
    dst.resize(span.size());
    for (std::size_t i = 0; i < span.size(); ++i) {
      if constexpr (std::is_integral_v<std::decay_t<decltype(span[i])>>)
        dst[i] = checked_device_offset32(span[i]);
      else
        dst[i] = checked_device_offset32(span[i].value());
    }
  };
  narrow(graph.owner_tiles(), L.owner_tiles);
  const auto out = graph.outgoing();
  const auto in = graph.incoming();
  narrow(out.row_offsets, L.csr_row_offsets);
  narrow(out.destinations, L.csr_destinations);
  L.csr_weights.assign(out.weights.begin(), out.weights.end());
  narrow(std::span<const EdgeOffset>(runs.csr_row_run_offsets), L.csr_row_run_offsets);
  narrow(std::span<const EdgeOffset>(runs.csr_run_edge_offsets), L.csr_run_edge_offsets);
  narrow(std::span<const TileId>(runs.csr_run_destination_tiles), L.csr_run_destination_tiles);
  narrow(in.column_offsets, L.csc_column_offsets);
  narrow(in.sources, L.csc_sources);
  L.csc_weights.assign(in.weights.begin(), in.weights.end());
  narrow(in.edge_ids, L.csc_edge_ids);
  narrow(std::span<const EdgeOffset>(runs.csc_column_run_offsets), L.csc_column_run_offsets);
  narrow(std::span<const EdgeOffset>(runs.csc_run_edge_offsets), L.csc_run_edge_offsets);
  narrow(std::span<const TileId>(runs.csc_run_source_tiles), L.csc_run_source_tiles);
  return L;
}

bool device_graph_layouts_deep_equal(const DeviceGraphLayout32& l, const DeviceGraphLayout32& r) {
  auto bits_equal = [](const std::vector<float>& a, const std::vector<float>& b) {
    if (a.size() != b.size()) return false;
    for (std::size_t i = 0; i < a.size(); ++i)
      if (std::bit_cast<std::uint32_t>(a[i]) != std::bit_cast<std::uint32_t>(b[i])) return false;
    return true;
  };
  return l.vertex_count == r.vertex_count && l.edge_count == r.edge_count &&
         l.tile_count == r.tile_count && l.owner_tiles == r.owner_tiles &&
         l.csr_row_offsets == r.csr_row_offsets && l.csr_destinations == r.csr_destinations &&
         bits_equal(l.csr_weights, r.csr_weights) && l.csr_row_run_offsets == r.csr_row_run_offsets &&
         l.csr_run_edge_offsets == r.csr_run_edge_offsets &&
         l.csr_run_destination_tiles == r.csr_run_destination_tiles &&
         l.csc_column_offsets == r.csc_column_offsets && l.csc_sources == r.csc_sources &&
         bits_equal(l.csc_weights, r.csc_weights) && l.csc_edge_ids == r.csc_edge_ids &&
         l.csc_column_run_offsets == r.csc_column_run_offsets &&
         l.csc_run_edge_offsets == r.csc_run_edge_offsets &&
         l.csc_run_source_tiles == r.csc_run_source_tiles;
}

DeviceGraphMemoryReport report_device_graph_memory(const DeviceGraphLayout32& L) noexcept {
  DeviceGraphMemoryReport m;
  m.owner_tiles_bytes = bytes_of(L.owner_tiles);
  m.csr_row_offsets_bytes = bytes_of(L.csr_row_offsets);
  m.csr_destinations_bytes = bytes_of(L.csr_destinations);
  m.csr_weights_bytes = bytes_of(L.csr_weights);
  m.csr_row_run_offsets_bytes = bytes_of(L.csr_row_run_offsets);
  m.csr_run_edge_offsets_bytes = bytes_of(L.csr_run_edge_offsets);
  m.csr_run_destination_tiles_bytes = bytes_of(L.csr_run_destination_tiles);
  m.csc_column_offsets_bytes = bytes_of(L.csc_column_offsets);
  m.csc_sources_bytes = bytes_of(L.csc_sources);
  m.csc_weights_bytes = bytes_of(L.csc_weights);
  m.csc_edge_ids_bytes = bytes_of(L.csc_edge_ids);
  m.csc_column_run_offsets_bytes = bytes_of(L.csc_column_run_offsets);
  m.csc_run_edge_offsets_bytes = bytes_of(L.csc_run_edge_offsets);
  m.csc_run_source_tiles_bytes = bytes_of(L.csc_run_source_tiles);
  m.total_bytes = m.owner_tiles_bytes + m.csr_row_offsets_bytes + m.csr_destinations_bytes +
                  m.csr_weights_bytes + m.csr_row_run_offsets_bytes + m.csr_run_edge_offsets_bytes +
                  m.csr_run_destination_tiles_bytes + m.csc_column_offsets_bytes +
                  m.csc_sources_bytes + m.csc_weights_bytes + m.csc_edge_ids_bytes +
                  m.csc_column_run_offsets_bytes + m.csc_run_edge_offsets_bytes +
                  m.csc_run_source_tiles_bytes;
  return m;
}

DeviceGraphFingerprint fingerprint_device_graph_layout32(const DeviceGraphLayout32& L) noexcept {
  Fingerprint f;
  f.component(0, 3);
  f.u32(L.vertex_count);
  f.u32(L.edge_count);
  f.u32(L.tile_count);
  f.values(1, L.owner_tiles);
  f.values(2, L.csr_row_offsets);
  f.values(3, L.csr_destinations);
  f.values(4, L.csr_weights);
  f.values(5, L.csr_row_run_offsets);
  f.values(6, L.csr_run_edge_offsets);
  f.values(7, L.csr_run_destination_tiles);
  f.values(8, L.csc_column_offsets);
  f.values(9, L.csc_sources);
  f.values(10, L.csc_weights);
  f.values(11, L.csc_column_run_offsets);
  f.values(12, L.csc_run_edge_offsets);
  f.values(13, L.csc_run_source_tiles);
  f.values(14, L.csc_edge_ids);
  return f.finish();
}

const TileRunLayout64& graph_tile_runs(const WeightedGraph& graph) {
  WeightedGraphMemo& memo = graph.memo();
  std::call_once(memo.tile_runs_once,
                 [&] { memo.tile_runs = std::make_shared<const TileRunLayout64>(build_tile_run_layout(graph)); });
  return *static_cast<const TileRunLayout64*>(memo.tile_runs.get());
}

DeviceGraphFingerprint fingerprint_device_graph_source32(const WeightedGraph& graph,
                                                         const TileRunLayout64& tile_runs) {
  // The tile runs are a function of the graph, so the graph's memo holds the
  // value once runs equal to the graph's own have been seen.  The graph's own
  // memoised runs need no element comparison (hundreds of MB on large graphs).
  const TileRunLayout64& own = graph_tile_runs(graph);
  if (&tile_runs != &own && tile_runs != own)
    throw std::invalid_argument{"device graph fingerprint: tile runs are not the graph's"};
  WeightedGraphMemo& memo = graph.memo();
  std::call_once(memo.device_fingerprint_once, [&] {
    const DeviceGraphFingerprint f = fingerprint_device_graph_layout32(build_device_graph_layout32(graph, tile_runs));
    memo.device_fingerprint[0] = f.first;
    memo.device_fingerprint[1] = f.second;
  });
  return DeviceGraphFingerprint{memo.device_fingerprint[0], memo.device_fingerprint[1]};
}

void compute_tile_run_lane_masks(const WeightedGraph& graph, const TileRunLayout64& runs,
                                 std::span<const std::uint32_t> tile_lane_masks,
                                 TileRunLaneMasks& output) {
  if (tile_lane_masks.size() != graph.tile_coordinates().size())
    throw std::invalid_argument{"tile lane masks or tile-run shape are invalid"};
  output.csr_run_masks.assign(runs.csr_run_destination_tiles.size(), 0U);
  output.csc_run_masks.assign(runs.csc_run_source_tiles.size(), 0U);
  const auto owner = graph.owner_tiles();
  for (std::uint32_t v = 0; v < graph.vertex_count(); ++v) {
    const std::uint32_t own = tile_lane_masks[owner[v].value()];
    for (auto r = runs.csr_row_run_offsets[v]; r < runs.csr_row_run_offsets[v + 1]; ++r)
      output.csr_run_masks[r] = own & tile_lane_masks[runs.csr_run_destination_tiles[r].value()];
    for (auto r = runs.csc_column_run_offsets[v]; r < runs.csc_column_run_offsets[v + 1]; ++r)
      output.csc_run_masks[r] = own & tile_lane_masks[runs.csc_run_source_tiles[r].value()];
  }
}

}  // namespace bfnew
