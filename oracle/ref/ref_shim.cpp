// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C entry layer over the reference's own CPU sources (compiled unmodified
// from the reference tree by oracle/ref/Makefile into oracle/_ref/).  The
// tests use it to pin the oracle restatement against the reference itself,
// and bench.py may time its Dijkstra as the reference CPU baseline.  Nothing
// in the product links it.
//
// The reference's query layer (query.cpp) does not compile with this
// toolchain's libstdc++ 11, so the induced-subgraph step of
// build_induced_query_graph (query.cpp:286-362) is restated here: keep the
// vertices owned by selected tiles in global order, keep the outgoing CSR
// edges whose two endpoints are kept, carry provenance, and hand the result to
// the reference's own build_weighted_graph.
#include "bfnew/device_layout.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/spatial.hpp"
#include "bfnew/sssp.hpp"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <atomic>
#include <memory>
#include <thread>
#include <vector>

namespace {

struct RefGraph {
  bfnew::WeightedGraph graph;
  bfnew::TileRunLayout64 runs;
  bfnew::DeviceGraphLayout32 layout;
  bfnew::DeviceGraphFingerprint fingerprint;
};

struct RefVertex {
  std::int32_t x;
  std::int32_t y;
  std::uint32_t has_location;
  std::uint32_t resource_class;
};

struct RefEdge {
  std::uint32_t source;
  std::uint32_t destination;
  float weight;
  std::uint32_t provenance_domain;
  std::uint32_t provenance_kind;
  std::uint32_t pad;
  std::uint64_t provenance_record;
};

bfnew::InputGraph make_input(const RefVertex* vertices, std::uint32_t vertex_count,
                             const RefEdge* edges, std::uint64_t edge_count) {
  std::vector<bfnew::VertexMetadata> v;
  v.reserve(vertex_count);
  for (std::uint32_t i = 0; i < vertex_count; ++i) {
    const bfnew::ResourceClassId cls{static_cast<std::uint16_t>(vertices[i].resource_class)};
    v.push_back(vertices[i].has_location
                    ? bfnew::VertexMetadata::located(vertices[i].x, vertices[i].y, cls)
                    : bfnew::VertexMetadata::unlocated(cls));
  }
  std::vector<bfnew::EdgeInputRecord> e;
  e.reserve(edge_count);
  for (std::uint64_t i = 0; i < edge_count; ++i) {
    e.push_back(bfnew::EdgeInputRecord{
        bfnew::VertexId{edges[i].source}, bfnew::VertexId{edges[i].destination},
        edges[i].weight,
        bfnew::PhysicalProvenance{edges[i].provenance_domain, edges[i].provenance_kind,
                                  edges[i].provenance_record}});
  }
  return bfnew::InputGraph{std::move(v), std::move(e)};
}

template <typename T, typename S>
void copy_out(T* dst, const S& src) {
  if (dst == nullptr) return;
  std::size_t i = 0;
  for (const auto& value : src) {
    if constexpr (std::is_same_v<std::decay_t<decltype(value)>, bfnew::VertexId> ||
                  std::is_same_v<std::decay_t<decltype(value)>, bfnew::TileId> ||
                  std::is_same_v<std::decay_t<decltype(value)>, bfnew::EdgeId>) {
      dst[i++] = static_cast<T>(value.value());
    } else {
      dst[i++] = static_cast<T>(value);
    }
  }
}

}  // namespace

// Induced query graph: the selected tiles' vertices and the edges with both
// ends selected, built by the reference's build_weighted_graph (provenance =
// CSR position, so EdgeIds follow CSR order), and the sources renumbered.
struct InducedQuery {
  bool valid = false;
  std::unique_ptr<bfnew::WeightedGraph> graph;
  std::vector<bfnew::VertexId> sources;
};

InducedQuery build_induced(std::uint32_t vertex_count, const std::uint64_t* row_offsets,
                           const std::uint32_t* destinations, const float* weights,
                           const std::uint32_t* owner_tiles, const std::uint8_t* tile_selected,
                           const std::uint32_t* sources, std::uint32_t source_count) {
  constexpr std::uint32_t unmapped = std::numeric_limits<std::uint32_t>::max();
  std::vector<std::uint32_t> local(vertex_count, unmapped);
  std::uint32_t kept = 0;
  for (std::uint32_t v = 0; v < vertex_count; ++v)
    if (tile_selected[owner_tiles[v]]) local[v] = kept++;
  std::vector<bfnew::VertexMetadata> vertices(kept,
                                              bfnew::VertexMetadata::unlocated(bfnew::ResourceClassId{1}));
  std::vector<bfnew::EdgeInputRecord> edges;
  for (std::uint32_t s = 0; s < vertex_count; ++s) {
    if (local[s] == unmapped) continue;
    for (std::uint64_t p = row_offsets[s]; p < row_offsets[s + 1]; ++p) {
      const std::uint32_t d = destinations[p];
      if (local[d] == unmapped) continue;
      edges.push_back(bfnew::EdgeInputRecord{bfnew::VertexId{local[s]}, bfnew::VertexId{local[d]},
                                             weights[p], bfnew::PhysicalProvenance{1U, 1U, p}});
    }
  }
  InducedQuery q;
  q.graph = std::make_unique<bfnew::WeightedGraph>(
      bfnew::build_weighted_graph(bfnew::InputGraph(std::move(vertices), std::move(edges))));
  for (std::uint32_t i = 0; i < source_count; ++i) {
    if (sources[i] >= vertex_count || local[sources[i]] == unmapped) return q;
    q.sources.push_back(bfnew::VertexId{local[sources[i]]});
  }
  q.valid = true;
  return q;
}

// Out-edges of reached vertices: what Dijkstra relaxes.
std::uint64_t reached_out_edges(const bfnew::WeightedGraph& g, const bfnew::SsspResult& result) {
  std::uint64_t count = 0;
  const auto out = g.outgoing();
  for (std::uint32_t s = 0; s < g.vertex_count(); ++s)
    if (result.distances[s] != std::numeric_limits<float>::infinity())
      count += out.row_offsets[s + 1] - out.row_offsets[s];
  return count;
}

template <typename F>
void run_threads(std::uint32_t n, F&& f) {
  std::vector<std::thread> pool;
  for (std::uint32_t i = 0; i < n; ++i) pool.emplace_back([&f, i] { f(i); });
  for (auto& t : pool) t.join();
}

extern "C" {

// Builds the spatially ordered graph (or the plain weighted graph when
// tile_width == 0), its tile runs, the 32-bit device layout and fingerprint.
void* ref_build(const RefVertex* vertices, std::uint32_t vertex_count, const RefEdge* edges,
                std::uint64_t edge_count, std::int32_t origin_x, std::int32_t origin_y,
                std::uint32_t tile_width, std::uint32_t tile_height) {
  try {
    bfnew::InputGraph input = make_input(vertices, vertex_count, edges, edge_count);
    if (tile_width == 0) {
      bfnew::WeightedGraph g = bfnew::build_weighted_graph(input);
      return new RefGraph{std::move(g), {}, {}, {}};
    }
    bfnew::WeightedGraph g = bfnew::build_spatially_ordered_graph(
        input, bfnew::SpatialOrderConfig{origin_x, origin_y, tile_width, tile_height});
    bfnew::TileRunLayout64 runs = bfnew::build_tile_run_layout(g);
    bfnew::DeviceGraphLayout32 layout = bfnew::build_device_graph_layout32(g, runs);
    const bfnew::DeviceGraphFingerprint fp = bfnew::fingerprint_device_graph_layout32(layout);
    return new RefGraph{std::move(g), std::move(runs), std::move(layout), fp};
  } catch (const std::exception&) {
    return nullptr;
  }
}

void ref_free(void* handle) { delete static_cast<RefGraph*>(handle); }

// counts[0..5] = V, E, tiles, csr runs, csc runs, has_layout
void ref_counts(void* handle, std::uint64_t* counts) {
  const auto* r = static_cast<RefGraph*>(handle);
  counts[0] = r->graph.vertex_count();
  counts[1] = r->graph.edge_count();
  counts[2] = r->graph.tile_coordinates().size();
  counts[3] = r->runs.csr_run_destination_tiles.size();
  counts[4] = r->runs.csc_run_source_tiles.size();
  counts[5] = r->layout.vertex_count != 0 || r->graph.vertex_count() == 0 ? 1 : 0;
}

void ref_graph_arrays(void* handle, std::uint64_t* row_offsets, std::uint32_t* destinations,
                      float* weights, std::uint64_t* edge_ids, std::uint64_t* column_offsets,
                      std::uint32_t* sources, float* in_weights, std::uint64_t* in_edge_ids,
                      std::uint32_t* old_to_new, std::uint32_t* new_to_old,
                      std::uint32_t* owner_tiles, std::uint64_t* tile_vertex_offsets) {
  const auto* r = static_cast<RefGraph*>(handle);
  const auto out = r->graph.outgoing();
  const auto in = r->graph.incoming();
  copy_out(row_offsets, out.row_offsets);
  copy_out(destinations, out.destinations);
  copy_out(weights, out.weights);
  copy_out(edge_ids, out.edge_ids);
  copy_out(column_offsets, in.column_offsets);
  copy_out(sources, in.sources);
  copy_out(in_weights, in.weights);
  copy_out(in_edge_ids, in.edge_ids);
  copy_out(old_to_new, r->graph.old_to_new());
  copy_out(new_to_old, r->graph.new_to_old());
  copy_out(owner_tiles, r->graph.owner_tiles());
  copy_out(tile_vertex_offsets, r->graph.tile_vertex_offsets());
}

void ref_run_arrays(void* handle, std::uint64_t* csr_row_runs, std::uint64_t* csr_run_edges,
                    std::uint32_t* csr_run_tiles, std::uint64_t* csc_col_runs,
                    std::uint64_t* csc_run_edges, std::uint32_t* csc_run_tiles) {
  const auto* r = static_cast<RefGraph*>(handle);
  copy_out(csr_row_runs, r->runs.csr_row_run_offsets);
  copy_out(csr_run_edges, r->runs.csr_run_edge_offsets);
  copy_out(csr_run_tiles, r->runs.csr_run_destination_tiles);
  copy_out(csc_col_runs, r->runs.csc_column_run_offsets);
  copy_out(csc_run_edges, r->runs.csc_run_edge_offsets);
  copy_out(csc_run_tiles, r->runs.csc_run_source_tiles);
}

void ref_fingerprint(void* handle, std::uint64_t* out) {
  const auto* r = static_cast<RefGraph*>(handle);
  out[0] = r->fingerprint.first;
  out[1] = r->fingerprint.second;
}

// Distances on the subgraph induced by tile_selected (nullptr = every tile),
// scattered back to global ids with +inf outside.  algorithm: 0 = Dijkstra,
// 1 = synchronous Bellman-Ford.  Returns the solver-only wall time in ns, or
// a negative value on error.  relaxations receives the admitted edges whose
// source is reachable (Dijkstra's relaxation count).
std::int64_t ref_sssp(void* handle, const std::uint8_t* tile_selected,
                      const std::uint32_t* sources, std::uint32_t source_count,
                      int algorithm, float* distances, std::uint64_t* relaxations) {
  try {
    const auto* r = static_cast<RefGraph*>(handle);
    const bfnew::WeightedGraph& g = r->graph;
    const std::uint32_t V = g.vertex_count();
    const auto owner = g.owner_tiles();
    const auto keep = [&](std::uint32_t v) {
      return tile_selected == nullptr || owner.empty() || tile_selected[owner[v].value()] != 0;
    };
    constexpr std::uint32_t unmapped = std::numeric_limits<std::uint32_t>::max();
    std::vector<std::uint32_t> global_to_local(V, unmapped);
    std::vector<std::uint32_t> local_to_global;
    std::vector<bfnew::VertexMetadata> vertices;
    for (std::uint32_t v = 0; v < V; ++v) {
      if (!keep(v)) continue;
      global_to_local[v] = static_cast<std::uint32_t>(local_to_global.size());
      local_to_global.push_back(v);
      vertices.push_back(g.vertices()[v]);
    }
    std::vector<bfnew::EdgeInputRecord> edges;
    const auto out = g.outgoing();
    for (std::uint32_t s = 0; s < V; ++s) {
      if (global_to_local[s] == unmapped) continue;
      for (auto p = out.row_offsets[s]; p < out.row_offsets[s + 1]; ++p) {
        const std::uint32_t d = out.destinations[p].value();
        if (global_to_local[d] == unmapped) continue;
        edges.push_back(bfnew::EdgeInputRecord{
            bfnew::VertexId{global_to_local[s]}, bfnew::VertexId{global_to_local[d]},
            out.weights[p], g.edge_provenance()[out.edge_ids[p].value()]});
      }
    }
    bfnew::InputGraph input(std::move(vertices), std::move(edges));
    const bfnew::WeightedGraph induced = bfnew::build_weighted_graph(input);
    std::vector<bfnew::VertexId> local_sources;
    for (std::uint32_t i = 0; i < source_count; ++i) {
      if (sources[i] >= V || global_to_local[sources[i]] == unmapped) return -1;
      local_sources.push_back(bfnew::VertexId{global_to_local[sources[i]]});
    }
    const auto start = std::chrono::steady_clock::now();
    const bfnew::SsspResult result = algorithm == 0
                                         ? bfnew::dijkstra_oracle(induced, local_sources)
                                         : bfnew::synchronous_bellman_ford(induced, local_sources);
    const auto stop = std::chrono::steady_clock::now();
    for (std::uint32_t v = 0; v < V; ++v) distances[v] = std::numeric_limits<float>::infinity();
    for (std::size_t l = 0; l < local_to_global.size(); ++l)
      distances[local_to_global[l]] = result.distances[l];
    if (relaxations != nullptr) {
      std::uint64_t count = 0;
      const auto iout = induced.outgoing();
      for (std::uint32_t s = 0; s < induced.vertex_count(); ++s) {
        if (result.distances[s] != std::numeric_limits<float>::infinity())
          count += iout.row_offsets[s + 1] - iout.row_offsets[s];
      }
      *relaxations = count;
    }
    return std::chrono::duration_cast<std::chrono::nanoseconds>(stop - start).count();
  } catch (const std::exception&) {
    return -2;
  }
}

// The reference CPU path on one query of a graph given as CSR arrays (the
// product's synthetic layout): the induced subgraph of the selected tiles
// (query.cpp:286-362 restated: kept vertices in global order, edges with both
// endpoints kept, provenance = CSR position so canonical order is kept) is
// built by the reference's build_weighted_graph, then dijkstra_oracle runs
// timed.  Returns the Dijkstra wall time in ns (<0 on error); relaxations
// receives the admitted edges whose source is reachable.
std::int64_t ref_csr_sssp(std::uint32_t vertex_count, const std::uint64_t* row_offsets,
                          const std::uint32_t* destinations, const float* weights,
                          const std::uint32_t* owner_tiles, const std::uint8_t* tile_selected,
                          const std::uint32_t* sources, std::uint32_t source_count,
                          std::uint64_t* relaxations) {
  try {
    InducedQuery q = build_induced(vertex_count, row_offsets, destinations, weights, owner_tiles,
                                   tile_selected, sources, source_count);
    if (!q.valid) return -1;
    const auto start = std::chrono::steady_clock::now();
    const bfnew::SsspResult result = bfnew::dijkstra_oracle(*q.graph, q.sources);
    const auto stop = std::chrono::steady_clock::now();
    if (relaxations != nullptr) *relaxations = reached_out_edges(*q.graph, result);
    return std::chrono::duration_cast<std::chrono::nanoseconds>(stop - start).count();
  } catch (const std::exception&) {
    return -2;
  }
}

// The same Dijkstra over many queries, one query per worker thread, in waves
// of `threads` queries: each wave builds its induced graphs in parallel
// (untimed), then runs their Dijkstras in parallel; the returned time is the
// sum of the waves' Dijkstra wall times.  Query i selects the tiles
// tiles[tile_offsets[i] .. tile_offsets[i+1]) and has the single source
// sources[i].  Returns -1 for an invalid query, -2 on error.
std::int64_t ref_csr_sssp_parallel(std::uint32_t vertex_count, const std::uint64_t* row_offsets,
                                   const std::uint32_t* destinations, const float* weights,
                                   const std::uint32_t* owner_tiles, std::uint32_t tile_count,
                                   std::uint32_t query_count, const std::uint32_t* tile_offsets,
                                   const std::uint32_t* tiles, const std::uint32_t* sources,
                                   std::uint32_t threads, std::uint64_t* relaxations) {
  try {
    if (threads == 0) threads = 1;
    std::uint64_t total_ns = 0, total_relax = 0;
    for (std::uint32_t first = 0; first < query_count; first += threads) {
      const std::uint32_t n = std::min(threads, query_count - first);
      std::vector<InducedQuery> wave(n);
      std::atomic<bool> bad{false};
      auto build = [&](std::uint32_t i) {
        std::vector<std::uint8_t> selected(tile_count, 0);
        const std::uint32_t q = first + i;
        for (std::uint32_t k = tile_offsets[q]; k < tile_offsets[q + 1]; ++k) selected[tiles[k]] = 1;
        wave[i] = build_induced(vertex_count, row_offsets, destinations, weights, owner_tiles,
                                selected.data(), &sources[q], 1);
        if (!wave[i].valid) bad = true;
      };
      run_threads(n, build);
      if (bad) return -1;
      std::vector<std::uint64_t> reached(n, 0);
      const auto start = std::chrono::steady_clock::now();
      run_threads(n, [&](std::uint32_t i) {
        const bfnew::SsspResult result = bfnew::dijkstra_oracle(*wave[i].graph, wave[i].sources);
        reached[i] = reached_out_edges(*wave[i].graph, result);
      });
      const auto stop = std::chrono::steady_clock::now();
      total_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(stop - start).count();
      for (std::uint64_t r : reached) total_relax += r;
    }
    if (relaxations != nullptr) *relaxations = total_relax;
    return static_cast<std::int64_t>(total_ns);
  } catch (const std::exception&) {
    return -2;
  }
}

// reconstruct_path_from_distances on the full graph (no tile restriction),
// for the sssp_test fixtures.  Returns the path length in edges, -1 when the
// target is unreachable, -2 on error.
std::int64_t ref_reconstruct(void* handle, const std::uint32_t* sources,
                             std::uint32_t source_count, const float* distances,
                             std::uint32_t target, std::uint32_t* vertices,
                             std::uint64_t* edge_ids, std::uint32_t capacity) {
  try {
    const auto* r = static_cast<RefGraph*>(handle);
    bfnew::SsspResult result;
    result.distances.assign(distances, distances + r->graph.vertex_count());
    for (std::uint32_t i = 0; i < source_count; ++i)
      result.sources.push_back(bfnew::VertexId{sources[i]});
    std::sort(result.sources.begin(), result.sources.end());
    result.sources.erase(std::unique(result.sources.begin(), result.sources.end()),
                         result.sources.end());
    const auto path =
        bfnew::reconstruct_path_from_distances(r->graph, result, bfnew::VertexId{target});
    if (!path) return -1;
    if (path->vertices.size() > capacity + 1U) return -2;
    for (std::size_t i = 0; i < path->vertices.size(); ++i) vertices[i] = path->vertices[i].value();
    for (std::size_t i = 0; i < path->edge_ids.size(); ++i) edge_ids[i] = path->edge_ids[i].value();
    return static_cast<std::int64_t>(path->edge_ids.size());
  } catch (const std::exception&) {
    return -2;
  }
}

}  // extern "C"
