// host_capi.cpp -- the C entry points of include/bfnew_host.h.
#include "bfnew_host.h"

#include "bfnew/device_layout.hpp"
#include "bfnew/query.hpp"
#include "bfnew/synthetic.hpp"

#include <cstring>
#include <exception>
#include <memory>
#include <optional>
#include <string>
#include <vector>

using namespace bfnew;

namespace {
thread_local std::string g_error;

int32_t fail(int32_t code, const std::string& message) {
  g_error = message;
  return code;
}

template <typename F>
int32_t guarded(F&& body) {
  try {
    return body();
  } catch (const std::invalid_argument& e) {
    return fail(BFN_E_INVALID_ARGUMENT, e.what());
  } catch (const std::out_of_range& e) {
    return fail(BFN_E_INVALID_ARGUMENT, e.what());
  } catch (const std::overflow_error& e) {
    return fail(BFN_E_NOT_REPRESENTABLE, e.what());
  } catch (const std::bad_alloc&) {
    return fail(BFN_E_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(BFN_E_DEVICE_RESULT, e.what());
  }
}
}  // namespace

struct bfnh_graph_s {
  std::optional<WeightedGraph> graph;
  std::optional<SyntheticTileGraph> synthetic;
  std::optional<DeviceGraphLayout32> own_layout;
  std::vector<std::uint32_t> tile_offsets32;
  std::vector<std::uint64_t> tile_offsets64;
  std::vector<std::int64_t> tile_x, tile_y;
  std::vector<std::uint32_t> tile_has;
  std::vector<std::uint32_t> owner32;

  const DeviceGraphLayout32* layout() const {
    if (synthetic) return &synthetic->layout();
    if (own_layout) return &*own_layout;
    return nullptr;
  }
  void index_graph() {
    if (!graph || !graph->has_spatial_ordering()) return;
    const auto coords = graph->tile_coordinates();
    tile_x.clear();
    tile_y.clear();
    tile_has.clear();
    for (const auto& c : coords) {
      tile_x.push_back(c.tile_x);
      tile_y.push_back(c.tile_y);
      tile_has.push_back(c.has_location ? 1U : 0U);
    }
    tile_offsets64.assign(graph->tile_vertex_offsets().begin(), graph->tile_vertex_offsets().end());
    tile_offsets32.assign(tile_offsets64.begin(), tile_offsets64.end());
    owner32.resize(graph->vertex_count());
    for (std::uint32_t v = 0; v < graph->vertex_count(); ++v) owner32[v] = graph->owner_tiles()[v].value();
  }
};

extern "C" {

const char* bfnh_last_error(void) { return g_error.c_str(); }

int32_t bfnh_graph_build(const bfnh_vertex* vertices, uint32_t vertex_count, const bfnh_edge* edges,
                         uint64_t edge_count, int32_t origin_x, int32_t origin_y,
                         uint32_t tile_width, uint32_t tile_height, bfnh_graph* out) {
  if (!out || (vertex_count && !vertices) || (edge_count && !edges))
    return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    std::vector<VertexMetadata> v(vertex_count);
    for (uint32_t i = 0; i < vertex_count; ++i) {
      const ResourceClassId cls{static_cast<std::uint16_t>(vertices[i].resource_class)};
      v[i] = vertices[i].has_location ? VertexMetadata::located(vertices[i].x, vertices[i].y, cls)
                                      : VertexMetadata::unlocated(cls);
    }
    std::vector<EdgeInputRecord> e(edge_count);
    for (uint64_t i = 0; i < edge_count; ++i)
      e[i] = EdgeInputRecord{VertexId{edges[i].source}, VertexId{edges[i].destination}, edges[i].weight,
                             PhysicalProvenance{edges[i].provenance_domain, edges[i].provenance_kind,
                                                edges[i].provenance_record}};
    InputGraph input(std::move(v), std::move(e));
    auto h = std::make_unique<bfnh_graph_s>();
    if (tile_width == 0) {
      h->graph.emplace(build_weighted_graph(input));
    } else {
      h->graph.emplace(build_spatially_ordered_graph(
          input, SpatialOrderConfig{origin_x, origin_y, tile_width, tile_height}));
      const TileRunLayout64 runs = build_tile_run_layout(*h->graph);
      h->own_layout.emplace(build_device_graph_layout32(*h->graph, runs));
    }
    h->index_graph();
    *out = h.release();
    return BFN_OK;
  });
}

int32_t bfnh_synthetic_graph(uint32_t tiles_x, uint32_t tiles_y, uint32_t nodes_per_tile,
                             uint64_t seed, uint32_t threads, uint32_t with_host_graph,
                             bfnh_graph* out) {
  if (!out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    SyntheticTileGridConfig c;
    c.tiles_x = tiles_x;
    c.tiles_y = tiles_y;
    c.nodes_per_tile = nodes_per_tile;
    c.seed = seed;
    auto h = std::make_unique<bfnh_graph_s>();
    h->synthetic.emplace(SyntheticTileGraph::generate(c, threads));
    if (with_host_graph) {
      h->graph.emplace(h->synthetic->weighted_graph());
      h->index_graph();
    } else {
      const auto offs = h->synthetic->tile_vertex_offsets();
      h->tile_offsets32.assign(offs.begin(), offs.end());
      h->tile_offsets64.assign(offs.begin(), offs.end());
      for (std::uint32_t ty = 0; ty < tiles_y; ++ty)
        for (std::uint32_t tx = 0; tx < tiles_x; ++tx) {
          h->tile_x.push_back(tx);
          h->tile_y.push_back(ty);
          h->tile_has.push_back(1);
        }
      h->tile_x.push_back(0);
      h->tile_y.push_back(0);
      h->tile_has.push_back(0);
    }
    *out = h.release();
    return BFN_OK;
  });
}

int32_t bfnh_synthetic_generic(bfnh_graph synthetic, bfnh_graph* out) {
  if (!synthetic || !synthetic->synthetic || !out) return fail(BFN_E_INVALID_ARGUMENT, "not a synthetic graph");
  return guarded([&] {
    auto h = std::make_unique<bfnh_graph_s>();
    const InputGraph input = synthetic->synthetic->input_graph();
    h->graph.emplace(build_spatially_ordered_graph(input, SpatialOrderConfig{0, 0, 1, 1}));
    const TileRunLayout64 runs = build_tile_run_layout(*h->graph);
    h->own_layout.emplace(build_device_graph_layout32(*h->graph, runs));
    h->index_graph();
    *out = h.release();
    return BFN_OK;
  });
}

int32_t bfnh_synthetic_input(bfnh_graph h, bfnh_vertex* vertices, bfnh_edge* edges) {
  if (!h || !h->synthetic || !vertices || !edges) return fail(BFN_E_INVALID_ARGUMENT, "not a synthetic graph");
  return guarded([&] {
    const InputGraph input = h->synthetic->input_graph();
    for (std::uint32_t v = 0; v < input.vertex_count(); ++v) {
      const VertexMetadata& m = input.vertices()[v];
      vertices[v] = bfnh_vertex{m.x, m.y, m.has_location ? 1U : 0U, m.resource_class.value()};
    }
    for (std::uint64_t e = 0; e < input.edge_count(); ++e) {
      const EdgeInputRecord& r = input.edges()[e];
      edges[e] = bfnh_edge{r.source.value(), r.destination.value(), r.weight, r.provenance.domain,
                           r.provenance.kind_and_flags, 0, r.provenance.source_record};
    }
    return BFN_OK;
  });
}

int32_t bfnh_graph_free(bfnh_graph graph) {
  delete graph;
  return BFN_OK;
}

int32_t bfnh_graph_view_get(bfnh_graph h, bfnh_graph_view* out) {
  if (!h || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  std::memset(out, 0, sizeof *out);
  if (h->graph) {
    const WeightedGraph& g = *h->graph;
    out->vertex_count = g.vertex_count();
    out->edge_count = g.edge_count();
    out->tile_count = static_cast<uint32_t>(g.tile_coordinates().size());
    const auto o = g.outgoing();
    const auto i = g.incoming();
    out->row_offsets = o.row_offsets.data();
    out->destinations = reinterpret_cast<const uint32_t*>(o.destinations.data());
    out->weights = o.weights.data();
    out->edge_ids = reinterpret_cast<const uint64_t*>(o.edge_ids.data());
    out->column_offsets = i.column_offsets.data();
    out->sources = reinterpret_cast<const uint32_t*>(i.sources.data());
    out->in_weights = i.weights.data();
    out->in_edge_ids = reinterpret_cast<const uint64_t*>(i.edge_ids.data());
    out->old_to_new = reinterpret_cast<const uint32_t*>(g.old_to_new().data());
    out->new_to_old = reinterpret_cast<const uint32_t*>(g.new_to_old().data());
    out->owner_tiles = h->owner32.empty() ? nullptr : h->owner32.data();
  } else if (h->synthetic) {
    const auto& L = h->synthetic->layout();
    out->vertex_count = L.vertex_count;
    out->edge_count = L.edge_count;
    out->tile_count = L.tile_count;
    out->owner_tiles = L.owner_tiles.data();
  }
  out->tile_vertex_offsets = h->tile_offsets64.data();
  out->tile_x = h->tile_x.data();
  out->tile_y = h->tile_y.data();
  out->tile_has_location = h->tile_has.data();
  out->node_costs = h->synthetic ? const_cast<float*>(h->synthetic->node_costs().data()) : nullptr;
  return BFN_OK;
}

int32_t bfnh_graph_layout(bfnh_graph h, bfn_graph_layout32* out) {
  if (!h || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  const DeviceGraphLayout32* L = h->layout();
  if (!L) return fail(BFN_E_STATE, "graph has no spatial layout");
  std::memset(out, 0, sizeof *out);
  out->vertex_count = L->vertex_count;
  out->edge_count = L->edge_count;
  out->tile_count = L->tile_count;
  out->csr_run_count = static_cast<uint32_t>(L->csr_run_destination_tiles.size());
  out->csc_run_count = static_cast<uint32_t>(L->csc_run_source_tiles.size());
  out->owner_tiles = L->owner_tiles.data();
  out->tile_vertex_offsets = h->tile_offsets32.data();
  out->csr_row_offsets = L->csr_row_offsets.data();
  out->csr_destinations = L->csr_destinations.data();
  out->csr_weights = L->csr_weights.data();
  out->csr_row_run_offsets = L->csr_row_run_offsets.data();
  out->csr_run_edge_offsets = L->csr_run_edge_offsets.data();
  out->csr_run_destination_tiles = L->csr_run_destination_tiles.data();
  out->csc_column_offsets = L->csc_column_offsets.data();
  out->csc_sources = L->csc_sources.data();
  out->csc_weights = L->csc_weights.data();
  out->csc_edge_ids = L->csc_edge_ids.data();
  out->csc_column_run_offsets = L->csc_column_run_offsets.data();
  out->csc_run_edge_offsets = L->csc_run_edge_offsets.data();
  out->csc_run_source_tiles = L->csc_run_source_tiles.data();
  return BFN_OK;
}

int32_t bfnh_graph_fingerprint(bfnh_graph h, uint64_t* first, uint64_t* second) {
  if (!h || !first || !second) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  const DeviceGraphLayout32* L = h->layout();
  if (!L) return fail(BFN_E_STATE, "graph has no spatial layout");
  const DeviceGraphFingerprint f = fingerprint_device_graph_layout32(*L);
  *first = f.first;
  *second = f.second;
  return BFN_OK;
}

int32_t bfnh_route_query_tiles(bfnh_graph h, const uint32_t* terminals, uint32_t terminal_count,
                               uint32_t padding, uint32_t* tiles, uint32_t capacity,
                               uint32_t* tile_count) {
  if (!h || !h->graph || !tile_count || (terminal_count && !terminals))
    return fail(BFN_E_INVALID_ARGUMENT, "route queries need a built host graph");
  return guarded([&] {
    std::vector<VertexId> t(terminal_count);
    for (uint32_t i = 0; i < terminal_count; ++i) t[i] = VertexId{terminals[i]};
    const std::vector<TileId> sel = select_route_tiles(*h->graph, t, padding);
    *tile_count = static_cast<uint32_t>(sel.size());
    if (tiles) {
      if (capacity < sel.size()) return fail(BFN_E_INVALID_ARGUMENT, "tile capacity too small");
      for (std::size_t i = 0; i < sel.size(); ++i) tiles[i] = sel[i].value();
    }
    return BFN_OK;
  });
}

int32_t bfnh_synthetic_queries(bfnh_graph h, uint32_t count, uint64_t seed, uint32_t side_min,
                               uint32_t side_max, double unbounded_fraction, uint32_t mix,
                               bfnh_query* out) {
  if (!h || !h->synthetic || (count && !out)) return fail(BFN_E_INVALID_ARGUMENT, "not a synthetic graph");
  if (mix > 1) return fail(BFN_E_INVALID_ARGUMENT, "invalid query mix");
  return guarded([&] {
    SyntheticQueryConfig qc;
    qc.count = count;
    qc.seed = seed;
    qc.side_min = side_min;
    qc.side_max = side_max;
    qc.unbounded_fraction = unbounded_fraction;
    qc.mix = static_cast<SyntheticBoxMix>(mix);
    const auto qs = generate_synthetic_queries(h->synthetic->config(), qc);
    for (uint32_t i = 0; i < count; ++i)
      out[i] = bfnh_query{qs[i].source, qs[i].target, qs[i].box_x, qs[i].box_y, qs[i].box_side,
                          qs[i].unbounded ? 1U : 0U};
    return BFN_OK;
  });
}

int32_t bfnh_synthetic_query_tiles(bfnh_graph h, const bfnh_query* q, uint32_t* tiles,
                                   uint32_t capacity, uint32_t* tile_count) {
  if (!h || !h->synthetic || !q || !tile_count) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  SyntheticQuery s{0, q->source, q->target, q->box_x, q->box_y, q->box_side, q->unbounded != 0};
  const auto sel = synthetic_query_tiles(h->synthetic->config(), s);
  *tile_count = static_cast<uint32_t>(sel.size());
  if (tiles) {
    if (capacity < sel.size()) return fail(BFN_E_INVALID_ARGUMENT, "tile capacity too small");
    for (std::size_t i = 0; i < sel.size(); ++i) tiles[i] = sel[i].value();
  }
  return BFN_OK;
}

}  // extern "C"
