/*
 * bfnew_host.h -- C ABI of the host-side graph library (libbfnew.so).
 *
 * These entry points expose the host half of the path to non-C++ callers
 * (the Python tests and benchmark bind them with ctypes):
 *   bfnh_graph_build    -- build_spatially_ordered_graph + build_tile_run_layout +
 *                          build_device_graph_layout32 (reorder.cpp:151,
 *                          device_layout.cpp:186, device_layout.cpp:395)
 *   bfnh_route_query    -- make_route_query's selected tiles (query.cpp:128-159)
 *   bfnh_synthetic_*    -- the synthetic tile-grid workloads (no reference
 *                          counterpart)
 * Returned array pointers stay valid until the owning handle is freed.
 */
#ifndef BFNEW_HOST_H
#define BFNEW_HOST_H

#include <stdint.h>

#include "bfnew_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfnh_graph_s *bfnh_graph;

typedef struct {
  int32_t x;
  int32_t y;
  uint32_t has_location;
  uint32_t resource_class;
} bfnh_vertex;

typedef struct {
  uint32_t source;
  uint32_t destination;
  float weight;
  uint32_t provenance_domain;
  uint32_t provenance_kind;
  uint32_t pad;
  uint64_t provenance_record;
} bfnh_edge;

typedef struct {
  uint32_t vertex_count;
  uint32_t tile_count;
  uint64_t edge_count;
  const uint64_t *row_offsets;      /* [V+1] */
  const uint32_t *destinations;     /* [E] */
  const float *weights;             /* [E] */
  const uint64_t *edge_ids;         /* [E] */
  const uint64_t *column_offsets;   /* [V+1] */
  const uint32_t *sources;          /* [E] */
  const float *in_weights;          /* [E] */
  const uint64_t *in_edge_ids;      /* [E] */
  const uint32_t *old_to_new;       /* [V] */
  const uint32_t *new_to_old;       /* [V] */
  const uint32_t *owner_tiles;      /* [V] */
  const uint64_t *tile_vertex_offsets; /* [T+1] */
  const int64_t *tile_x;            /* [T] */
  const int64_t *tile_y;            /* [T] */
  const uint32_t *tile_has_location;/* [T] */
  float *node_costs;                /* [V] synthetic graphs only, else NULL */
} bfnh_graph_view;

const char *bfnh_last_error(void);

/* Generic build from reference input records; tile_width == 0 builds the
 * unordered weighted graph (no tiles, no layout). */
int32_t bfnh_graph_build(const bfnh_vertex *vertices, uint32_t vertex_count,
                         const bfnh_edge *edges, uint64_t edge_count, int32_t origin_x,
                         int32_t origin_y, uint32_t tile_width, uint32_t tile_height,
                         bfnh_graph *out);
/* Synthetic tile grid (threads == 0: all hardware threads). with_host_graph
 * also materialises the WeightedGraph view (bfnh_graph_view); the layout is
 * always available. */
int32_t bfnh_synthetic_graph(uint32_t tiles_x, uint32_t tiles_y, uint32_t nodes_per_tile,
                             uint64_t seed, uint32_t threads, uint32_t with_host_graph,
                             bfnh_graph *out);
/* Rebuilds a synthetic graph through the generic reference-order builder
 * (InputGraph -> build_spatially_ordered_graph); for cross-checks. */
int32_t bfnh_synthetic_generic(bfnh_graph synthetic, bfnh_graph *out);
/* The synthetic graph's reference input records (vertex_count vertices,
 * edge_count edges, input order); for cross-checks with other builders. */
int32_t bfnh_synthetic_input(bfnh_graph synthetic, bfnh_vertex *vertices, bfnh_edge *edges);
int32_t bfnh_graph_free(bfnh_graph graph);
int32_t bfnh_graph_view_get(bfnh_graph graph, bfnh_graph_view *out);
/* The reference DeviceGraphLayout32 of the graph, ready for bfn_graph_upload. */
int32_t bfnh_graph_layout(bfnh_graph graph, bfn_graph_layout32 *out);
int32_t bfnh_graph_fingerprint(bfnh_graph graph, uint64_t *first, uint64_t *second);

/* make_route_query(..., tile_padding): canonical sorted sources/targets and
 * the padded terminal rectangle (+ spill when a terminal is unlocated).
 * Writes up to capacity tiles; returns the count in *tile_count. */
int32_t bfnh_route_query_tiles(bfnh_graph graph, const uint32_t *terminals,
                               uint32_t terminal_count, uint32_t padding, uint32_t *tiles,
                               uint32_t capacity, uint32_t *tile_count);

/* Synthetic query set: mix 0 = box side log-uniform in [side_min, side_max]
 * with unbounded_fraction of unbounded queries, mix 1 = serving mix. */
typedef struct {
  uint32_t source;
  uint32_t target;
  uint32_t box_x;
  uint32_t box_y;
  uint32_t box_side;
  uint32_t unbounded;
} bfnh_query;
int32_t bfnh_synthetic_queries(bfnh_graph synthetic, uint32_t count, uint64_t seed,
                               uint32_t side_min, uint32_t side_max, double unbounded_fraction,
                               uint32_t mix, bfnh_query *out);
/* Selected tiles of one synthetic query (count first with tiles == NULL). */
int32_t bfnh_synthetic_query_tiles(bfnh_graph synthetic, const bfnh_query *query,
                                   uint32_t *tiles, uint32_t capacity, uint32_t *tile_count);

#ifdef __cplusplus
}
#endif

#endif
