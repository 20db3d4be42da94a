/*
 * oracle.c -- CPU restatement of the bfnew shortest-path path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Build with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NONE32 0xFFFFFFFFu

static uint32_t float_bits(float value) {
  uint32_t bits;
  memcpy(&bits, &value, sizeof bits);
  return bits;
}

/* graph_builder.cpp:91-93: -0 is stored as +0. */
static float canonicalize_weight(float weight) {
  return weight == 0.0f ? 0.0f : weight;
}

/* ---------------------------------------------------------------------------
 * Graph build: canonical EdgeIds (graph_builder.cpp:182-214), spatial vertex
 * order (reorder.cpp:151-273) and the tile-aware CSR/CSC orders
 * (reorder.cpp:293-342; without a spatial config every vertex shares one
 * owner tile, which reduces to graph_builder.cpp:216-273).
 * ------------------------------------------------------------------------- */

typedef struct {
  uint32_t source, destination, weight_bits, domain, kind;
  uint64_t record;
  uint64_t input_index;
} canonical_key;

static int cmp_u64(uint64_t a, uint64_t b) { return a < b ? -1 : (a > b); }

static int cmp_canonical(const void *pa, const void *pb) {
  const canonical_key *a = pa, *b = pb;
  int c;
  if ((c = cmp_u64(a->source, b->source))) return c;
  if ((c = cmp_u64(a->destination, b->destination))) return c;
  if ((c = cmp_u64(a->weight_bits, b->weight_bits))) return c;
  if ((c = cmp_u64(a->domain, b->domain))) return c;
  if ((c = cmp_u64(a->kind, b->kind))) return c;
  if ((c = cmp_u64(a->record, b->record))) return c;
  return cmp_u64(a->input_index, b->input_index);
}

typedef struct {
  int located;
  int64_t tile_y, tile_x;
  uint32_t resource_class;
  uint64_t locality;
  uint32_t original;
} vertex_key;

static int cmp_vertex(const void *pa, const void *pb) {
  const vertex_key *a = pa, *b = pb;
  int c;
  if (a->located != b->located) return a->located ? -1 : 1;
  if (a->located) {
    if (a->tile_y != b->tile_y) return a->tile_y < b->tile_y ? -1 : 1;
    if (a->tile_x != b->tile_x) return a->tile_x < b->tile_x ? -1 : 1;
    if ((c = cmp_u64(a->resource_class, b->resource_class))) return c;
    if ((c = cmp_u64(a->locality, b->locality))) return c;
    return cmp_u64(a->original, b->original);
  }
  if ((c = cmp_u64(a->resource_class, b->resource_class))) return c;
  return cmp_u64(a->original, b->original);
}

typedef struct {
  int64_t y, x;
} tile_coord;

static int cmp_tile(const void *pa, const void *pb) {
  const tile_coord *a = pa, *b = pb;
  if (a->y != b->y) return a->y < b->y ? -1 : 1;
  if (a->x != b->x) return a->x < b->x ? -1 : 1;
  return 0;
}

typedef struct {
  uint32_t major, tile, minor;
  float weight;
  uint64_t edge_id;
} layout_edge;

static int cmp_layout(const void *pa, const void *pb) {
  const layout_edge *a = pa, *b = pb;
  int c;
  if ((c = cmp_u64(a->major, b->major))) return c;
  if ((c = cmp_u64(a->tile, b->tile))) return c;
  if ((c = cmp_u64(a->minor, b->minor))) return c;
  return cmp_u64(a->edge_id, b->edge_id);
}

/* reorder.cpp:54-66 */
static int64_t floor_div(int64_t numerator, int64_t denominator) {
  int64_t q = numerator / denominator;
  if (numerator % denominator < 0) --q;
  return q;
}

/* reorder.cpp:133-149: x bits on even positions, y bits on odd positions. */
static uint64_t spread_bits(uint32_t input) {
  uint64_t v = input;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFULL;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFULL;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FULL;
  v = (v | (v << 2)) & 0x3333333333333333ULL;
  v = (v | (v << 1)) & 0x5555555555555555ULL;
  return v;
}

void orc_graph_free(orc_graph *g) {
  if (!g) return;
  free(g->row_offsets);
  free(g->destinations);
  free(g->weights);
  free(g->edge_ids);
  free(g->column_offsets);
  free(g->sources);
  free(g->in_weights);
  free(g->in_edge_ids);
  free(g->old_to_new);
  free(g->new_to_old);
  free(g->owner_tiles);
  free(g->tile_x);
  free(g->tile_y);
  free(g->tile_has_location);
  free(g->tile_vertex_offsets);
  memset(g, 0, sizeof *g);
}

int orc_build_graph(const orc_vertex *vertices, uint32_t vertex_count,
                    const orc_edge *edges, uint64_t edge_count,
                    const orc_spatial_config *config, orc_graph *out) {
  memset(out, 0, sizeof *out);
  for (uint64_t e = 0; e < edge_count; ++e) {
    float w = edges[e].weight;
    if (!isfinite(w) || w < 0.0f || edges[e].source >= vertex_count ||
        edges[e].destination >= vertex_count)
      return ORC_INVALID;
  }
  if (config && (config->tile_width == 0 || config->tile_height == 0))
    return ORC_INVALID;

  canonical_key *keys = malloc(sizeof *keys * (edge_count ? edge_count : 1));
  uint64_t *edge_id_of_input = malloc(sizeof(uint64_t) * (edge_count ? edge_count : 1));
  if (!keys || !edge_id_of_input) return ORC_NOMEM;
  for (uint64_t e = 0; e < edge_count; ++e) {
    keys[e].source = edges[e].source;
    keys[e].destination = edges[e].destination;
    keys[e].weight_bits = float_bits(canonicalize_weight(edges[e].weight));
    keys[e].domain = edges[e].provenance_domain;
    keys[e].kind = edges[e].provenance_kind;
    keys[e].record = edges[e].provenance_record;
    keys[e].input_index = e;
  }
  qsort(keys, edge_count, sizeof *keys, cmp_canonical);
  for (uint64_t rank = 0; rank < edge_count; ++rank)
    edge_id_of_input[keys[rank].input_index] = rank;
  free(keys);

  const uint32_t V = vertex_count;
  out->vertex_count = V;
  out->edge_count = edge_count;
  out->old_to_new = malloc(sizeof(uint32_t) * (V ? V : 1));
  out->new_to_old = malloc(sizeof(uint32_t) * (V ? V : 1));
  out->owner_tiles = malloc(sizeof(uint32_t) * (V ? V : 1));

  vertex_key *order = malloc(sizeof *order * (V ? V : 1));
  tile_coord *tiles = malloc(sizeof *tiles * (V ? V : 1));
  uint32_t tile_total = 0;
  for (uint32_t v = 0; v < V; ++v) {
    order[v].original = v;
    order[v].resource_class = vertices[v].resource_class;
    order[v].located = config ? (vertices[v].has_location != 0) : 0;
    order[v].tile_x = order[v].tile_y = 0;
    order[v].locality = 0;
    if (config && vertices[v].has_location) {
      int64_t dx = (int64_t)vertices[v].x - config->origin_x;
      int64_t dy = (int64_t)vertices[v].y - config->origin_y;
      int64_t tx = floor_div(dx, config->tile_width);
      int64_t ty = floor_div(dy, config->tile_height);
      uint32_t lx = (uint32_t)(dx - tx * (int64_t)config->tile_width);
      uint32_t ly = (uint32_t)(dy - ty * (int64_t)config->tile_height);
      order[v].tile_x = tx;
      order[v].tile_y = ty;
      order[v].locality = spread_bits(lx) | (spread_bits(ly) << 1);
      tiles[tile_total].x = tx;
      tiles[tile_total].y = ty;
      ++tile_total;
    }
  }
  uint32_t located_tiles = 0;
  if (config) {
    qsort(tiles, tile_total, sizeof *tiles, cmp_tile);
    for (uint32_t i = 0; i < tile_total; ++i)
      if (located_tiles == 0 || cmp_tile(&tiles[located_tiles - 1], &tiles[i]) != 0)
        tiles[located_tiles++] = tiles[i];
    /* Unordered graphs keep the plain vertex order (graph_builder.cpp:189-196). */
    qsort(order, V, sizeof *order, cmp_vertex);
  }
  /* Located tiles in (y, x) order, then one spill tile (reorder.cpp:195-249).
   * An unordered graph is modelled as a single tile holding everything. */
  const uint32_t tile_count = config ? located_tiles + 1 : 1;
  out->tile_count = tile_count;
  out->tile_x = calloc(tile_count, sizeof(int64_t));
  out->tile_y = calloc(tile_count, sizeof(int64_t));
  out->tile_has_location = calloc(tile_count, sizeof(uint32_t));
  out->tile_vertex_offsets = calloc(tile_count + 1, sizeof(uint64_t));
  for (uint32_t t = 0; t < located_tiles; ++t) {
    out->tile_x[t] = tiles[t].x;
    out->tile_y[t] = tiles[t].y;
    out->tile_has_location[t] = 1;
  }
  for (uint32_t n = 0; n < V; ++n) {
    const vertex_key *k = &order[n];
    uint32_t owner = 0;
    if (config) {
      owner = located_tiles; /* spill */
      if (k->located) {
        tile_coord probe = {k->tile_y, k->tile_x};
        tile_coord *hit = bsearch(&probe, tiles, located_tiles, sizeof *tiles, cmp_tile);
        owner = (uint32_t)(hit - tiles);
      }
    }
    out->new_to_old[n] = k->original;
    out->old_to_new[k->original] = n;
    out->owner_tiles[n] = owner;
    out->tile_vertex_offsets[owner + 1]++;
  }
  for (uint32_t t = 1; t <= tile_count; ++t)
    out->tile_vertex_offsets[t] += out->tile_vertex_offsets[t - 1];
  free(order);
  free(tiles);

  /* Outgoing CSR by (source, destination tile, destination, EdgeId). */
  const uint64_t E = edge_count;
  layout_edge *le = malloc(sizeof *le * (E ? E : 1));
  out->row_offsets = calloc((size_t)V + 1, sizeof(uint64_t));
  out->destinations = malloc(sizeof(uint32_t) * (E ? E : 1));
  out->weights = malloc(sizeof(float) * (E ? E : 1));
  out->edge_ids = malloc(sizeof(uint64_t) * (E ? E : 1));
  for (uint64_t e = 0; e < E; ++e) {
    uint32_t s = out->old_to_new[edges[e].source];
    uint32_t d = out->old_to_new[edges[e].destination];
    le[e].major = s;
    le[e].tile = out->owner_tiles[d];
    le[e].minor = d;
    le[e].weight = canonicalize_weight(edges[e].weight);
    le[e].edge_id = edge_id_of_input[e];
  }
  qsort(le, E, sizeof *le, cmp_layout);
  for (uint64_t e = 0; e < E; ++e) {
    out->row_offsets[le[e].major + 1]++;
    out->destinations[e] = le[e].minor;
    out->weights[e] = le[e].weight;
    out->edge_ids[e] = le[e].edge_id;
  }
  for (uint32_t v = 1; v <= V; ++v) out->row_offsets[v] += out->row_offsets[v - 1];

  /* Incoming CSC by (destination, source tile, source, EdgeId). */
  out->column_offsets = calloc((size_t)V + 1, sizeof(uint64_t));
  out->sources = malloc(sizeof(uint32_t) * (E ? E : 1));
  out->in_weights = malloc(sizeof(float) * (E ? E : 1));
  out->in_edge_ids = malloc(sizeof(uint64_t) * (E ? E : 1));
  for (uint64_t e = 0; e < E; ++e) {
    uint32_t s = out->old_to_new[edges[e].source];
    uint32_t d = out->old_to_new[edges[e].destination];
    le[e].major = d;
    le[e].tile = out->owner_tiles[s];
    le[e].minor = s;
    le[e].weight = canonicalize_weight(edges[e].weight);
    le[e].edge_id = edge_id_of_input[e];
  }
  qsort(le, E, sizeof *le, cmp_layout);
  for (uint64_t e = 0; e < E; ++e) {
    out->column_offsets[le[e].major + 1]++;
    out->sources[e] = le[e].minor;
    out->in_weights[e] = le[e].weight;
    out->in_edge_ids[e] = le[e].edge_id;
  }
  for (uint32_t v = 1; v <= V; ++v) out->column_offsets[v] += out->column_offsets[v - 1];
  free(le);
  free(edge_id_of_input);
  return ORC_OK;
}

/* device_layout.cpp:186-251: maximal runs of one remote tile per row/column. */
static int build_runs(uint32_t V, const uint64_t *offsets, const uint32_t *ends,
                      const uint32_t *owner, uint64_t **row_runs,
                      uint64_t **run_edges, uint32_t **run_tiles, uint64_t *count) {
  const uint64_t E = offsets[V];
  *row_runs = calloc((size_t)V + 1, sizeof(uint64_t));
  *run_edges = malloc(sizeof(uint64_t) * (E + 1));
  *run_tiles = malloc(sizeof(uint32_t) * (E ? E : 1));
  if (!*row_runs || !*run_edges || !*run_tiles) return ORC_NOMEM;
  uint64_t runs = 0;
  (*run_edges)[0] = 0;
  for (uint32_t v = 0; v < V; ++v) {
    uint64_t p = offsets[v];
    while (p < offsets[v + 1]) {
      uint32_t tile = owner[ends[p]];
      ++p;
      while (p < offsets[v + 1] && owner[ends[p]] == tile) ++p;
      (*run_tiles)[runs] = tile;
      (*run_edges)[++runs] = p;
    }
    (*row_runs)[v + 1] = runs;
  }
  *count = runs;
  return ORC_OK;
}

int orc_tile_runs(const orc_graph *g, orc_tile_runs_t *out) {
  memset(out, 0, sizeof *out);
  int rc = build_runs(g->vertex_count, g->row_offsets, g->destinations, g->owner_tiles,
                      &out->csr_row_run_offsets, &out->csr_run_edge_offsets,
                      &out->csr_run_tiles, &out->csr_run_count);
  if (rc) return rc;
  return build_runs(g->vertex_count, g->column_offsets, g->sources, g->owner_tiles,
                    &out->csc_column_run_offsets, &out->csc_run_edge_offsets,
                    &out->csc_run_tiles, &out->csc_run_count);
}

void orc_tile_runs_free(orc_tile_runs_t *r) {
  if (!r) return;
  free(r->csr_row_run_offsets);
  free(r->csr_run_edge_offsets);
  free(r->csr_run_tiles);
  free(r->csc_column_run_offsets);
  free(r->csc_run_edge_offsets);
  free(r->csc_run_tiles);
  memset(r, 0, sizeof *r);
}

/* ---------------------------------------------------------------------------
 * SSSP.
 * ------------------------------------------------------------------------- */

static int selected(const uint32_t *owner, const uint8_t *tile_selected, uint32_t v) {
  return tile_selected == NULL || tile_selected[owner[v]] != 0;
}

/* sssp_common.hpp:34-45: +inf everywhere, 0 at every (deduplicated) source. */
static int init_distances(uint32_t V, const uint32_t *sources, uint32_t count,
                          float *distances) {
  if (count == 0) return ORC_INVALID;
  for (uint32_t v = 0; v < V; ++v) distances[v] = INFINITY;
  for (uint32_t i = 0; i < count; ++i) {
    if (sources[i] >= V) return ORC_INVALID;
    distances[sources[i]] = 0.0f;
  }
  return ORC_OK;
}

typedef struct {
  float distance;
  uint32_t vertex;
} heap_entry;

/* reference_dijkstra.cpp:19-28: smaller distance first, then smaller vertex. */
static int heap_less(heap_entry a, heap_entry b) {
  if (a.distance != b.distance) return a.distance < b.distance;
  return a.vertex < b.vertex;
}

typedef struct {
  heap_entry *items;
  size_t size, capacity;
} heap;

static int heap_push(heap *h, heap_entry e) {
  if (h->size == h->capacity) {
    size_t cap = h->capacity ? h->capacity * 2 : 1024;
    heap_entry *n = realloc(h->items, cap * sizeof *n);
    if (!n) return ORC_NOMEM;
    h->items = n;
    h->capacity = cap;
  }
  size_t i = h->size++;
  while (i > 0) {
    size_t parent = (i - 1) / 2;
    if (!heap_less(e, h->items[parent])) break;
    h->items[i] = h->items[parent];
    i = parent;
  }
  h->items[i] = e;
  return ORC_OK;
}

static heap_entry heap_pop(heap *h) {
  heap_entry top = h->items[0];
  heap_entry last = h->items[--h->size];
  size_t i = 0;
  for (;;) {
    size_t child = 2 * i + 1;
    if (child >= h->size) break;
    if (child + 1 < h->size && heap_less(h->items[child + 1], h->items[child])) ++child;
    if (!heap_less(h->items[child], last)) break;
    h->items[i] = h->items[child];
    i = child;
  }
  if (h->size) h->items[i] = last;
  return top;
}

/* reference_dijkstra.cpp:32-63 on the induced subgraph of the selected tiles
 * (query.cpp:286-362 keeps vertices of selected tiles and edges whose two
 * endpoints are selected; masking the global CSR admits the same edges). */
static int cmp_u32(const void *pa, const void *pb) {
  uint32_t a = *(const uint32_t *)pa, b = *(const uint32_t *)pb;
  return a < b ? -1 : (a > b);
}

/* sssp_common.hpp:14-32: sources sorted and deduplicated. */
static uint32_t *canonical_sources(const uint32_t *sources, uint32_t count, uint32_t *out_count) {
  uint32_t *c = malloc(sizeof(uint32_t) * (count ? count : 1));
  if (!c) return NULL;
  memcpy(c, sources, sizeof(uint32_t) * count);
  qsort(c, count, sizeof(uint32_t), cmp_u32);
  uint32_t n = 0;
  for (uint32_t i = 0; i < count; ++i)
    if (n == 0 || c[n - 1] != c[i]) c[n++] = c[i];
  *out_count = n;
  return c;
}

int orc_dijkstra(uint32_t V, const uint64_t *row_offsets, const uint32_t *destinations,
                 const float *weights, const uint32_t *owner, const uint8_t *tile_selected,
                 const uint32_t *sources, uint32_t source_count, float *distances,
                 uint64_t *relaxations_out) {
  int rc = init_distances(V, sources, source_count, distances);
  if (rc) return rc;
  uint32_t n_src = 0;
  uint32_t *src = canonical_sources(sources, source_count, &n_src);
  if (!src) return ORC_NOMEM;
  heap h = {0};
  uint64_t relaxations = 0;
  for (uint32_t i = 0; i < n_src; ++i) {
    if (!selected(owner, tile_selected, src[i])) {
      free(h.items);
      free(src);
      return ORC_INVALID;
    }
    if ((rc = heap_push(&h, (heap_entry){0.0f, src[i]}))) return rc;
  }
  free(src);
  while (h.size) {
    heap_entry e = heap_pop(&h);
    if (e.distance != distances[e.vertex]) continue;
    for (uint64_t p = row_offsets[e.vertex]; p < row_offsets[e.vertex + 1]; ++p) {
      uint32_t d = destinations[p];
      if (!selected(owner, tile_selected, d)) continue;
      ++relaxations;
      float candidate = e.distance + weights[p];
      if (candidate < distances[d]) {
        distances[d] = candidate;
        if ((rc = heap_push(&h, (heap_entry){candidate, d}))) return rc;
      }
    }
  }
  free(h.items);
  if (relaxations_out) *relaxations_out = relaxations;
  return ORC_OK;
}

/* reference_bellman_ford.cpp:14-47: synchronous rounds over a copied column. */
int orc_bellman_ford(uint32_t V, const uint64_t *row_offsets, const uint32_t *destinations,
                     const float *weights, const uint32_t *owner, const uint8_t *tile_selected,
                     const uint32_t *sources, uint32_t source_count, float *distances) {
  int rc = init_distances(V, sources, source_count, distances);
  if (rc) return rc;
  float *next = malloc(sizeof(float) * (V ? V : 1));
  if (!next) return ORC_NOMEM;
  for (uint64_t round = 1; round < V; ++round) {
    memcpy(next, distances, sizeof(float) * V);
    int changed = 0;
    for (uint32_t s = 0; s < V; ++s) {
      float ds = distances[s];
      if (!isfinite(ds) || !selected(owner, tile_selected, s)) continue;
      for (uint64_t p = row_offsets[s]; p < row_offsets[s + 1]; ++p) {
        uint32_t d = destinations[p];
        if (!selected(owner, tile_selected, d)) continue;
        float candidate = ds + weights[p];
        if (candidate < next[d]) {
          next[d] = candidate;
          changed = 1;
        }
      }
    }
    memcpy(distances, next, sizeof(float) * V);
    if (!changed) break;
  }
  free(next);
  return ORC_OK;
}

/* Two-column Jacobi pull (jacobi_pull.hpp:153-220 controller,
 * batched_jacobi_pull.cpp:715-800 per-destination strict-< minimum). */
int64_t orc_jacobi_rounds(uint32_t V, const uint64_t *column_offsets,
                          const uint32_t *sources_csc, const float *in_weights,
                          const uint32_t *owner, const uint8_t *tile_selected,
                          const uint32_t *sources, uint32_t source_count,
                          uint64_t maximum_rounds, float *distances) {
  if (init_distances(V, sources, source_count, distances)) return -1;
  float *next = malloc(sizeof(float) * (V ? V : 1));
  if (!next) return -2;
  int64_t rounds = 0;
  for (;;) {
    ++rounds;
    int changed = 0;
    for (uint32_t v = 0; v < V; ++v) {
      float best = distances[v];
      if (selected(owner, tile_selected, v)) {
        for (uint64_t p = column_offsets[v]; p < column_offsets[v + 1]; ++p) {
          uint32_t u = sources_csc[p];
          if (!selected(owner, tile_selected, u)) continue;
          float candidate = distances[u] + in_weights[p];
          if (candidate < best) best = candidate;
        }
      }
      if (best != distances[v]) changed = 1;
      next[v] = best;
    }
    memcpy(distances, next, sizeof(float) * V);
    if (!changed) break;
    if (maximum_rounds && (uint64_t)rounds == maximum_rounds) break;
  }
  free(next);
  return rounds;
}

static int is_source(const uint32_t *sources, uint32_t count, uint32_t v) {
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] == v) return 1;
  return 0;
}

/* The first candidate of compact_paths.cpp:86-115 / path_reconstruct.cpp:44-78:
 * the smallest stable EdgeId among tight admitted incoming edges. */
int orc_predecessors(uint32_t V, const uint64_t *column_offsets, const uint32_t *sources_csc,
                     const float *in_weights, const uint64_t *in_edge_ids,
                     const uint32_t *owner, const uint8_t *tile_selected,
                     const float *distances, const uint32_t *sources, uint32_t source_count,
                     uint32_t *pred_vertex, uint64_t *pred_edge) {
  uint8_t *src_mask = calloc(V ? V : 1, 1);
  if (!src_mask) return ORC_NOMEM;
  for (uint32_t i = 0; i < source_count; ++i) src_mask[sources[i]] = 1;
  for (uint32_t v = 0; v < V; ++v) {
    pred_vertex[v] = ORC_NONE32;
    pred_edge[v] = UINT64_MAX;
    if (src_mask[v] || !selected(owner, tile_selected, v) || !isfinite(distances[v]))
      continue;
    for (uint64_t p = column_offsets[v]; p < column_offsets[v + 1]; ++p) {
      uint32_t u = sources_csc[p];
      if (!selected(owner, tile_selected, u) || !isfinite(distances[u])) continue;
      if (distances[u] + in_weights[p] == distances[v] && in_edge_ids[p] < pred_edge[v]) {
        pred_edge[v] = in_edge_ids[p];
        pred_vertex[v] = u;
      }
    }
  }
  free(src_mask);
  return ORC_OK;
}

typedef struct {
  uint64_t edge_id;
  uint32_t predecessor;
} choice;

static int cmp_choice(const void *pa, const void *pb) {
  const choice *a = pa, *b = pb;
  return cmp_u64(a->edge_id, b->edge_id);
}

typedef struct {
  uint32_t vertex;
  choice *choices;
  uint32_t count, cursor;
} frame;

static choice *tight_choices(const uint64_t *column_offsets, const uint32_t *sources_csc,
                             const float *in_weights, const uint64_t *in_edge_ids,
                             const uint32_t *owner, const uint8_t *tile_selected,
                             const float *distances, uint32_t v, uint32_t *count) {
  uint64_t degree = column_offsets[v + 1] - column_offsets[v];
  choice *out = malloc(sizeof *out * (degree ? degree : 1));
  uint32_t n = 0;
  for (uint64_t p = column_offsets[v]; p < column_offsets[v + 1]; ++p) {
    uint32_t u = sources_csc[p];
    if (!selected(owner, tile_selected, u) || !isfinite(distances[u])) continue;
    if (distances[u] + in_weights[p] == distances[v]) {
      out[n].edge_id = in_edge_ids[p];
      out[n].predecessor = u;
      ++n;
    }
  }
  qsort(out, n, sizeof *out, cmp_choice);
  *count = n;
  return out;
}

/* compact_paths.cpp:117-206 (equivalently path_reconstruct.cpp:113-191). */
int orc_reconstruct(uint32_t V, const uint64_t *column_offsets, const uint32_t *sources_csc,
                    const float *in_weights, const uint64_t *in_edge_ids,
                    const uint32_t *owner, const uint8_t *tile_selected,
                    const float *distances, const uint32_t *sources, uint32_t source_count,
                    uint32_t target, uint32_t capacity, uint32_t *path_vertices,
                    uint64_t *path_edge_ids, uint32_t *path_length) {
  *path_length = 0;
  if (target >= V) return ORC_INVALID;
  if (!isfinite(distances[target])) return 1;
  uint8_t *on_path = calloc(V ? V : 1, 1);
  frame *stack = malloc(sizeof *stack * (V ? V : 1));
  uint32_t *rev_vertices = malloc(sizeof(uint32_t) * (V + 1));
  uint64_t *rev_edges = malloc(sizeof(uint64_t) * (V + 1));
  if (!on_path || !stack || !rev_vertices || !rev_edges) return ORC_NOMEM;
  uint32_t depth = 0, nv = 0, ne = 0;
  int rc = ORC_NO_TIGHT_PATH;
  rev_vertices[nv++] = target;
  on_path[target] = 1;
  stack[depth].vertex = target;
  stack[depth].cursor = 0;
  stack[depth].choices = tight_choices(column_offsets, sources_csc, in_weights, in_edge_ids,
                                       owner, tile_selected, distances, target,
                                       &stack[depth].count);
  ++depth;
  while (depth) {
    frame *f = &stack[depth - 1];
    if (is_source(sources, source_count, f->vertex)) {
      if (ne > capacity || nv > capacity + 1) {
        rc = ORC_CAPACITY;
        break;
      }
      for (uint32_t i = 0; i < nv; ++i) path_vertices[i] = rev_vertices[nv - 1 - i];
      for (uint32_t i = 0; i < ne; ++i) path_edge_ids[i] = rev_edges[ne - 1 - i];
      *path_length = ne;
      rc = ORC_OK;
      break;
    }
    int advanced = 0;
    while (f->cursor < f->count) {
      choice c = f->choices[f->cursor++];
      if (on_path[c.predecessor]) continue;
      rev_edges[ne++] = c.edge_id;
      rev_vertices[nv++] = c.predecessor;
      on_path[c.predecessor] = 1;
      stack[depth].vertex = c.predecessor;
      stack[depth].cursor = 0;
      stack[depth].choices =
          tight_choices(column_offsets, sources_csc, in_weights, in_edge_ids, owner,
                        tile_selected, distances, c.predecessor, &stack[depth].count);
      ++depth;
      advanced = 1;
      break;
    }
    if (advanced) continue;
    on_path[f->vertex] = 0;
    free(f->choices);
    --depth;
    --nv;
    if (depth) --ne;
  }
  for (uint32_t i = 0; i < depth; ++i) free(stack[i].choices);
  free(on_path);
  free(stack);
  free(rev_vertices);
  free(rev_edges);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Exactness certificate of one query's labels over its selected tiles
 * (no reference counterpart: it checks the result that
 * reference_dijkstra.cpp:32-63 defines, without running Dijkstra).
 *
 * labels[] holds the query's labels tile by tile in the order of tiles[]
 * (sorted, unique): tile tiles[i] contributes tile_vertex_offsets[t+1] -
 * tile_vertex_offsets[t] labels.  Admitted edges are those whose two
 * endpoints lie in selected tiles (query.cpp:286-362).  The labels are the
 * shortest-path fixed point exactly when
 *   (1) every source is +0 and every label is +0, finite positive or +inf;
 *   (2) no admitted edge lowers a label: fl(d[u] + w) >= d[v];
 *   (3) every finite non-source label d[v] has a tight admitted in-edge,
 *       fl(d[u] + w) == d[v], with d[u] < d[v].
 * (3) with the strict decrease makes every finite label the left-to-right
 * fp32 sum of an admitted path from a source; (2) with the monotonicity of
 * fl(x + w) in x makes it no larger than any admitted path's sum; an
 * unreached +inf label next to a finite one fails (2).  Absorption
 * (fl(d[u] + w) == d[u]) can never count as tight, so zero or tiny weights
 * cannot certify a false label.  Returns 0 or the failed rule (1-3) and, in
 * *where, the compact index of the offending label.
 * ------------------------------------------------------------------------- */
int orc_certify_region(uint32_t tile_count, const uint32_t *tile_vertex_offsets,
                       const uint32_t *column_offsets, const uint32_t *sources_csc,
                       const float *in_weights, const uint32_t *owner_tiles,
                       uint32_t n_tiles, const uint32_t *tiles, const float *labels,
                       uint32_t source_count, const uint32_t *sources, uint64_t *where) {
  int64_t *base = malloc(sizeof(int64_t) * (tile_count ? tile_count : 1));
  if (!base) return ORC_NOMEM;
  for (uint32_t t = 0; t < tile_count; ++t) base[t] = -1;
  uint64_t n = 0;
  for (uint32_t i = 0; i < n_tiles; ++i) {
    const uint32_t t = tiles[i];
    if (t >= tile_count || (i && tiles[i - 1] >= t)) {
      free(base);
      return ORC_INVALID;
    }
    base[t] = (int64_t)n;
    n += tile_vertex_offsets[t + 1] - tile_vertex_offsets[t];
  }
  int rule = 0;
  uint64_t at = 0;
  uint8_t *is_source = calloc(n ? n : 1, 1);
  if (!is_source) {
    free(base);
    return ORC_NOMEM;
  }
  for (uint32_t i = 0; i < source_count && !rule; ++i) {
    const uint32_t s = sources[i];
    const uint32_t t = owner_tiles[s];
    if (base[t] < 0) {
      rule = 1;
      break;
    }
    const uint64_t k = (uint64_t)base[t] + (s - tile_vertex_offsets[t]);
    is_source[k] = 1;
    if (float_bits(labels[k]) != 0u) {
      rule = 1;
      at = k;
    }
  }
  for (uint64_t k = 0; k < n && !rule; ++k) {
    const uint32_t b = float_bits(labels[k]);
    if (b > 0x7F800000u) {
      rule = 1;
      at = k;
    }
  }
  for (uint32_t i = 0; i < n_tiles && !rule; ++i) {
    const uint32_t t = tiles[i];
    const uint32_t v0 = tile_vertex_offsets[t], v1 = tile_vertex_offsets[t + 1];
    for (uint32_t v = v0; v < v1 && !rule; ++v) {
      const uint64_t kv = (uint64_t)base[t] + (v - v0);
      const float dv = labels[kv];
      int tight = 0;
      for (uint32_t e = column_offsets[v]; e < column_offsets[v + 1]; ++e) {
        const uint32_t u = sources_csc[e];
        const uint32_t tu = owner_tiles[u];
        if (base[tu] < 0) continue;
        const float du = labels[(uint64_t)base[tu] + (u - tile_vertex_offsets[tu])];
        const float cand = du + in_weights[e];
        if (cand < dv) {
          rule = 2;
          at = kv;
          break;
        }
        if (cand == dv && du < dv) tight = 1;
      }
      if (!rule && dv != INFINITY && !is_source[kv] && !tight) {
        rule = 3;
        at = kv;
      }
    }
  }
  free(is_source);
  free(base);
  if (where) *where = at;
  return rule;
}
