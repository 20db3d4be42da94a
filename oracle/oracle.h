/*
 * oracle.h -- CPU restatement of the bfnew shortest-path path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the tests,
 * the smoke entry point and the CPU-baseline leg of the benchmark compare the
 * GPU engine against.  The product library never links, loads or calls it,
 * and nothing here is ever the thing being measured on the GPU path.
 *
 * Every routine restates one reference routine (cited per function in
 * oracle.c) in plain C11 with IEEE fp32 arithmetic, no FMA contraction
 * (built with -ffp-contract=off) and the reference's tie-breaks.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_build_graph    -- reorder literals from tests/reorder_test.cpp:86-160
 *                         (golden fixture), and bit-equality with the
 *                         reference's own build_spatially_ordered_graph
 *                         compiled into oracle/_ref.
 *   orc_tile_runs      -- run counts/lengths from
 *                         tests/device_layout_test.cpp:124-183 and equality
 *                         with the reference build_tile_run_layout (_ref).
 *   orc_dijkstra       -- literal distances tests/sssp_test.cpp:108-117,
 *                         equality with orc_bellman_ford (a different
 *                         algorithm), and with the reference dijkstra (_ref).
 *   orc_jacobi_rounds  -- literal round counts 33/2/3/3/3/1 from
 *                         tests/jacobi_fixture_suite.hpp:195-355.
 *   orc_predecessors /
 *   orc_reconstruct    -- literal paths tests/sssp_test.cpp:174-227 and
 *                         tests/compact_paths_test.cpp:114-205, and the
 *                         reference reconstruct_path_from_distances (_ref).
 *   orc_certify_region -- accepts the oracle's own Dijkstra labels on every
 *                         fixture and synthetic case, and rejects each of a
 *                         set of planted mistakes (a label one ulp low or
 *                         high, a spurious finite label on an unreachable
 *                         cycle, an absorbed zero-weight edge, a nonzero
 *                         source) in tests/test_oracle.py.
 */
#ifndef BFNEW_ORACLE_H
#define BFNEW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t x;
  int32_t y;
  uint32_t has_location;
  uint32_t resource_class;
} orc_vertex;

typedef struct {
  uint32_t source;
  uint32_t destination;
  float weight;
  uint32_t provenance_domain;
  uint32_t provenance_kind;
  uint32_t pad;
  uint64_t provenance_record;
} orc_edge;

typedef struct {
  int32_t origin_x;
  int32_t origin_y;
  uint32_t tile_width;
  uint32_t tile_height;
} orc_spatial_config;

/* A weighted graph in the reference's reordered numbering.  All arrays are
 * malloc'ed by orc_build_graph and released by orc_graph_free. */
typedef struct {
  uint32_t vertex_count;
  uint64_t edge_count;
  uint32_t tile_count; /* located tiles + 1 spill tile, or 1 when unordered */
  uint64_t *row_offsets;    /* [V+1] */
  uint32_t *destinations;   /* [E]   */
  float *weights;           /* [E]   */
  uint64_t *edge_ids;       /* [E]   */
  uint64_t *column_offsets; /* [V+1] */
  uint32_t *sources;        /* [E]   */
  float *in_weights;        /* [E]   */
  uint64_t *in_edge_ids;    /* [E]   */
  uint32_t *old_to_new;     /* [V]   */
  uint32_t *new_to_old;     /* [V]   */
  uint32_t *owner_tiles;    /* [V]   */
  int64_t *tile_x;          /* [tile_count] */
  int64_t *tile_y;          /* [tile_count] */
  uint32_t *tile_has_location;        /* [tile_count] */
  uint64_t *tile_vertex_offsets;      /* [tile_count+1] */
} orc_graph;

typedef struct {
  uint64_t *csr_row_run_offsets;   /* [V+1] */
  uint64_t *csr_run_edge_offsets;  /* [Rcsr+1] */
  uint32_t *csr_run_tiles;         /* [Rcsr] */
  uint64_t csr_run_count;
  uint64_t *csc_column_run_offsets;
  uint64_t *csc_run_edge_offsets;
  uint32_t *csc_run_tiles;
  uint64_t csc_run_count;
} orc_tile_runs_t;

/* Return codes. */
enum {
  ORC_OK = 0,
  ORC_INVALID = -1,
  ORC_NOMEM = -2,
  ORC_NO_TIGHT_PATH = -3,
  ORC_CAPACITY = -4
};

/* config == NULL builds the plain (unordered) weighted graph. */
int orc_build_graph(const orc_vertex *vertices, uint32_t vertex_count,
                    const orc_edge *edges, uint64_t edge_count,
                    const orc_spatial_config *config, orc_graph *out);
void orc_graph_free(orc_graph *graph);

int orc_tile_runs(const orc_graph *graph, orc_tile_runs_t *out);
void orc_tile_runs_free(orc_tile_runs_t *runs);

/* Multi-source Dijkstra on the subgraph induced by the selected tiles
 * (tile_selected == NULL selects every tile).  distances[V] out. */
int orc_dijkstra(uint32_t vertex_count, const uint64_t *row_offsets,
                 const uint32_t *destinations, const float *weights,
                 const uint32_t *owner_tiles, const uint8_t *tile_selected,
                 const uint32_t *sources, uint32_t source_count,
                 float *distances, uint64_t *relaxations_out);

int orc_bellman_ford(uint32_t vertex_count, const uint64_t *row_offsets,
                     const uint32_t *destinations, const float *weights,
                     const uint32_t *owner_tiles, const uint8_t *tile_selected,
                     const uint32_t *sources, uint32_t source_count,
                     float *distances);

/* Synchronous (two-column) Jacobi pull over the incoming CSC restricted to
 * the selected tiles.  Returns the controller's rounds_completed: the number
 * of rounds with a change plus the final no-change round, capped by
 * maximum_rounds (0 = unlimited).  distances[V] out. */
int64_t orc_jacobi_rounds(uint32_t vertex_count, const uint64_t *column_offsets,
                          const uint32_t *sources_csc, const float *in_weights,
                          const uint32_t *owner_tiles,
                          const uint8_t *tile_selected,
                          const uint32_t *sources, uint32_t source_count,
                          uint64_t maximum_rounds, float *distances);

/* Smallest-stable-EdgeId tight admitted predecessor of every vertex, or
 * 0xFFFFFFFF for sources, unreached and unselected vertices.  The matching
 * EdgeId goes to pred_edge (UINT64_MAX when none). */
int orc_predecessors(uint32_t vertex_count, const uint64_t *column_offsets,
                     const uint32_t *sources_csc, const float *in_weights,
                     const uint64_t *in_edge_ids, const uint32_t *owner_tiles,
                     const uint8_t *tile_selected, const float *distances,
                     const uint32_t *sources, uint32_t source_count,
                     uint32_t *pred_vertex, uint64_t *pred_edge);

/* Backtracking tight-edge DFS from target to any source, trying tight
 * admitted predecessors in stable EdgeId order and skipping vertices already
 * on the path.  Writes the path source->target.  Returns ORC_OK,
 * ORC_NO_TIGHT_PATH, ORC_CAPACITY, or 1 when the target is unreachable. */
int orc_reconstruct(uint32_t vertex_count, const uint64_t *column_offsets,
                    const uint32_t *sources_csc, const float *in_weights,
                    const uint64_t *in_edge_ids, const uint32_t *owner_tiles,
                    const uint8_t *tile_selected, const float *distances,
                    const uint32_t *sources, uint32_t source_count,
                    uint32_t target, uint32_t capacity, uint32_t *path_vertices,
                    uint64_t *path_edge_ids, uint32_t *path_length);

/* Size-independent exactness certificate of one query's labels, given tile
 * by tile over its sorted selected tiles (see oracle.c for the argument).
 * Returns 0 when the labels are the exact fp32 shortest-path fixed point of
 * the admitted subgraph, else the failed rule 1-3 with the compact index of
 * the offending label in *where. */
int orc_certify_region(uint32_t tile_count, const uint32_t *tile_vertex_offsets,
                       const uint32_t *column_offsets, const uint32_t *sources_csc,
                       const float *in_weights, const uint32_t *owner_tiles,
                       uint32_t n_tiles, const uint32_t *tiles, const float *labels,
                       uint32_t source_count, const uint32_t *sources, uint64_t *where);

#ifdef __cplusplus
}
#endif

#endif
