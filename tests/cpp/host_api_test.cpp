// host_api_test.cpp -- host-only parts of the reference-named C++ API.
//
//   host_api_test tiles            tile directory of the spatial fixture
//                                  (reference tests/tile_test.cpp:80-260
//                                  literals) and validation of planted
//                                  mistakes
//   host_api_test paths            path reconstruction from labels on the
//                                  reference tests' literal fixtures
//                                  (sssp_test.cpp:108-227,
//                                  compact_paths_test.cpp:114-205)
//   host_api_test workspace        batch_workspace byte model, reservations
//                                  and decisions on the reference's workspace
//                                  fixture (tests/batch_workspace_test.cpp:
//                                  77-960 literal totals)
//   host_api_test paths-file IN OUT
//                                  reconstructs paths for the graph, labels
//                                  and targets in IN (written by
//                                  tests/test_host_api.py from the oracle) and
//                                  writes them to OUT for comparison with the
//                                  oracle's reconstruction
//
// Labels of the literal fixtures come from a Bellman-Ford written here (test
// code: the library has no CPU shortest-path engine).
#include "bfnew/batch_workspace.hpp"
#include "bfnew/batched_expansion.hpp"
#include "bfnew/compact_paths.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"
#include "bfnew/spatial.hpp"
#include "bfnew/sssp.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

using namespace bfnew;

namespace {

int failures = 0;

void check(bool ok, const char* what, int line) {
  if (!ok) {
    std::fprintf(stderr, "FAIL line %d: %s\n", line, what);
    ++failures;
  }
}
#define CHECK(x) check((x), #x, __LINE__)

template <typename E, typename F>
bool throws(F&& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

constexpr float kInf = std::numeric_limits<float>::infinity();
const ResourceClassId kClass1{1U}, kClass2{2U};

EdgeInputRecord edge(std::uint32_t s, std::uint32_t d, float w, std::uint64_t record) {
  return EdgeInputRecord{VertexId{s}, VertexId{d}, w,
                         PhysicalProvenance{provenance_domain::synthetic,
                                            provenance_kind::synthetic_edge, record}};
}

// Labels of the selected-region fixed point (Bellman-Ford to convergence);
// vertices outside the region stay +inf.
std::vector<float> region_labels(const WeightedGraph& g, std::span<const VertexId> sources,
                                 const std::vector<TileId>* tiles) {
  const std::uint32_t V = g.vertex_count();
  std::vector<std::uint8_t> in_region(V, 1U);
  if (tiles) {
    for (std::uint32_t v = 0; v < V; ++v)
      in_region[v] = std::binary_search(tiles->begin(), tiles->end(), g.owner_tiles()[v]);
  }
  std::vector<float> d(V, kInf);
  for (const VertexId s : sources) d[s.value()] = 0.0F;
  const OutgoingCsrView out = g.outgoing();
  for (bool changed = true; changed;) {
    changed = false;
    for (std::uint32_t u = 0; u < V; ++u) {
      if (!in_region[u] || !std::isfinite(d[u])) continue;
      for (auto p = out.row_offsets[u]; p < out.row_offsets[u + 1]; ++p) {
        const std::uint32_t v = out.destinations[p].value();
        const float c = d[u] + out.weights[p];
        if (in_region[v] && c < d[v]) {
          d[v] = c;
          changed = true;
        }
      }
    }
  }
  return d;
}

std::vector<std::uint32_t> old_ids(const WeightedGraph& g, const std::vector<VertexId>& vs) {
  std::vector<std::uint32_t> out;
  for (const VertexId v : vs) out.push_back(g.new_to_old()[v.value()].value());
  return out;
}

// ---------------------------------------------------------------------------
// Tile directory.
// ---------------------------------------------------------------------------
InputGraph spatial_fixture() {  // graph_fixtures.hpp:72-110
  std::vector<VertexMetadata> v{
      VertexMetadata::located(20, 1, kClass2), VertexMetadata::located(1, 1, kClass1),
      VertexMetadata::located(11, 1, kClass1), VertexMetadata::located(2, 1, kClass2),
      VertexMetadata::located(12, 1, kClass1), VertexMetadata::located(21, 1, kClass1),
      VertexMetadata::unlocated(kClass2),      VertexMetadata::located(3, 1, kClass1),
      VertexMetadata::located(3, 1, kClass1),  VertexMetadata::unlocated(kClass1)};
  std::vector<EdgeInputRecord> e{edge(1, 0, 0.5F, 100),   edge(1, 2, 0.75F, 101),
                                 edge(1, 3, 1.25F, 102),  edge(1, 4, 1.5F, 103),
                                 edge(1, 5, 2.25F, 104),  edge(1, 7, 0.0F, 105),
                                 edge(1, 8, 0.25F, 106),  edge(1, 9, 3.5F, 107),
                                 edge(3, 1, 0.625F, 108), edge(6, 1, 4.0F, 109),
                                 edge(2, 2, 0.125F, 110), edge(5, 0, 1.75F, 111)};
  return InputGraph{std::move(v), std::move(e)};
}

template <typename T>
std::vector<T> group(std::span<const EdgeOffset> off, std::span<const T> values, std::size_t t) {
  return std::vector<T>(values.begin() + off[t], values.begin() + off[t + 1]);
}

EdgeId edge_of_record(const WeightedGraph& g, std::uint64_t record) {
  for (std::size_t id = 0; id < g.edge_provenance().size(); ++id)
    if (g.edge_provenance()[id].source_record == record) return EdgeId{id};
  throw std::logic_error{"record not in graph"};
}

template <typename T>
bool contains(const std::vector<T>& xs, T x) {
  return std::find(xs.begin(), xs.end(), x) != xs.end();
}

void tiles() {
  const UniformGridPartitioner partitioner{SpatialOrderConfig{0, 0, 10U, 10U}};
  const PartitionedGraph pg = partitioner.partition(spatial_fixture());
  const WeightedGraph& g = pg.graph;
  const TileDirectory& d = pg.tiles;
  using Code = TileDirectoryValidationErrorCode;

  // tile_test.cpp:80-134: counts and classification.
  CHECK(g.has_spatial_ordering());
  CHECK(d.tile_count() == 4U);
  CHECK(d.spill_tile() == TileId{3U});
  CHECK(validate_tile_directory(g, d).ok());
  CHECK(d.internal_edge_ids().size() == 6U);
  CHECK(d.outgoing_cross_edge_ids().size() == 6U);
  CHECK(d.incoming_cross_edge_ids().size() == 6U);
  CHECK(d.halo_vertices().size() == 9U);
  std::vector<int> classified(g.edge_count(), 0);
  for (const EdgeId e : d.internal_edge_ids()) ++classified[e.value()];
  for (const EdgeId e : d.outgoing_cross_edge_ids()) ++classified[e.value()];
  CHECK(std::all_of(classified.begin(), classified.end(), [](int c) { return c == 1; }));
  for (std::size_t t = 0; t < d.tile_count(); ++t)
    for (const VertexId h : group(d.halo_vertex_offsets(), d.halo_vertices(), t))
      CHECK(h.value() < g.vertex_count() && g.owner_tiles()[h.value()].value() != t);

  // tile_test.cpp:136-201: neighbours, long edges and the spill tile.
  auto nbr = [&](std::size_t t) { return group(d.neighbor_tile_offsets(), d.neighbor_tiles(), t); };
  CHECK(nbr(0) == (std::vector<TileId>{TileId{1U}, TileId{3U}}));
  CHECK(nbr(1) == (std::vector<TileId>{TileId{0U}, TileId{2U}}));
  CHECK(nbr(2) == (std::vector<TileId>{TileId{1U}}));
  CHECK(nbr(3) == (std::vector<TileId>{TileId{0U}}));
  const EdgeId adjacent = edge_of_record(g, 101), long_edge = edge_of_record(g, 100),
               to_spill = edge_of_record(g, 107);
  const auto out0 = group(d.outgoing_cross_edge_offsets(), d.outgoing_cross_edge_ids(), 0);
  CHECK(contains(out0, adjacent) && contains(out0, long_edge) && contains(out0, to_spill));
  CHECK(contains(group(d.incoming_cross_edge_offsets(), d.incoming_cross_edge_ids(), 1), adjacent));
  CHECK(contains(group(d.incoming_cross_edge_offsets(), d.incoming_cross_edge_ids(), 2), long_edge));
  CHECK(contains(group(d.incoming_cross_edge_offsets(), d.incoming_cross_edge_ids(), 3), to_spill));
  CHECK(contains(group(d.halo_vertex_offsets(), d.halo_vertices(), 0), g.old_to_new()[0]));
  CHECK(contains(group(d.halo_vertex_offsets(), d.halo_vertices(), 2), g.old_to_new()[1]));

  // tile_test.cpp:203-228: diagonal tiles are neighbours without an edge.
  {
    const InputGraph diag{{VertexMetadata::located(0, 0, kClass1), VertexMetadata::located(10, 10, kClass1)},
                          {}};
    const PartitionedGraph dp = partitioner.partition(diag);
    const auto& dd = dp.tiles;
    CHECK(group(dd.neighbor_tile_offsets(), dd.neighbor_tiles(), 0) == (std::vector<TileId>{TileId{1U}}));
    CHECK(group(dd.neighbor_tile_offsets(), dd.neighbor_tiles(), 1) == (std::vector<TileId>{TileId{0U}}));
    CHECK(group(dd.neighbor_tile_offsets(), dd.neighbor_tiles(), 2).empty());
    CHECK(validate_tile_directory(dp.graph, dd).ok());
  }

  // tile_test.cpp:230-262: determinism; a graph without spatial order.
  const PartitionedGraph again = partitioner.partition(spatial_fixture());
  const TileDirectory& d2 = again.tiles;
  auto same = [](auto a, auto b) { return std::equal(a.begin(), a.end(), b.begin(), b.end()); };
  CHECK(same(d.neighbor_tile_offsets(), d2.neighbor_tile_offsets()) &&
        same(d.neighbor_tiles(), d2.neighbor_tiles()) &&
        same(d.internal_edge_offsets(), d2.internal_edge_offsets()) &&
        same(d.internal_edge_ids(), d2.internal_edge_ids()) &&
        same(d.outgoing_cross_edge_offsets(), d2.outgoing_cross_edge_offsets()) &&
        same(d.outgoing_cross_edge_ids(), d2.outgoing_cross_edge_ids()) &&
        same(d.incoming_cross_edge_offsets(), d2.incoming_cross_edge_offsets()) &&
        same(d.incoming_cross_edge_ids(), d2.incoming_cross_edge_ids()) &&
        same(d.halo_vertex_offsets(), d2.halo_vertex_offsets()) &&
        same(d.halo_vertices(), d2.halo_vertices()));
  const WeightedGraph plain = build_weighted_graph(spatial_fixture());
  CHECK(throws<std::invalid_argument>([&] { (void)build_tile_directory(plain); }));
  CHECK(validate_tile_directory(plain, d).code == Code::graph_is_not_spatially_ordered);

  // Planted mistakes: each swap breaks one ordering rule.
  CHECK(validate_tile_directory(g, tile_directory_with_swapped_entries(d, 0, 0, 1)).code ==
        Code::neighbor_order_violation);
  CHECK(validate_tile_directory(g, tile_directory_with_swapped_entries(d, 4, 0, 1)).code ==
        Code::halo_order_violation);
  // Two entries of one tile's internal group swapped break the CSR order.
  {
    const auto off = d.internal_edge_offsets();
    std::size_t t = 0;
    while (t < d.tile_count() && off[t + 1] - off[t] < 2) ++t;
    CHECK(t < d.tile_count());
    if (t < d.tile_count())
      CHECK(validate_tile_directory(g, tile_directory_with_swapped_entries(d, 1, off[t], off[t] + 1)).code ==
            Code::edge_metadata_order_violation);
  }
  {
    // An outgoing edge moved into another tile's group is misclassified.
    const auto off = d.outgoing_cross_edge_offsets();
    std::vector<std::size_t> nonempty;
    for (std::size_t t = 0; t < d.tile_count(); ++t)
      if (off[t + 1] > off[t]) nonempty.push_back(t);
    CHECK(nonempty.size() >= 2U);
    if (nonempty.size() >= 2U)
      CHECK(validate_tile_directory(
                g, tile_directory_with_swapped_entries(d, 2, off[nonempty[0]], off[nonempty[1]]))
                .code == Code::edge_classification_mismatch);
  }
}

// ---------------------------------------------------------------------------
// Path reconstruction.
// ---------------------------------------------------------------------------
std::vector<VertexMetadata> plain_vertices(std::uint32_t n) {
  std::vector<VertexMetadata> v;
  for (std::uint32_t i = 0; i < n; ++i) v.push_back(VertexMetadata::located(static_cast<int>(i), 0, kClass1));
  return v;
}

void paths() {
  // sssp_test.cpp:108-133: core fixture; the path to 5 takes the zero edge.
  {
    std::vector<VertexMetadata> v{VertexMetadata::located(0, 0, kClass1), VertexMetadata::located(1, 0, kClass2),
                                  VertexMetadata::located(2, 0, kClass1), VertexMetadata::unlocated(kClass2),
                                  VertexMetadata::located(3, 1, kClass2), VertexMetadata::located(4, 1, kClass1),
                                  VertexMetadata::located(9, 9, kClass2)};
    std::vector<EdgeInputRecord> e{edge(0, 1, 0.5F, 0),  edge(0, 2, 2.25F, 1), edge(1, 2, 0.0F, 2),
                                   edge(1, 2, 0.75F, 3), edge(2, 2, 0.125F, 4), edge(2, 3, 1.5F, 5),
                                   edge(3, 4, 2.0F, 6),  edge(0, 4, 7.75F, 7), edge(4, 5, 0.25F, 8)};
    const WeightedGraph g = build_weighted_graph(InputGraph{std::move(v), std::move(e)});
    const std::vector<VertexId> src{g.old_to_new()[0]};
    SsspResult r{region_labels(g, src, nullptr), src};
    const float expect[] = {0.0F, 0.5F, 0.5F, 2.0F, 4.0F, 4.25F, kInf};
    for (std::uint32_t old = 0; old < 7; ++old) CHECK(r.distances[g.old_to_new()[old].value()] == expect[old]);
    const auto p = reconstruct_path_from_distances(g, r, g.old_to_new()[5]);
    CHECK(p.has_value());
    if (p) {
      CHECK(old_ids(g, p->vertices) == (std::vector<std::uint32_t>{0, 1, 2, 3, 4, 5}));
      CHECK(p->cost == 4.25F);
      CHECK(validate_reconstructed_path(g, r, g.old_to_new()[5], *p));
      ReconstructedPath broken = *p;
      broken.cost = 4.0F;
      CHECK(!validate_reconstructed_path(g, r, g.old_to_new()[5], broken));
      broken = *p;
      std::swap(broken.edge_ids[0], broken.edge_ids[1]);
      CHECK(!validate_reconstructed_path(g, r, g.old_to_new()[5], broken));
    }
    CHECK(!reconstruct_path_from_distances(g, r, g.old_to_new()[6]).has_value());
    CHECK(throws<std::out_of_range>([&] { (void)reconstruct_path_from_distances(g, r, VertexId{7U}); }));
    SsspResult bad = r;
    bad.distances[g.old_to_new()[0].value()] = 1.0F;
    CHECK(throws<std::invalid_argument>([&] { (void)reconstruct_path_from_distances(g, bad, VertexId{1U}); }));
    bad = r;
    bad.distances.pop_back();
    CHECK(throws<std::invalid_argument>([&] { (void)reconstruct_path_from_distances(g, bad, VertexId{1U}); }));
    // A finite label with no tight in-edge has no path.
    bad = r;
    bad.distances[g.old_to_new()[5].value()] = 4.5F;
    CHECK(throws<std::logic_error>([&] { (void)reconstruct_path_from_distances(g, bad, g.old_to_new()[5]); }));
  }
  // sssp_test.cpp:174-200: an equal-cost tie takes the smaller stable EdgeId.
  {
    const WeightedGraph g = build_weighted_graph(InputGraph{
        plain_vertices(4),
        {edge(0, 1, 0.5F, 200), edge(0, 2, 0.5F, 201), edge(1, 3, 0.5F, 202), edge(2, 3, 0.5F, 203)}});
    const std::vector<VertexId> src{g.old_to_new()[0]};
    SsspResult r{region_labels(g, src, nullptr), src};
    const auto p = reconstruct_path_from_distances(g, r, g.old_to_new()[3]);
    CHECK(p && old_ids(g, p->vertices) == (std::vector<std::uint32_t>{0, 1, 3}) && p->cost == 1.0F);
  }
  // sssp_test.cpp:202-227: a zero-weight cycle forces backtracking.
  {
    const WeightedGraph g = build_weighted_graph(InputGraph{
        plain_vertices(4),
        {edge(0, 1, 0.0F, 300), edge(1, 0, 0.0F, 301), edge(1, 2, 1.0F, 302), edge(3, 1, 1.0F, 303)}});
    const std::vector<VertexId> src{g.old_to_new()[3]};
    SsspResult r{region_labels(g, src, nullptr), src};
    const auto p = reconstruct_path_from_distances(g, r, g.old_to_new()[2]);
    CHECK(p && old_ids(g, p->vertices) == (std::vector<std::uint32_t>{3, 1, 2}) && p->cost == 2.0F);
    CHECK(p && validate_reconstructed_path(g, r, g.old_to_new()[2], *p));
  }
  CHECK(nonnegative_distance_within_ulps(1.0F, std::nextafter(1.0F, 2.0F), 1U));
  CHECK(!nonnegative_distance_within_ulps(1.0F, 1.5F));
  CHECK(nonnegative_distance_within_ulps(kInf, kInf) && !nonnegative_distance_within_ulps(kInf, 1.0F));
  CHECK(!nonnegative_distance_within_ulps(-1.0F, -1.0F));

  // compact_paths_fixture_suite.hpp:45-145 / compact_paths_test.cpp:114-205.
  const int xs[16][2] = {{10, 0}, {20, 0}, {30, 0}, {0, 0},   {40, 0},  {50, 0},  {50, 0},  {60, 0},
                         {80, 0}, {100, 0}, {90, 0}, {90, 20}, {120, 0}, {121, 0}, {140, 0}, {160, 0}};
  std::vector<VertexMetadata> cv;
  for (const auto& x : xs) cv.push_back(VertexMetadata::located(x[0], x[1], kClass1));
  std::vector<EdgeInputRecord> ce{
      edge(0, 1, 0.0F, 18000),  edge(1, 0, 0.0F, 18001),  edge(1, 2, 1.0F, 18002),
      edge(3, 1, 1.0F, 18003),  edge(4, 5, 0.5F, 18004),  edge(4, 6, 0.5F, 18005),
      edge(5, 7, 0.5F, 18006),  edge(6, 7, 0.5F, 18007),  edge(5, 7, 0.5F, 18014),
      edge(8, 10, 2.0F, 18008), edge(10, 9, 2.0F, 18009), edge(8, 11, 0.5F, 18010),
      edge(11, 9, 0.5F, 18011), edge(12, 14, 3.0F, 18012), edge(13, 14, 1.0F, 18013)};
  const WeightedGraph g = build_spatially_ordered_graph(InputGraph{std::move(cv), std::move(ce)},
                                                        SpatialOrderConfig{0, 0, 10U, 10U});
  auto ids = [&](std::initializer_list<std::uint32_t> old) {
    std::vector<VertexId> out;
    for (auto o : old) out.push_back(g.old_to_new()[o]);
    return out;
  };
  struct Case {
    std::uint32_t id;
    std::vector<std::uint32_t> sources, targets;
  };
  const Case cases[] = {{1800, {3}, {2}}, {1801, {4}, {7}}, {1802, {8}, {9}}, {1803, {12, 13, 12}, {14, 14, 12}},
                        {1804, {12}, {15}}};
  std::vector<ExpansionQueryOutcome> outcomes;
  for (const Case& c : cases) {
    std::vector<VertexId> s, t;
    for (auto o : c.sources) s.push_back(g.old_to_new()[o]);
    for (auto o : c.targets) t.push_back(g.old_to_new()[o]);
    const RouteQuery q = make_route_query(QueryId{c.id}, g, s, t, 0U, 0U);
    const std::vector<float> d = region_labels(g, q.sources, &q.selected_tiles);
    const CompactPathPayload p = reconstruct_compact_path_payload(g, q, d);
    CHECK(p.query_id == q.query_id && p.targets.size() == q.targets.size());
    CHECK(validate_compact_path_payload(g, q, p).ok());
    auto path_of = [&](std::uint32_t old_target) -> const CompactTargetPath* {
      for (const auto& tp : p.targets)
        if (tp.summary.target == g.old_to_new()[old_target]) return &tp;
      return nullptr;
    };
    auto labels = [](const CompactTargetPath* tp) { return tp->distance_labels; };
    switch (c.id) {
      case 1800: {  // zero cycle: old {3,1,2}, labels {0,1,2}
        const auto* tp = path_of(2);
        CHECK(tp && tp->vertices == ids({3, 1, 2}) && labels(tp) == (std::vector<float>{0.0F, 1.0F, 2.0F}));
        break;
      }
      case 1801: {  // tie: old {4,5,7} at 1.0, the smaller of two parallel edge ids
        const auto* tp = path_of(7);
        CHECK(tp && tp->vertices == ids({4, 5, 7}) && tp->summary.distance == 1.0F);
        CHECK(tp && tp->edge_ids.size() == 2 && tp->edge_ids[1] == edge_of_record(g, 18006));
        break;
      }
      case 1802: {  // bounded 4.0 through old 10 (the unbounded path costs 1.0)
        const auto* tp = path_of(9);
        CHECK(tp && tp->vertices == ids({8, 10, 9}) && tp->summary.distance == 4.0F);
        RouteQuery full = q;
        full.selected_tiles.clear();
        for (std::uint32_t t = 0; t < g.tile_coordinates().size(); ++t) full.selected_tiles.push_back(TileId{t});
        const auto pf = reconstruct_compact_path_payload(g, full, region_labels(g, full.sources, nullptr));
        CHECK(pf.targets[0].summary.distance == 1.0F && pf.targets[0].vertices == ids({8, 11, 9}));
        break;
      }
      case 1803: {  // a target that is a source has a zero-edge path
        const auto* t12 = path_of(12);
        const auto* t14 = path_of(14);
        CHECK(t12 && t12->vertices == ids({12}) && t12->summary.path_length == 0U);
        CHECK(t14 && t14->vertices == ids({13, 14}) && labels(t14) == (std::vector<float>{0.0F, 1.0F}));
        CHECK(p.target_terminal_to_target == q.target_terminal_to_target);
        break;
      }
      case 1804: {
        const auto* tp = path_of(15);
        CHECK(tp && tp->summary.reconstruction == CompactPathStatus::unreachable && tp->vertices.empty());
        break;
      }
    }
    const bool all = compact_path_payload_complete(p);
    CHECK(all == (c.id != 1804));
    const auto disposition =
        all ? ExpansionQueryDisposition::reached : ExpansionQueryDisposition::unreachable_in_full_region;
    const CompactQueryResult r = reconstruct_compact_query_paths(g, q, d, disposition);
    CHECK(r.targets == p.targets && r.disposition == disposition);
    CHECK(throws<std::logic_error>([&] {
      (void)reconstruct_compact_query_paths(g, q, d, all ? ExpansionQueryDisposition::unreachable_in_full_region
                                                        : ExpansionQueryDisposition::reached);
    }));
    std::vector<float> bad = d;
    bad[q.sources[0].value()] = 0.5F;
    CHECK(throws<std::invalid_argument>([&] { (void)reconstruct_compact_path_payload(g, q, bad); }));
    bad = d;
    bad[0] = -1.0F;
    CHECK(throws<std::invalid_argument>([&] { (void)reconstruct_compact_path_payload(g, q, bad); }));
    // A finite label without a tight predecessor: no_tight_path.
    if (c.id == 1801) {
      bad = d;
      bad[g.old_to_new()[7].value()] = 0.75F;
      const auto np = reconstruct_compact_path_payload(g, q, bad);
      CHECK(np.targets[0].summary.reconstruction == CompactPathStatus::no_tight_path &&
            np.targets[0].summary.reached == CompactTargetReachStatus::reached);
    }
    // Outcomes for extraction: labels only, a retained payload, a failure.
    ExpansionQueryOutcome o;
    o.final_query = q;
    o.disposition = disposition;
    if (c.id == 1800) {
      o.compact_paths = p;
    } else if (c.id == 1803) {
      o.disposition = ExpansionQueryDisposition::expansion_limit;
    } else {
      o.final_distances = d;
    }
    outcomes.push_back(std::move(o));
  }
  const auto failed = make_failed_compact_query_result(outcomes[3].final_query, ExpansionQueryDisposition::region_stalled);
  CHECK(std::all_of(failed.targets.begin(), failed.targets.end(), [](const CompactTargetPath& t) {
    return t.summary.reconstruction == CompactPathStatus::query_terminal_failure && !std::isfinite(t.summary.distance);
  }));
  CHECK(throws<std::invalid_argument>(
      [&] { (void)make_failed_compact_query_result(outcomes[0].final_query, ExpansionQueryDisposition::reached); }));
  const std::vector<ExpansionQueryOutcome> kept = outcomes;
  const auto results = extract_host_compact_paths(g, outcomes);
  CHECK(results.size() == 5U);
  CHECK(results[0].targets[0].vertices == ids({3, 1, 2}) && !outcomes[0].compact_paths.has_value());
  CHECK(results[1].targets[0].vertices == ids({4, 5, 7}) && outcomes[1].final_distances.empty());
  CHECK(results[3].disposition == ExpansionQueryDisposition::expansion_limit &&
        results[3].targets[0].summary.reconstruction == CompactPathStatus::query_terminal_failure);
  CHECK(results[4].disposition == ExpansionQueryDisposition::unreachable_in_full_region);
  for (std::size_t i = 0; i < results.size(); ++i)
    CHECK(validate_compact_query_result(g, kept[i].final_query, results[i]).ok());
  std::vector<ExpansionQueryOutcome> unordered{kept[1], kept[0]};
  CHECK(throws<std::invalid_argument>([&] { (void)extract_host_compact_paths(g, unordered); }));
  std::vector<ExpansionQueryOutcome> empty_reached{kept[1]};
  empty_reached[0].final_distances.clear();
  CHECK(throws<std::logic_error>([&] { (void)extract_host_compact_paths(g, empty_reached); }));
}

// ---------------------------------------------------------------------------
// Workspace byte model (batch_workspace_test.cpp:77-960).
// ---------------------------------------------------------------------------
struct WorkspaceFixture {
  PartitionedGraph partitioned;
  TileRunLayout64 tile_runs;
  std::vector<RouteQuery> dense, sparse;
};

WorkspaceFixture workspace_fixture() {
  std::vector<VertexMetadata> v{
      VertexMetadata::located(0, 0, kClass1),  VertexMetadata::located(1, 0, kClass1),
      VertexMetadata::located(10, 0, kClass1), VertexMetadata::located(30, 0, kClass1),
      VertexMetadata::unlocated(kClass1),      VertexMetadata::located(11, 0, kClass1),
      VertexMetadata::located(2, 0, kClass1)};
  std::vector<EdgeInputRecord> e{edge(0, 1, 1.0F, 0), edge(0, 2, 2.0F, 1), edge(0, 3, 3.0F, 2),
                                 edge(0, 3, 4.0F, 3), edge(0, 4, 5.0F, 4), edge(4, 0, 6.0F, 5),
                                 edge(2, 0, 7.0F, 6), edge(1, 2, 8.0F, 7)};
  const UniformGridPartitioner partitioner{SpatialOrderConfig{0, 0, 10U, 10U}};
  WorkspaceFixture f{partitioner.partition(InputGraph{std::move(v), std::move(e)}), {}, {}, {}};
  const WeightedGraph& g = f.partitioned.graph;
  auto vid = [&](std::uint32_t old) { return g.old_to_new()[old]; };
  auto tile = [&](std::uint32_t old) { return g.owner_tiles()[vid(old).value()]; };
  auto query = [&](std::uint32_t id, std::vector<std::uint32_t> s, std::vector<std::uint32_t> t,
                   std::uint32_t generation, std::vector<TileId> tiles) {
    std::vector<VertexId> sv, tv;
    for (auto x : s) sv.push_back(vid(x));
    for (auto x : t) tv.push_back(vid(x));
    RouteQuery q = make_route_query(QueryId{id}, g, sv, tv, 0U, generation);
    std::sort(tiles.begin(), tiles.end());
    q.selected_tiles = std::move(tiles);
    if (!validate_route_query(g, q).ok()) throw std::runtime_error{"invalid workspace fixture query"};
    return q;
  };
  const TileId t0 = tile(0), t1 = tile(2), t2 = tile(3), spill = tile(4);
  f.dense = {query(10, {0, 2}, {5}, 0, {t0, t1}), query(20, {0}, {4}, 1, {t0, spill}),
             query(30, {4}, {4}, 2, {spill}), query(40, {2}, {3}, 3, {t1, t2})};
  f.sparse = {query(50, {2}, {5}, 0, {t1}), query(60, {4}, {4}, 0, {spill})};
  f.tile_runs = build_tile_run_layout(g);
  return f;
}

struct PreparedBatch {
  std::vector<BatchQueryFeatures> features;
  BatchPlan plan;
  BatchDeviceDescription retained, descriptors;
};

// Cold preparation, then the same preparation again (the warmed reusable
// path the reported write counters model).
PreparedBatch prepare_warm(const WorkspaceFixture& f, const SelectedRegionIndex& regions,
                           std::span<const RouteQuery> queries, std::uint32_t width, bool zero_overlap) {
  const WeightedGraph& g = f.partitioned.graph;
  PreparedBatch p;
  p.features = make_batch_query_features(g, regions, queries);
  BatchPlannerPolicy policy;
  policy.lane_width = width;
  policy.maximum_union_inflation_numerator = 3U;
  if (zero_overlap) {
    policy.minimum_jaccard_numerator = 0U;
    policy.minimum_jaccard_denominator = 1U;
  }
  p.plan = make_overlapping_batch_plan(regions, p.features, policy);
  if (p.plan.batches.size() != 1U) throw std::runtime_error{"fixture split into several batches"};
  const BatchPlanEntry& b = p.plan.batches.front();
  for (int pass = 0; pass < 2; ++pass) {
    prepare_batch_device_description(g, f.tile_runs, queries, p.features, b,
                                     BatchRunRepresentation::retained_per_run_masks, p.retained);
    prepare_batch_device_description(g, f.tile_runs, queries, p.features, b,
                                     BatchRunRepresentation::compact_nonzero_descriptors, p.descriptors);
    prepare_compact_vertex_mapping(g, b, p.retained);
    prepare_compact_vertex_mapping(g, b, p.descriptors);
  }
  if (!validate_compact_vertex_mapping(g, b, p.retained).ok() ||
      !validate_compact_vertex_mapping(g, b, p.descriptors).ok())
    throw std::runtime_error{"fixture compact mapping is invalid"};
  return p;
}

const BatchWorkspaceEstimate& row(const std::vector<BatchWorkspaceEstimate>& rows, BatchVertexStorageStrategy s,
                                  BatchRunRepresentation r) {
  for (const auto& e : rows)
    if (e.vertex_storage == s && e.run_representation == r) return e;
  throw std::runtime_error{"missing strategy row"};
}

void workspace() {
  using S = BatchVertexStorageStrategy;
  using R = BatchRunRepresentation;
  const WorkspaceFixture f = workspace_fixture();
  const WeightedGraph& g = f.partitioned.graph;
  const SelectedRegionIndex regions{g, f.tile_runs};

  // Dense matrix at widths 8/16/32 and one or two slots
  // (batch_workspace_test.cpp:275-378).
  const BatchWorkspaceBudget budget{1'000'000U, 1'000U, 500U};
  for (const std::uint32_t width : {8U, 16U, 32U}) {
    const PreparedBatch p = prepare_warm(f, regions, f.dense, width, false);
    const BatchPlanEntry& b = p.plan.batches.front();
    CHECK(b.union_vertex_count == 7U && b.selected_lane_vertex_count == 13U);
    const std::uint64_t metadata = 32U * static_cast<std::uint64_t>(width) + 280U;
    for (const std::uint32_t slots : {1U, 2U}) {
      const auto rows = compare_batch_workspace_strategies(g, f.tile_runs, b, p.retained, p.descriptors, slots, budget);
      CHECK(rows.size() == 4U);
      const std::uint64_t dist = 7U * static_cast<std::uint64_t>(width) * slots * 4U;
      const std::uint64_t reset = 13U * slots * 4U;
      for (const S storage : {S::full_graph_vertex_major, S::compact_union_tiles}) {
        const bool compact = storage == S::compact_union_tiles;
        const auto& ret = row(rows, storage, R::retained_per_run_masks);
        const auto& des = row(rows, storage, R::compact_nonzero_descriptors);
        CHECK(ret.distance_bytes == dist && des.distance_bytes == dist);
        CHECK(ret.distance_reset_bytes == reset && des.distance_reset_bytes == reset);
        CHECK(ret.allocated_lane_vertex_count == 7U * width && ret.wasted_lane_vertex_count == 7U * width - 13U);
        CHECK(ret.batch_metadata_bytes == metadata && des.batch_metadata_bytes == metadata);
        CHECK(ret.run_storage_bytes == 28U && ret.descriptor_offset_bytes == 0U &&
              ret.tile_mapping_bytes == (compact ? 16U : 0U));
        CHECK(des.run_storage_bytes == 48U && des.descriptor_offset_bytes == 32U &&
              des.tile_mapping_bytes == (compact ? 16U : 0U));
        CHECK(ret.active_csr_runs == 6U && ret.active_csc_runs == 5U && ret.zero_csr_runs == 1U &&
              ret.zero_csc_runs == 1U && des.active_csr_runs == 6U && des.active_csc_runs == 5U);
        CHECK(ret.run_preparation_write_bytes == 48U && des.run_preparation_write_bytes == 80U);
        CHECK(ret.tile_mapping_write_bytes == (compact ? 32U : 0U) &&
              des.tile_mapping_write_bytes == (compact ? 32U : 0U));
        CHECK(ret.total_workspace_bytes == dist + (compact ? 16U : 0U) + 28U + metadata);
        CHECK(des.total_workspace_bytes == dist + (compact ? 16U : 0U) + 48U + 32U + metadata);
        CHECK(ret.maximum_concurrent_workspaces == (1'000'000U - 1'500U) / ret.total_workspace_bytes);
      }
    }
  }

  // Sparse, noncontiguous compact union (batch_workspace_test.cpp:380-488).
  {
    const PreparedBatch p = prepare_warm(f, regions, f.sparse, 8U, true);
    const BatchPlanEntry& b = p.plan.batches.front();
    CHECK(b.union_tiles == (std::vector<TileId>{TileId{1U}, TileId{3U}}) && b.union_vertex_count == 3U &&
          b.selected_lane_vertex_count == 3U);
    CHECK(p.retained.touched_compact_tiles == (std::vector<std::uint32_t>{1U, 3U}));
    CHECK(p.retained.compact_vertex_biases_by_tile == (std::vector<std::uint32_t>{0U, 3U, 0U, 4U}));
    BatchDeviceDescription stale = p.retained;
    stale.compact_vertex_biases_by_tile[0] = 9U;
    CHECK(validate_compact_vertex_mapping(g, b, stale).code ==
          BatchLayoutValidationErrorCode::invalid_compact_vertex_mapping);
    CHECK(p.retained.run_report.active_csr_runs == 0U && p.retained.run_report.active_csc_runs == 0U &&
          p.descriptors.csr_run_descriptors.empty() && p.descriptors.csc_run_descriptors.empty());
    const BatchWorkspaceBudget small{100'000U, 0U, 0U};
    const auto rows = compare_batch_workspace_strategies(g, f.tile_runs, b, p.retained, p.descriptors, 2U, small);
    BatchDeviceDescription cold;
    prepare_batch_device_description(g, f.tile_runs, f.sparse, p.features, b, R::retained_per_run_masks, cold);
    prepare_compact_vertex_mapping(g, b, cold);
    const auto c = estimate_batch_workspace(g, f.tile_runs, b, cold, S::compact_union_tiles,
                                            R::retained_per_run_masks, 2U, small);
    CHECK(c.tile_mapping_write_bytes == 24U && c.run_preparation_write_bytes == 28U &&
          c.total_preparation_write_bytes == 76U);
    const auto& fr = row(rows, S::full_graph_vertex_major, R::retained_per_run_masks);
    const auto& cr = row(rows, S::compact_union_tiles, R::retained_per_run_masks);
    const auto& fd = row(rows, S::full_graph_vertex_major, R::compact_nonzero_descriptors);
    const auto& cd = row(rows, S::compact_union_tiles, R::compact_nonzero_descriptors);
    CHECK(fr.distance_bytes == 448U && cr.distance_bytes == 192U && fr.wasted_lane_vertex_count == 53U &&
          cr.wasted_lane_vertex_count == 21U);
    CHECK(fr.distance_reset_bytes == 24U && cr.distance_reset_bytes == 24U);
    CHECK(fr.total_workspace_bytes == 960U && cr.total_workspace_bytes == 720U &&
          fd.total_workspace_bytes == 948U && cd.total_workspace_bytes == 708U);
  }

  // Budget boundaries, rejections, overflow and reuse (:490-670).
  const PreparedBatch p = prepare_warm(f, regions, f.dense, 8U, false);
  const BatchPlanEntry& b = p.plan.batches.front();
  const BatchWorkspaceBudget generous{1'000'000U, 0U, 0U};
  auto estimate = [&](const BatchPlanEntry& batch, const BatchDeviceDescription& d, S s, R r, std::uint32_t slots,
                      const BatchWorkspaceBudget& budget_) {
    return estimate_batch_workspace(g, f.tile_runs, batch, d, s, r, slots, budget_);
  };
  const auto base = estimate(b, p.retained, S::full_graph_vertex_major, R::retained_per_run_masks, 2U, generous);
  const std::uint64_t bytes = base.total_workspace_bytes;
  for (const auto& [capacity, expected] : {std::pair{bytes - 1U, 0U}, std::pair{bytes, 1U},
                                           std::pair{3U * bytes - 1U, 2U}, std::pair{3U * bytes, 3U}})
    CHECK(estimate(b, p.retained, S::full_graph_vertex_major, R::retained_per_run_masks, 2U,
                   BatchWorkspaceBudget{capacity, 0U, 0U})
              .maximum_concurrent_workspaces == expected);
  CHECK(throws<std::invalid_argument>(
      [&] { (void)estimate(b, p.retained, S::full_graph_vertex_major, R::retained_per_run_masks, 0U, generous); }));
  CHECK(throws<std::invalid_argument>(
      [&] { (void)estimate(b, p.retained, static_cast<S>(255U), R::retained_per_run_masks, 2U, generous); }));
  CHECK(throws<std::invalid_argument>(
      [&] { (void)estimate(b, p.retained, S::full_graph_vertex_major, static_cast<R>(255U), 2U, generous); }));
  BatchDeviceDescription unmapped = p.retained;
  unmapped.compact_vertex_mapping_valid = false;
  CHECK(throws<std::invalid_argument>(
      [&] { (void)estimate(b, unmapped, S::compact_union_tiles, R::retained_per_run_masks, 2U, generous); }));
  CHECK(throws<std::overflow_error>([&] {
    (void)estimate(b, p.retained, S::full_graph_vertex_major, R::retained_per_run_masks, 2U,
                   BatchWorkspaceBudget{~0ULL, ~0ULL, 1U});
  }));
  BatchPlanEntry overflow = b;
  overflow.union_vertex_count = ~0ULL;
  overflow.selected_lane_vertex_count = 0U;
  CHECK(throws<std::overflow_error>(
      [&] { (void)estimate(overflow, p.retained, S::compact_union_tiles, R::retained_per_run_masks, 2U, generous); }));
  BatchPlanEntry impossible = b;
  impossible.selected_lane_vertex_count = static_cast<std::uint64_t>(g.vertex_count()) * b.lane_width + 1U;
  CHECK(throws<std::invalid_argument>([&] {
    (void)estimate(impossible, p.retained, S::full_graph_vertex_major, R::retained_per_run_masks, 2U, generous);
  }));
  ReusableBatchWorkspaceReservation reservation;
  const auto first = reservation.reserve(base);
  const auto repeated = reservation.reserve(base);
  CHECK(first.capacity_grew && !repeated.capacity_grew && first.generation == 1U && repeated.generation == 2U &&
        reservation.growth_events() == 1U);
  const PreparedBatch wide = prepare_warm(f, regions, f.dense, 16U, false);
  const auto larger = estimate(wide.plan.batches.front(), wide.retained, S::full_graph_vertex_major,
                               R::retained_per_run_masks, 2U, generous);
  const auto grown = reservation.reserve(larger);
  CHECK(grown.capacity_grew && grown.generation == 3U && reservation.growth_events() == 2U &&
        reservation.capacity().total_bytes() >= larger.total_workspace_bytes);
  BatchWorkspaceEstimate inconsistent = base;
  ++inconsistent.total_workspace_bytes;
  CHECK(throws<std::invalid_argument>([&] { (void)reservation.reserve(inconsistent); }));

  // Device-materialised Jacobi rows (:672-748).
  {
    BatchDeviceDescription dm;
    prepare_batch_device_description(g, f.tile_runs, f.dense, p.features, b, R::device_materialized_run_masks, dm);
    const auto e = estimate(b, dm, S::full_graph_vertex_major, R::device_materialized_run_masks, 2U, generous);
    CHECK(e.run_storage_bytes == f.tile_runs.csc_run_source_tiles.size() * 4U && e.descriptor_offset_bytes == 0U &&
          e.active_csr_runs == 0U && e.zero_csr_runs == 0U &&
          e.active_csc_runs == p.retained.run_report.active_csc_runs &&
          e.zero_csc_runs + e.active_csc_runs == f.tile_runs.csc_run_source_tiles.size());
    CHECK(e.run_preparation_write_bytes == p.retained.run_report.csc_runs_visited * 4U && e.run_build_nanoseconds == 0U);
    CHECK(throws<std::invalid_argument>([&] {
      (void)estimate_batch_workspace(g, f.tile_runs, b, dm, S::full_graph_vertex_major,
                                     R::device_materialized_run_masks, 2U, generous, {0U, 1U});
    }));
    CHECK(throws<std::invalid_argument>(
        [&] { (void)estimate(b, dm, S::compact_union_tiles, R::device_materialized_run_masks, 2U, generous); }));
    CHECK(throws<std::invalid_argument>(
        [&] { (void)estimate(b, dm, S::full_graph_vertex_major, R::device_materialized_run_masks, 1U, generous); }));
  }

  // Warmed preparation counters and the recorded decision (:750-960).
  CHECK(p.retained.compact_mapping_report.entries_cleared == 4U &&
        p.retained.compact_mapping_report.entries_written == 4U &&
        p.retained.run_report.csr_retained_entries_cleared == 6U &&
        p.retained.run_report.csr_retained_entries_written == 6U);
  CHECK(p.retained.run_report.csr_lane_edge_pairs == 7U && p.retained.run_report.csc_lane_edge_pairs == 7U &&
        p.descriptors.run_report.csr_lane_edge_pairs == 7U && p.descriptors.run_report.csc_lane_edge_pairs == 7U);
  const BatchWorkspaceBudget decision_budget{65'536U, 4'096U, 4'096U};
  const BatchWorkspaceBuildMeasurements measured{11U, 22U, 33U};
  const auto rows = compare_batch_workspace_strategies(g, f.tile_runs, b, p.retained, p.descriptors, 2U,
                                                       decision_budget, measured);
  for (const auto& e : rows) {
    const bool compact = e.vertex_storage == S::compact_union_tiles;
    const bool retained = e.run_representation == R::retained_per_run_masks;
    CHECK(e.mapping_build_nanoseconds == (compact ? 11U : 0U) && e.run_build_nanoseconds == (retained ? 22U : 33U));
  }
  CHECK(row(rows, S::full_graph_vertex_major, R::retained_per_run_masks).total_preparation_write_bytes == 152U);
  CHECK(row(rows, S::full_graph_vertex_major, R::compact_nonzero_descriptors).total_preparation_write_bytes == 184U);
  CHECK(row(rows, S::compact_union_tiles, R::retained_per_run_masks).total_preparation_write_bytes == 184U);
  CHECK(row(rows, S::compact_union_tiles, R::compact_nonzero_descriptors).total_preparation_write_bytes == 216U);
  BatchWorkspaceDecision decision;
  decision.plan_fingerprint = fingerprint_batch_plan(p.plan);
  decision.budget = decision_budget;
  decision.graph_vertex_count = g.vertex_count();
  decision.graph_tile_count = g.tile_coordinates().size();
  decision.lane_width = 8U;
  decision.distance_slot_count = 2U;
  decision.compared_strategies = rows;
  decision.selected_run_build_nanoseconds = 22U;
  decision.quantitative_reason = "full/retained uses the fewest bytes and needs no mapping table";
  CHECK(validate_batch_workspace_decision(decision));
  const std::string text = serialize_batch_workspace_decision(decision);
  CHECK(text.rfind("bfnew_phase13_workspace_decision_v1\nplan_fingerprint=", 0) == 0);
  CHECK(text.find("selected_vertex_storage=full_graph_vertex_major\n") != std::string::npos);
  CHECK(std::count(text.begin(), text.end(), '\n') == 17 + 4);  // 16 header lines, column names, 4 rows
  // Each planted inconsistency is refused.
  auto refused = [&](auto&& plant) {
    BatchWorkspaceDecision d = decision;
    plant(d);
    return !validate_batch_workspace_decision(d) &&
           throws<std::invalid_argument>([&] { (void)serialize_batch_workspace_decision(d); });
  };
  CHECK(refused([](BatchWorkspaceDecision& d) { d.quantitative_reason = "tab\there"; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { d.compared_strategies.pop_back(); }));
  CHECK(refused([](BatchWorkspaceDecision& d) { d.compared_strategies[1] = d.compared_strategies[0]; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { ++d.compared_strategies[0].distance_bytes; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { ++d.compared_strategies[2].run_storage_bytes; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { d.selected_run_build_nanoseconds = 23U; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { d.budget.device_capacity_bytes = 8'192U; }));
  CHECK(refused([](BatchWorkspaceDecision& d) { d.plan_fingerprint = 0U; }));
}

// ---------------------------------------------------------------------------
// Paths for a graph written by the Python test (oracle labels).
//   IN : u32 V, u32 E, then V x (i32 x, i32 y, u32 located, u32 class),
//        E x (u32 src, u32 dst, f32 w, u64 record), u32 nsrc, nsrc x u32
//        source (old id), V x f32 labels (in the built graph's order),
//        u32 ntgt, ntgt x u32 target (built-graph id)
//   OUT: per target: u32 status (0 complete, 1 unreachable), u32 n,
//        n x u32 vertex, (n-1) x u64 edge id, f32 cost
// ---------------------------------------------------------------------------
template <typename T>
T get(std::ifstream& f) {
  T x{};
  f.read(reinterpret_cast<char*>(&x), sizeof x);
  if (!f) throw std::runtime_error{"short input"};
  return x;
}

template <typename T>
void put(std::ofstream& f, T x) {
  f.write(reinterpret_cast<const char*>(&x), sizeof x);
}

int paths_file(const char* in_path, const char* out_path) {
  std::ifstream f(in_path, std::ios::binary);
  const auto V = get<std::uint32_t>(f), E = get<std::uint32_t>(f);
  std::vector<VertexMetadata> v;
  for (std::uint32_t i = 0; i < V; ++i) {
    const auto x = get<std::int32_t>(f), y = get<std::int32_t>(f);
    const auto located = get<std::uint32_t>(f), cls = get<std::uint32_t>(f);
    const ResourceClassId rc{static_cast<std::uint16_t>(cls)};
    v.push_back(located ? VertexMetadata::located(x, y, rc) : VertexMetadata::unlocated(rc));
  }
  std::vector<EdgeInputRecord> e;
  for (std::uint32_t i = 0; i < E; ++i) {
    const auto s = get<std::uint32_t>(f), d = get<std::uint32_t>(f);
    const auto w = get<float>(f);
    const auto rec = get<std::uint64_t>(f);
    e.push_back(edge(s, d, w, rec));
  }
  const WeightedGraph g = build_weighted_graph(InputGraph{std::move(v), std::move(e)});
  SsspResult r;
  const auto ns = get<std::uint32_t>(f);
  for (std::uint32_t i = 0; i < ns; ++i) r.sources.push_back(g.old_to_new()[get<std::uint32_t>(f)]);
  std::sort(r.sources.begin(), r.sources.end());
  r.distances.resize(V);
  for (auto& x : r.distances) x = get<float>(f);
  const auto nt = get<std::uint32_t>(f);
  std::ofstream out(out_path, std::ios::binary);
  for (std::uint32_t i = 0; i < nt; ++i) {
    const VertexId t{get<std::uint32_t>(f)};
    const auto p = reconstruct_path_from_distances(g, r, t);
    put<std::uint32_t>(out, p ? 0U : 1U);
    if (!p) continue;
    CHECK(validate_reconstructed_path(g, r, t, *p));
    put<std::uint32_t>(out, static_cast<std::uint32_t>(p->vertices.size()));
    for (const VertexId x : p->vertices) put<std::uint32_t>(out, x.value());
    for (const EdgeId x : p->edge_ids) put<std::uint64_t>(out, x.value());
    put<float>(out, p->cost);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  try {
    if (mode == "tiles") {
      tiles();
    } else if (mode == "paths") {
      paths();
    } else if (mode == "workspace") {
      workspace();
    } else if (mode == "paths-file" && argc == 4) {
      paths_file(argv[2], argv[3]);
    } else {
      std::fprintf(stderr, "usage: host_api_test tiles|paths|workspace|paths-file IN OUT\n");
      return 2;
    }
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "unexpected exception: %s\n", ex.what());
    ++failures;
  }
  std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
  return failures == 0 ? 0 : 1;
}
