// artifact_test.cpp -- BFGRPH07 / BFQRY007 artifacts.
//
//   artifact_test write DIR   writes DIR/graph.bfgrph and DIR/queries.bfqry
//                             for a small fixture and prints the fixture as
//                             JSON (tests/test_artifacts.py decodes the files
//                             independently and compares)
//   artifact_test check DIR   round trips and every rejection path
//   artifact_test recode IN OUT
//                             reads the graph artifact IN and writes it to OUT
//   artifact_test bench DIR X Y N
//                             write and read timing of the X x Y x N
//                             synthetic graph's input records
#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"
#include "bfnew/spatial.hpp"
#include "bfnew/synthetic.hpp"
#include "bfnew/workload_artifact.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
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

EdgeInputRecord edge(std::uint32_t s, std::uint32_t d, float w, std::uint32_t domain, std::uint32_t kind,
                     std::uint64_t record) {
  EdgeInputRecord e{};
  e.source = VertexId{s};
  e.destination = VertexId{d};
  e.weight = w;
  e.provenance = PhysicalProvenance{domain, kind, record};
  return e;
}

InputGraph fixture_input() {
  const ResourceClassId a{1U}, b{513U};
  std::vector<VertexMetadata> v{VertexMetadata::located(-3, 7, a),   VertexMetadata::located(12, 0, b),
                                VertexMetadata::unlocated(a),         VertexMetadata::located(25, 31, a),
                                VertexMetadata::located(-20, -9, b),  VertexMetadata::located(5, 5, a)};
  std::vector<EdgeInputRecord> e{edge(0, 1, 1.5F, 1, 2, 3),          edge(1, 3, 0.25F, 7, 0x80000001U, 1ULL << 40),
                                 edge(3, 2, 2.0F, 0, 0, 0),          edge(2, 4, 1e-3F, 2, 5, 99),
                                 edge(4, 5, 0.0F, 3, 3, 12345678901ULL), edge(5, 0, 7.75F, 4, 1, 17),
                                 edge(0, 1, 1.5F, 1, 2, 4)};
  return InputGraph{std::move(v), std::move(e)};
}

struct Fixture {
  InputGraph input;
  PartitionedGraph partitioned;
  std::vector<MappedRouteQuery> queries;
};

MappedTerminal terminal(const WeightedGraph& g, std::uint32_t old, const char* site, const char* pin) {
  MappedTerminal t;
  t.site = site;
  t.pin = pin;
  t.original_vertex = VertexId{old};
  t.vertex = g.old_to_new()[old];
  const VertexMetadata& m = g.vertices()[t.vertex.value()];
  t.x = m.x;
  t.y = m.y;
  t.has_location = m.has_location;
  t.owner_tile = g.owner_tiles()[t.vertex.value()];
  return t;
}

Fixture make_fixture() {
  const UniformGridPartitioner grid{SpatialOrderConfig{-32, -32, 8U, 8U}};
  Fixture f{fixture_input(), grid.partition(fixture_input()), {}};
  const WeightedGraph& g = f.partitioned.graph;
  const auto route = [&](std::uint32_t id, const char* net, std::vector<std::uint32_t> src,
                         std::vector<std::uint32_t> tgt, std::uint32_t padding, std::uint32_t generation) {
    MappedRouteQuery m;
    m.net_name = net;
    std::vector<VertexId> s, t;
    for (std::uint32_t x : src) {
      s.push_back(g.old_to_new()[x]);
      m.source_terminals.push_back(terminal(g, x, "SLICE_X1Y2", "A1"));
    }
    for (std::uint32_t x : tgt) {
      t.push_back(g.old_to_new()[x]);
      m.target_terminals.push_back(terminal(g, x, "SLICE_X7Y9", "BQ"));
    }
    m.query = make_route_query(QueryId{id}, g, s, t, padding, generation);
    f.queries.push_back(std::move(m));
  };
  route(11, "net/a", {0}, {3}, 0, 0);
  route(0xFFFFFFF0U, "", {5, 0, 5}, {2, 4}, 2, 7);  // duplicate terminal, spill target
  route(12, "n\xc3\xa9t \"c\"", {4}, {1}, 1, 3);
  return f;
}

std::vector<unsigned char> slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  return {std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
}

void spit(const std::string& path, const std::vector<unsigned char>& bytes) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  out.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
}

std::string json_ids(const std::vector<VertexId>& v) {
  std::string s = "[";
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i].value());
  return s + "]";
}
std::string json_tiles(const std::vector<TileId>& v) {
  std::string s = "[";
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i].value());
  return s + "]";
}
std::string json_words(const std::vector<std::uint32_t>& v) {
  std::string s = "[";
  for (std::size_t i = 0; i < v.size(); ++i) s += (i ? ", " : "") + std::to_string(v[i]);
  return s + "]";
}
std::string json_text(const std::string& t) {
  std::string s = "[";
  for (std::size_t i = 0; i < t.size(); ++i)
    s += (i ? ", " : "") + std::to_string(static_cast<unsigned char>(t[i]));
  return s + "]";
}
std::string json_terminal(const MappedTerminal& t) {
  return "{\"site\": " + json_text(t.site) + ", \"pin\": " + json_text(t.pin) +
         ", \"original_vertex\": " + std::to_string(t.original_vertex.value()) +
         ", \"vertex\": " + std::to_string(t.vertex.value()) + ", \"x\": " + std::to_string(t.x) +
         ", \"y\": " + std::to_string(t.y) + ", \"has_location\": " + (t.has_location ? "1" : "0") +
         ", \"owner_tile\": " + std::to_string(t.owner_tile.value()) + "}";
}

int write_mode(const std::string& dir) {
  const Fixture f = make_fixture();
  write_graph_artifact(dir + "/graph.bfgrph", f.input);
  write_query_artifact(dir + "/queries.bfqry", f.queries);
  std::string out = "{\"vertices\": [";
  const auto vs = f.input.vertices();
  for (std::size_t i = 0; i < vs.size(); ++i) {
    out += (i ? ", " : "") + std::string("[") + std::to_string(vs[i].x) + ", " + std::to_string(vs[i].y) + ", " +
           (vs[i].has_location ? "1" : "0") + ", " + std::to_string(vs[i].resource_class.value()) + "]";
  }
  out += "], \"edges\": [";
  const auto es = f.input.edges();
  for (std::size_t i = 0; i < es.size(); ++i) {
    std::uint32_t bits;
    std::memcpy(&bits, &es[i].weight, 4);
    out += (i ? ", " : "") + std::string("[") + std::to_string(es[i].source.value()) + ", " +
           std::to_string(es[i].destination.value()) + ", " + std::to_string(bits) + ", " +
           std::to_string(es[i].provenance.domain) + ", " + std::to_string(es[i].provenance.kind_and_flags) + ", " +
           std::to_string(es[i].provenance.source_record) + "]";
  }
  out += "], \"queries\": [";
  for (std::size_t i = 0; i < f.queries.size(); ++i) {
    const MappedRouteQuery& m = f.queries[i];
    const RouteQuery& q = m.query;
    out += (i ? ", " : "") + std::string("{\"net_name\": ") + json_text(m.net_name) +
           ", \"query_id\": " + std::to_string(q.query_id.value()) +
           ", \"source_terminals\": " + json_ids(q.source_terminals) +
           ", \"target_terminals\": " + json_ids(q.target_terminals) + ", \"sources\": " + json_ids(q.sources) +
           ", \"targets\": " + json_ids(q.targets) +
           ", \"source_terminal_to_source\": " + json_words(q.source_terminal_to_source) +
           ", \"target_terminal_to_target\": " + json_words(q.target_terminal_to_target) +
           ", \"selected_tiles\": " + json_tiles(q.selected_tiles) +
           ", \"expansion_generation\": " + std::to_string(q.expansion_generation) + ", \"source_descriptors\": [";
    for (std::size_t k = 0; k < m.source_terminals.size(); ++k)
      out += (k ? ", " : "") + json_terminal(m.source_terminals[k]);
    out += "], \"target_descriptors\": [";
    for (std::size_t k = 0; k < m.target_terminals.size(); ++k)
      out += (k ? ", " : "") + json_terminal(m.target_terminals[k]);
    out += "]}";
  }
  out += "]}";
  std::printf("%s\n", out.c_str());
  return 0;
}

void check_mode(const std::string& dir) {
  const Fixture f = make_fixture();
  const WeightedGraph& g = f.partitioned.graph;
  const std::string gp = dir + "/check.bfgrph", qp = dir + "/check.bfqry";
  write_graph_artifact(gp, f.input);
  write_query_artifact(qp, f.queries);

  // Round trips: the same records, the same queries, and the same bytes when
  // written again.
  const InputGraph back = read_graph_artifact(gp);
  CHECK(back.vertex_count() == f.input.vertex_count() && back.edge_count() == f.input.edge_count());
  CHECK(std::equal(back.vertices().begin(), back.vertices().end(), f.input.vertices().begin()));
  bool same_edges = true;
  for (std::size_t i = 0; i < back.edges().size(); ++i) {
    const EdgeInputRecord &x = back.edges()[i], &y = f.input.edges()[i];
    same_edges = same_edges && x.source == y.source && x.destination == y.destination &&
                 std::memcmp(&x.weight, &y.weight, 4) == 0 && x.provenance == y.provenance;
  }
  CHECK(same_edges);
  const std::vector<MappedRouteQuery> qback = read_query_artifact(qp, g);
  CHECK(qback == f.queries);
  write_graph_artifact(dir + "/again.bfgrph", back);
  write_query_artifact(dir + "/again.bfqry", qback);
  CHECK(files_are_byte_identical(gp, dir + "/again.bfgrph"));
  CHECK(files_are_byte_identical(qp, dir + "/again.bfqry"));
  CHECK(!files_are_byte_identical(gp, qp));
  // The spatial graph built from the read records is the original's.
  const UniformGridPartitioner grid{SpatialOrderConfig{-32, -32, 8U, 8U}};
  const PartitionedGraph rebuilt = grid.partition(back);
  CHECK(rebuilt.graph.owner_tiles().size() == g.owner_tiles().size() &&
        std::equal(rebuilt.graph.owner_tiles().begin(), rebuilt.graph.owner_tiles().end(), g.owner_tiles().begin()));
  // An empty query list and a graph without edges.
  write_query_artifact(dir + "/none.bfqry", {});
  CHECK(read_query_artifact(dir + "/none.bfqry", g).empty());
  const InputGraph lone{{VertexMetadata::located(1, 1, ResourceClassId{2U})}, {}};
  write_graph_artifact(dir + "/lone.bfgrph", lone);
  CHECK(read_graph_artifact(dir + "/lone.bfgrph").vertex_count() == 1U);
  CHECK(slurp(dir + "/lone.bfgrph").size() == 24U + 11U);

  // Rejections.
  const std::vector<unsigned char> graph_bytes = slurp(gp), query_bytes = slurp(qp);
  const auto graph_fails = [&](std::vector<unsigned char> bytes, const char* expected) {
    spit(dir + "/bad.bfgrph", bytes);
    try {
      (void)read_graph_artifact(dir + "/bad.bfgrph");
    } catch (const std::exception& e) {
      return std::strstr(e.what(), expected) != nullptr;
    }
    return false;
  };
  const auto query_fails = [&](std::vector<unsigned char> bytes, const char* expected) {
    spit(dir + "/bad.bfqry", bytes);
    try {
      (void)read_query_artifact(dir + "/bad.bfqry", g);
    } catch (const std::exception& e) {
      return std::strstr(e.what(), expected) != nullptr;
    }
    return false;
  };
  auto b = graph_bytes;
  b[7] = '8';
  CHECK(graph_fails(b, "magic"));
  b = graph_bytes;
  b[8] = 2;
  CHECK(graph_fails(b, "version"));
  b = graph_bytes;
  b.pop_back();
  CHECK(graph_fails(b, "truncated"));
  b = graph_bytes;
  b.push_back(0);
  CHECK(graph_fails(b, "trailing"));
  b = graph_bytes;
  b[24 + 8] = 2;  // first vertex's has_location
  CHECK(graph_fails(b, "boolean"));
  b = graph_bytes;
  b[24 + 11 * 6 + 4] = 9;  // first edge's destination out of range
  CHECK(graph_fails(b, "destination"));
  b = graph_bytes;
  const std::uint32_t nan_bits = 0x7FC00000U;
  std::memcpy(&b[24 + 11 * 6 + 8], &nan_bits, 4);
  CHECK(graph_fails(b, "finite"));
  CHECK(throws<std::runtime_error>([&] { (void)read_graph_artifact(dir + "/missing.bfgrph"); }));
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/missing.bfqry", g); }));
  b = query_bytes;
  b[0] = 'X';
  CHECK(query_fails(b, "magic"));
  b = query_bytes;
  b.pop_back();
  CHECK(query_fails(b, "truncated"));
  b = query_bytes;
  b.push_back(1);
  CHECK(query_fails(b, "trailing"));

  // Descriptor and query consistency.
  auto wrong = f.queries;
  wrong[0].source_terminals[0].x += 1;
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
  wrong = f.queries;
  wrong[0].source_terminals[0].original_vertex = VertexId{1U};
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
  wrong = f.queries;
  wrong[1].target_terminals.pop_back();
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
  wrong = f.queries;
  wrong[2].query.selected_tiles.clear();
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
  wrong = f.queries;
  std::swap(wrong[1].source_terminals[0], wrong[1].source_terminals[1]);
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
  wrong = f.queries;
  wrong[0].source_terminals[0].has_location = false;
  write_query_artifact(dir + "/bad.bfqry", wrong);
  CHECK(throws<std::runtime_error>([&] { (void)read_query_artifact(dir + "/bad.bfqry", g); }));
}

int bench_mode(const std::string& dir, std::uint32_t x, std::uint32_t y, std::uint32_t n) {
  SyntheticTileGridConfig cfg;
  cfg.tiles_x = x;
  cfg.tiles_y = y;
  cfg.nodes_per_tile = n;
  const InputGraph input = SyntheticTileGraph::generate(cfg).input_graph();
  const std::string path = dir + "/bench.bfgrph";
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  write_graph_artifact(path, input);
  const auto t1 = Clock::now();
  const InputGraph back = read_graph_artifact(path);
  const auto t2 = Clock::now();
  const double bytes = 24.0 + 11.0 * input.vertex_count() + 28.0 * static_cast<double>(input.edge_count());
  const double w_s = std::chrono::duration<double>(t1 - t0).count();
  const double r_s = std::chrono::duration<double>(t2 - t1).count();
  CHECK(back.edge_count() == input.edge_count() && back.vertex_count() == input.vertex_count());
  std::printf("{\"workload\": \"artifact_%ux%ux%u\", \"vertices\": %u, \"edges\": %llu, \"bytes\": %.0f, "
              "\"write_s\": %.3f, \"read_s\": %.3f, \"write_gb_per_s\": %.3f, \"read_gb_per_s\": %.3f}\n",
              x, y, n, input.vertex_count(), static_cast<unsigned long long>(input.edge_count()), bytes, w_s, r_s,
              bytes / w_s / 1e9, bytes / r_s / 1e9);
  if (!std::getenv("KEEP_ARTIFACT")) std::remove(path.c_str());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: artifact_test write|check|bench DIR [X Y N]\n");
    return 2;
  }
  try {
    const std::string mode = argv[1], dir = argv[2];
    if (mode == "write") return write_mode(dir);
    if (mode == "check") {
      check_mode(dir);
      std::printf("check: %d failure(s)\n", failures);
      return failures ? 1 : 0;
    }
    if (mode == "recode" && argc >= 4) {
      const InputGraph g = read_graph_artifact(dir);
      write_graph_artifact(argv[3], g);
      return 0;
    }
    if (mode == "bench" && argc >= 6)
      return bench_mode(dir, static_cast<std::uint32_t>(std::atoi(argv[3])),
                        static_cast<std::uint32_t>(std::atoi(argv[4])), static_cast<std::uint32_t>(std::atoi(argv[5])));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 3;
  }
  std::fprintf(stderr, "unknown mode\n");
  return 2;
}
