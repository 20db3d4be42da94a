// path_reconstruct.cpp -- paths from label images on the host.
//
// Rule (reference src/path_reconstruct.cpp:44-191 and
// src/compact_paths.cpp:94-218): walk back from the target over tight
// incoming edges, fl(d[u] + w) == d[v] with d[u] finite, trying them in
// stable-EdgeId order; a vertex already on the current path is skipped (so
// zero-weight cycles cannot loop) and a vertex with no usable edge is
// abandoned (backtracking); the first source reached ends the path.  With
// strictly positive weights the labels strictly decrease along tight edges
// and the walk never backtracks.
//
// One search object serves every target of a call: its on-path marks are a
// byte per vertex allocated once, and exactly the current path is marked at
// any time, so a finished search clears only its own path.  Candidate edges
// of all open frames share one stack-like pool.
#include "bfnew/compact_paths.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/query.hpp"
#include "bfnew/sssp.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace bfnew {
namespace {

struct TightInEdge {
  EdgeId edge;
  VertexId from;
  float weight;
};

class TightPathSearch {
 public:
  explicit TightPathSearch(std::uint32_t vertex_count) : on_path_(vertex_count, 0U) {}

  // Searches from `target`; `admit(u)` says whether u may enter the path,
  // `is_source(v)` ends it.  On success `vertices` holds target ... source
  // and `steps[i]` is the edge from vertices[i + 1] into vertices[i].
  template <typename Admit, typename IsSource>
  bool find(const WeightedGraph& graph, std::span<const float> d, VertexId target, Admit admit,
            IsSource is_source, std::vector<VertexId>& vertices, std::vector<TightInEdge>& steps) {
    vertices.assign(1U, target);
    steps.clear();
    pool_.clear();
    frames_.clear();
    on_path_[target.value()] = 1U;
    open(graph, d, target, admit);
    while (!frames_.empty()) {
      const VertexId v = vertices.back();
      if (is_source(v)) {
        for (const VertexId u : vertices) on_path_[u.value()] = 0U;
        return true;
      }
      Frame& f = frames_.back();
      bool advanced = false;
      while (f.next < f.end) {
        const TightInEdge c = pool_[f.next++];
        if (on_path_[c.from.value()]) continue;
        steps.push_back(c);
        vertices.push_back(c.from);
        on_path_[c.from.value()] = 1U;
        open(graph, d, c.from, admit);  // may move frames_: f is not used again
        advanced = true;
        break;
      }
      if (advanced) continue;
      on_path_[v.value()] = 0U;
      pool_.resize(f.begin);
      frames_.pop_back();
      vertices.pop_back();
      if (!steps.empty()) steps.pop_back();
    }
    return false;
  }

 private:
  struct Frame {
    std::size_t begin, end, next;
  };

  template <typename Admit>
  void open(const WeightedGraph& graph, std::span<const float> d, VertexId v, Admit admit) {
    const IncomingCscView in = graph.incoming();
    const std::size_t begin = pool_.size();
    const float dv = d[v.value()];
    for (std::uint64_t p = in.column_offsets[v.value()]; p < in.column_offsets[v.value() + 1U]; ++p) {
      const VertexId u = in.sources[p];
      const float du = d[u.value()];
      if (!std::isfinite(du) || !admit(u)) continue;
      if (du + in.weights[p] == dv) pool_.push_back(TightInEdge{in.edge_ids[p], u, in.weights[p]});
    }
    std::sort(pool_.begin() + static_cast<std::ptrdiff_t>(begin), pool_.end(),
              [](const TightInEdge& a, const TightInEdge& b) { return a.edge < b.edge; });
    frames_.push_back(Frame{begin, pool_.size(), begin});
  }

  std::vector<std::uint8_t> on_path_;
  std::vector<TightInEdge> pool_;
  std::vector<Frame> frames_;
};

bool canonical_sources_at_zero(const WeightedGraph& graph, std::span<const float> d,
                               std::span<const VertexId> sources) {
  if (sources.empty()) return false;
  for (std::size_t i = 0; i < sources.size(); ++i) {
    if (sources[i].value() >= graph.vertex_count() || (i && !(sources[i - 1] < sources[i])))
      return false;
    if (d[sources[i].value()] != 0.0F) return false;
  }
  return true;
}

bool known_disposition(ExpansionQueryDisposition d) noexcept {
  return static_cast<std::uint32_t>(d) <= static_cast<std::uint32_t>(ExpansionQueryDisposition::engine_failure);
}

// The stable edge `e` in the CSR row of `from`: its destination and weight.
bool find_out_edge(const WeightedGraph& graph, VertexId from, EdgeId e, VertexId& to, float& w) {
  const OutgoingCsrView out = graph.outgoing();
  for (std::uint64_t p = out.row_offsets[from.value()]; p < out.row_offsets[from.value() + 1U]; ++p)
    if (out.edge_ids[p] == e) {
      to = out.destinations[p];
      w = out.weights[p];
      return true;
    }
  return false;
}

CompactPathPayload empty_payload(const RouteQuery& query) {
  CompactPathPayload p;
  p.query_id = query.query_id;
  p.expansion_generation = query.expansion_generation;
  p.target_terminal_to_target = query.target_terminal_to_target;
  p.targets.reserve(query.targets.size());
  return p;
}

}  // namespace

std::optional<ReconstructedPath> reconstruct_path_from_distances(const WeightedGraph& graph,
                                                                 const SsspResult& result,
                                                                 VertexId target) {
  if (result.distances.size() != graph.vertex_count() ||
      !canonical_sources_at_zero(graph, result.distances, result.sources))
    throw std::invalid_argument{"path reconstruction requires a distance-only result of this graph"};
  if (target.value() >= graph.vertex_count())
    throw std::out_of_range{"path target is outside the graph"};
  if (!std::isfinite(result.distances[target.value()])) return std::nullopt;
  TightPathSearch search(graph.vertex_count());
  std::vector<VertexId> vertices;
  std::vector<TightInEdge> steps;
  const bool found = search.find(
      graph, result.distances, target, [](VertexId) { return true; },
      [&](VertexId v) { return std::binary_search(result.sources.begin(), result.sources.end(), v); },
      vertices, steps);
  if (!found) throw std::logic_error{"a finite label has no tight path back to a source"};
  ReconstructedPath path;
  path.vertices.assign(vertices.rbegin(), vertices.rend());
  path.edge_ids.reserve(steps.size());
  path.cost = 0.0F;
  for (auto s = steps.rbegin(); s != steps.rend(); ++s) {
    path.edge_ids.push_back(s->edge);
    path.cost = path.cost + s->weight;
  }
  return path;
}

bool validate_reconstructed_path(const WeightedGraph& graph, const SsspResult& result,
                                 VertexId target, const ReconstructedPath& path) {
  const std::uint32_t V = graph.vertex_count();
  if (result.distances.size() != V || !canonical_sources_at_zero(graph, result.distances, result.sources) ||
      target.value() >= V || path.vertices.empty() || path.vertices.back() != target ||
      path.edge_ids.size() + 1U != path.vertices.size() ||
      !std::binary_search(result.sources.begin(), result.sources.end(), path.vertices.front()))
    return false;
  std::vector<VertexId> sorted = path.vertices;
  std::sort(sorted.begin(), sorted.end());
  if (sorted.back().value() >= V || std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return false;
  float cost = 0.0F;
  for (std::size_t i = 0; i < path.edge_ids.size(); ++i) {
    if (path.edge_ids[i].value() >= graph.edge_count()) return false;
    VertexId to;
    float w = 0.0F;
    if (!find_out_edge(graph, path.vertices[i], path.edge_ids[i], to, w) || to != path.vertices[i + 1U])
      return false;
    if (result.distances[path.vertices[i].value()] + w != result.distances[to.value()]) return false;
    cost = cost + w;
  }
  return cost == path.cost && path.cost == result.distances[target.value()];
}

bool nonnegative_distance_within_ulps(float left, float right, std::uint32_t maximum_ulps) noexcept {
  if (std::isnan(left) || std::isnan(right) || left < 0.0F || right < 0.0F) return false;
  if (left == right) return true;
  if (!std::isfinite(left) || !std::isfinite(right)) return false;
  // Non-negative floats order like their bit patterns.
  const std::uint32_t a = std::bit_cast<std::uint32_t>(left), b = std::bit_cast<std::uint32_t>(right);
  return (a > b ? a - b : b - a) <= maximum_ulps;
}

CompactPathPayload reconstruct_compact_path_payload(const WeightedGraph& graph, const RouteQuery& query,
                                                    std::span<const float> final_distances) {
  if (!weighted_graph_validation(graph).empty() || !validate_route_query(graph, query).ok())
    throw std::invalid_argument{"compact reconstruction requires a valid graph and query"};
  const std::uint32_t V = graph.vertex_count();
  bool image_ok = final_distances.size() == V;
  for (std::size_t i = 0; image_ok && i < final_distances.size(); ++i)
    image_ok = !std::isnan(final_distances[i]) && final_distances[i] >= 0.0F;
  for (const VertexId s : query.sources)
    image_ok = image_ok && final_distances[s.value()] == 0.0F;
  if (!image_ok)
    throw std::invalid_argument{"compact reconstruction requires a valid graph-sized label image"};

  std::vector<std::uint8_t> tile_selected(graph.tile_coordinates().size(), 0U);
  for (const TileId t : query.selected_tiles) tile_selected[t.value()] = 1U;
  const auto owner = graph.owner_tiles();
  auto admit = [&](VertexId u) { return tile_selected[owner[u.value()].value()] != 0U; };
  auto is_source = [&](VertexId v) {
    return std::binary_search(query.sources.begin(), query.sources.end(), v);
  };

  CompactPathPayload payload = empty_payload(query);
  TightPathSearch search(V);
  std::vector<VertexId> vertices;
  std::vector<TightInEdge> steps;
  for (const VertexId target : query.targets) {
    CompactTargetPath out;
    out.summary.target = target;
    out.summary.distance = final_distances[target.value()];
    if (!std::isfinite(out.summary.distance)) {
      out.summary.reached = CompactTargetReachStatus::not_reached;
      out.summary.reconstruction = CompactPathStatus::unreachable;
      payload.targets.push_back(std::move(out));
      continue;
    }
    out.summary.reached = CompactTargetReachStatus::reached;
    if (!search.find(graph, final_distances, target, admit, is_source, vertices, steps)) {
      out.summary.reconstruction = CompactPathStatus::no_tight_path;
    } else if (steps.size() > std::numeric_limits<std::uint32_t>::max()) {
      out.summary.reconstruction = CompactPathStatus::path_length_overflow;
    } else {
      out.vertices.assign(vertices.rbegin(), vertices.rend());
      out.distance_labels.reserve(out.vertices.size());
      for (const VertexId v : out.vertices) out.distance_labels.push_back(final_distances[v.value()]);
      out.edge_ids.reserve(steps.size());
      for (auto s = steps.rbegin(); s != steps.rend(); ++s) out.edge_ids.push_back(s->edge);
      out.summary.selected_source = out.vertices.front();
      out.summary.has_selected_source = 1U;
      out.summary.path_length = static_cast<std::uint32_t>(out.edge_ids.size());
      out.summary.reconstruction = CompactPathStatus::complete;
    }
    payload.targets.push_back(std::move(out));
  }
  if (!detail::validate_compact_payload_trusted_graph(graph, query, payload).ok())
    throw std::logic_error{"compact reconstruction produced an invalid payload"};
  return payload;
}

CompactQueryResult reconstruct_compact_query_paths(const WeightedGraph& graph, const RouteQuery& query,
                                                   std::span<const float> final_distances,
                                                   ExpansionQueryDisposition disposition) {
  if (!known_disposition(disposition))
    throw std::invalid_argument{"compact result disposition is invalid"};
  CompactQueryResult result = make_compact_query_result(
      reconstruct_compact_path_payload(graph, query, final_distances), disposition);
  if (!validate_compact_query_result(graph, query, result).ok())
    throw std::logic_error{"reconstructed paths disagree with the query's disposition"};
  return result;
}

CompactQueryResult make_failed_compact_query_result(const RouteQuery& query,
                                                    ExpansionQueryDisposition disposition) {
  if (!known_disposition(disposition) || disposition == ExpansionQueryDisposition::reached ||
      disposition == ExpansionQueryDisposition::unreachable_in_full_region)
    throw std::invalid_argument{"only a failed query publishes a label-free result"};
  CompactPathPayload payload = empty_payload(query);
  for (const VertexId target : query.targets) {
    CompactTargetPath t;
    t.summary.target = target;
    t.summary.reconstruction = CompactPathStatus::query_terminal_failure;
    payload.targets.push_back(std::move(t));
  }
  return make_compact_query_result(std::move(payload), disposition);
}

std::vector<CompactQueryResult> extract_host_compact_paths(const WeightedGraph& graph,
                                                           std::span<ExpansionQueryOutcome> outcomes) {
  if (!weighted_graph_validation(graph).empty())
    throw std::invalid_argument{"compact extraction requires a valid graph"};
  for (std::size_t i = 0; i < outcomes.size(); ++i)
    if (!validate_route_query(graph, outcomes[i].final_query).ok() ||
        (i && !(outcomes[i - 1].final_query.query_id < outcomes[i].final_query.query_id)))
      throw std::invalid_argument{"compact extraction requires canonical, valid query outcomes"};
  std::vector<CompactQueryResult> results;
  results.reserve(outcomes.size());
  for (ExpansionQueryOutcome& o : outcomes) {
    const RouteQuery& q = o.final_query;
    CompactQueryResult r;
    if (o.compact_paths.has_value()) {
      r = make_compact_query_result(std::move(*o.compact_paths), o.disposition);
      o.compact_paths.reset();
      if (!validate_compact_query_result(graph, q, r).ok())
        throw std::logic_error{"a retained compact payload fails its checks"};
      if (!o.final_distances.empty() &&
          !validate_compact_query_result_against_distances(graph, q, r, o.final_distances).ok())
        throw std::logic_error{"a retained compact payload disagrees with the outcome's labels"};
    } else if (!o.final_distances.empty()) {
      const bool classified = o.disposition == ExpansionQueryDisposition::reached ||
                              o.disposition == ExpansionQueryDisposition::unreachable_in_full_region;
      r = classified ? reconstruct_compact_query_paths(graph, q, o.final_distances, o.disposition)
                     : make_failed_compact_query_result(q, o.disposition);
    } else {
      if (o.reached()) throw std::logic_error{"a reached outcome has neither paths nor labels"};
      r = make_failed_compact_query_result(q, o.disposition);
    }
    std::vector<float>{}.swap(o.final_distances);
    results.push_back(std::move(r));
  }
  return results;
}

}  // namespace bfnew
