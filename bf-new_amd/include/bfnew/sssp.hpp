// sssp.hpp -- distance-only results and path reconstruction on the host
// (reference: include/bfnew/sssp.hpp:14-61, src/path_reconstruct.cpp:20-257).
//
// The engines of this library produce labels on the device; this header
// turns a graph-sized label image back into a path with the reference's
// rule: from the target, follow tight incoming edges (fl(d[u] + w) == d[v],
// u finite) tried in stable-EdgeId order, never revisiting a vertex of the
// current path, backtracking out of dead ends, until a source is reached.
// The CPU shortest-path engines of the reference (dijkstra_oracle,
// synchronous_bellman_ford) are not part of this library: labels come from
// the device engines.
#pragma once

#include "bfnew/types.hpp"

#include <cstdint>
#include <optional>
#include <span>
#include <vector>

namespace bfnew {

class WeightedGraph;

struct SsspResult {
  std::vector<float> distances;   // one label per graph vertex
  std::vector<VertexId> sources;  // canonical (sorted, unique), each labelled 0
};

struct ReconstructedPath {
  std::vector<VertexId> vertices;  // source first, target last
  std::vector<EdgeId> edge_ids;    // vertices.size() - 1 stable edge ids
  float cost{};                    // left-to-right fp32 sum of the path's weights
};

// std::nullopt when the target's label is +inf.  Throws
// std::invalid_argument for a result whose shape is not a distance-only
// result of `graph` (size, canonical sources labelled 0), std::out_of_range
// for a target outside the graph, and std::logic_error when a finite label
// has no tight path back to a source.
[[nodiscard]] std::optional<ReconstructedPath> reconstruct_path_from_distances(
    const WeightedGraph& graph, const SsspResult& result, VertexId target);

// The path starts at a source, ends at `target`, is simple, follows the
// graph's edges by stable id, every edge is tight under result.distances,
// and cost equals both the fp32 sum of its weights and the target's label.
[[nodiscard]] bool validate_reconstructed_path(const WeightedGraph& graph, const SsspResult& result,
                                               VertexId target, const ReconstructedPath& path);

// Both values non-negative (or +inf) and equal, or finite and at most
// maximum_ulps representable floats apart.
[[nodiscard]] bool nonnegative_distance_within_ulps(float left, float right,
                                                    std::uint32_t maximum_ulps = 4U) noexcept;

}  // namespace bfnew
