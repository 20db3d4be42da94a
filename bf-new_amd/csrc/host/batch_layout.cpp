// batch_layout.cpp -- host description of one batch of route queries
// (reference: src/batch_layout.cpp:1056-1200 prepares the same lane fields
// from a plan entry; here the lanes are the given queries in order).
#include "bfnew/batch_layout.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace bfnew {

void prepare_batch_device_description(const WeightedGraph& graph,
                                      std::span<const RouteQuery> queries,
                                      std::uint32_t lane_width, BatchDeviceDescription& out) {
  if (queries.empty() || lane_width == 0U || lane_width > maximum_batch_lanes ||
      queries.size() > lane_width)
    throw std::invalid_argument{"batch description: need 1 <= queries <= lane_width <= 32"};
  const std::uint32_t tile_count = static_cast<std::uint32_t>(graph.tile_vertex_offsets().size()) - 1U;
  for (std::size_t l = 0; l < queries.size(); ++l) {
    const RouteQueryValidationResult r = validate_route_query(graph, queries[l]);
    if (!r.ok())
      throw std::invalid_argument{"batch description: invalid route query in lane " +
                                  std::to_string(l)};
  }
  out = BatchDeviceDescription{};
  out.lane_width = lane_width;
  out.valid_lane_mask = queries.size() == 32U ? 0xFFFFFFFFu : ((1U << queries.size()) - 1U);
  out.query_ids_by_lane.assign(lane_width, 0U);
  out.expansion_generations_by_lane.assign(lane_width, 0U);
  out.selected_vertex_counts_by_lane.assign(lane_width, 0U);
  out.selected_edge_estimates_by_lane.assign(lane_width, 0U);
  out.source_offsets.assign(lane_width + 1U, 0U);
  out.target_offsets.assign(lane_width + 1U, 0U);
  out.tile_lane_masks.assign(tile_count, 0U);
  for (std::uint32_t l = 0; l < lane_width; ++l) {
    if (l < queries.size()) {
      const RouteQuery& q = queries[l];
      out.query_ids_by_lane[l] = q.query_id.value();
      out.expansion_generations_by_lane[l] = q.expansion_generation;
      for (VertexId v : q.sources) out.sources.push_back(v.value());
      for (VertexId v : q.targets) out.targets.push_back(v.value());
      for (TileId t : q.selected_tiles) out.tile_lane_masks[t.value()] |= 1U << l;
      out.selected_vertex_counts_by_lane[l] = estimate_selected_vertex_count(graph, q.selected_tiles);
      out.selected_edge_estimates_by_lane[l] = estimate_selected_edge_count(graph, q.selected_tiles);
    }
    out.source_offsets[l + 1U] = static_cast<std::uint32_t>(out.sources.size());
    out.target_offsets[l + 1U] = static_cast<std::uint32_t>(out.targets.size());
  }
  const auto offsets = graph.tile_vertex_offsets();
  for (std::uint32_t t = 0; t < tile_count; ++t) {
    if (out.tile_lane_masks[t] == 0U) continue;
    out.union_tiles.push_back(t);
    out.selected_vertex_ranges.push_back(BatchVertexRange{static_cast<std::uint32_t>(offsets[t]),
                                                          static_cast<std::uint32_t>(offsets[t + 1U]),
                                                          out.tile_lane_masks[t]});
  }
}

BatchLayoutValidationResult validate_batch_device_description(
    const WeightedGraph& graph, const BatchDeviceDescription& d) noexcept {
  using C = BatchLayoutValidationErrorCode;
  auto bad = [](C code, std::size_t pos = BatchLayoutValidationResult::no_position) {
    return BatchLayoutValidationResult{code, pos};
  };
  const auto offsets = graph.tile_vertex_offsets();
  if (offsets.empty()) return bad(C::invalid_batch_shape);
  const std::uint32_t tile_count = static_cast<std::uint32_t>(offsets.size()) - 1U;
  const std::uint32_t V = graph.vertex_count();
  const std::uint32_t W = d.lane_width;
  if (W == 0U || W > maximum_batch_lanes) return bad(C::invalid_batch_shape);
  const LaneMask lanes = W == 32U ? 0xFFFFFFFFu : ((1U << W) - 1U);
  if (d.valid_lane_mask == 0U || (d.valid_lane_mask & ~lanes)) return bad(C::invalid_batch_shape);
  if (d.query_ids_by_lane.size() != W || d.expansion_generations_by_lane.size() != W)
    return bad(C::invalid_lane_identity);
  if (d.reached_lane_mask != 0U || d.miss_lane_mask != 0U) return bad(C::nonzero_initial_result_mask);
  auto check_slices = [&](const std::vector<std::uint32_t>& off, const std::vector<std::uint32_t>& val,
                          C code) -> BatchLayoutValidationResult {
    if (off.size() != W + 1U || off[0] != 0U || off[W] != val.size()) return bad(code);
    for (std::uint32_t l = 0; l < W; ++l) {
      if (off[l + 1U] < off[l]) return bad(code, l);
      const bool valid = (d.valid_lane_mask >> l) & 1U;
      if (valid == (off[l + 1U] == off[l])) return bad(code, l);  // valid lanes non-empty, others empty
      for (std::uint32_t i = off[l]; i < off[l + 1U]; ++i) {
        if (val[i] >= V) return bad(C::invalid_terminal_payload, i);
        if (i > off[l] && val[i] <= val[i - 1U]) return bad(C::invalid_terminal_payload, i);
        const std::uint32_t owner = graph.owner_tiles()[val[i]].value();
        if (!((d.tile_lane_masks.size() == tile_count) && ((d.tile_lane_masks[owner] >> l) & 1U)))
          return bad(C::invalid_terminal_payload, i);
      }
    }
    return {};
  };
  if (d.tile_lane_masks.size() != tile_count) return bad(C::invalid_tile_lane_masks);
  if (auto r = check_slices(d.source_offsets, d.sources, C::invalid_source_offsets); !r.ok()) return r;
  if (auto r = check_slices(d.target_offsets, d.targets, C::invalid_target_offsets); !r.ok()) return r;
  std::size_t u = 0;
  for (std::uint32_t t = 0; t < tile_count; ++t) {
    const LaneMask m = d.tile_lane_masks[t];
    if (m & ~d.valid_lane_mask) return bad(C::invalid_tile_lane_masks, t);
    if (m == 0U) continue;
    if (u >= d.union_tiles.size() || d.union_tiles[u] != t) return bad(C::invalid_union_tiles, u);
    if (u >= d.selected_vertex_ranges.size()) return bad(C::invalid_vertex_ranges, u);
    const BatchVertexRange& r = d.selected_vertex_ranges[u];
    if (r.begin != offsets[t] || r.end != offsets[t + 1U] || r.lane_mask != m)
      return bad(C::invalid_vertex_ranges, u);
    ++u;
  }
  if (u != d.union_tiles.size()) return bad(C::invalid_union_tiles, u);
  if (u != d.selected_vertex_ranges.size()) return bad(C::invalid_vertex_ranges, u);
  return {};
}

}  // namespace bfnew
