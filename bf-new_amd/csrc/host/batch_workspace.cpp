// batch_workspace.cpp -- run images of a lane-batched description, the
// compact vertex mapping, and the workspace byte model over them
// (reference behaviour: src/batch_layout.cpp:373-603, 1056-1470 and
// src/batch_workspace.cpp:13-1021).
//
// All byte totals are checked against 64-bit overflow.  The preparation
// counters model a reused description: a retained run image and a compact
// mapping keep their storage between batches and clear only the entries the
// previous batch wrote (its ledger), so a warm preparation writes
// cleared + written entries and a cold one initialises the whole image.
#include "bfnew/batch_workspace.hpp"

#include <algorithm>
#include <array>
#include <bit>
#include <limits>
#include <stdexcept>
#include <utility>

namespace bfnew {
namespace {

constexpr std::uint64_t kMax = std::numeric_limits<std::uint64_t>::max();

std::uint64_t add(std::uint64_t a, std::uint64_t b) {
  if (a > kMax - b) throw std::overflow_error{"batch workspace byte total overflow"};
  return a + b;
}

std::uint64_t mul(std::uint64_t a, std::uint64_t b) {
  if (b != 0U && a > kMax / b) throw std::overflow_error{"batch workspace component overflow"};
  return a * b;
}

std::uint32_t narrow32(std::uint64_t x, const char* what) {
  if (x > std::numeric_limits<std::uint32_t>::max()) throw std::overflow_error{what};
  return static_cast<std::uint32_t>(x);
}

bool known(BatchRunRepresentation r) {
  return r == BatchRunRepresentation::retained_per_run_masks ||
         r == BatchRunRepresentation::compact_nonzero_descriptors ||
         r == BatchRunRepresentation::device_materialized_run_masks;
}

bool known(BatchVertexStorageStrategy s) {
  return s == BatchVertexStorageStrategy::full_graph_vertex_major ||
         s == BatchVertexStorageStrategy::compact_union_tiles;
}

bool supported_width(std::uint32_t w) { return w == 1U || w == 8U || w == 16U || w == 32U; }

// Clears the previous batch's nonzero retained masks (sorted, unique run
// ids that must still be nonzero) and counts the clears.
void clear_ledger(std::vector<LaneMask>& masks, std::vector<std::uint32_t>& touched,
                  std::uint64_t& all, std::uint64_t& orientation) {
  for (std::size_t i = 0; i < touched.size(); ++i) {
    const std::uint32_t run = touched[i];
    if (run >= masks.size() || (i && run <= touched[i - 1]) || masks[run] == 0U)
      throw std::invalid_argument{"reused retained run-mask ledger is inconsistent"};
    masks[run] = 0U;
    all = add(all, 1U);
    orientation = add(orientation, 1U);
  }
  touched.clear();
}

// Admitted lanes and edges over one orientation's runs.
struct Coverage {
  std::uint64_t visited{}, active{}, lane_edge_pairs{}, union_edges{};
  std::array<std::uint64_t, maximum_batch_lanes> edges_by_lane{};
};

// One orientation of the run images: CSR (out-runs, remote = destination
// tile) or CSC (in-runs, remote = source tile).
struct Orientation {
  bool csr;
  const std::vector<EdgeOffset>& vertex_runs;
  const std::vector<EdgeOffset>& run_edges;
  const std::vector<TileId>& run_tiles;
};

void record(const Orientation& o, std::size_t run, LaneMask mask, BatchDeviceDescription& d) {
  if (mask == 0U) return;
  BatchRunPreparationReport& r = d.run_report;
  switch (d.run_representation) {
    case BatchRunRepresentation::retained_per_run_masks:
      (o.csr ? d.csr_run_lane_masks : d.csc_run_lane_masks)[run] = mask;
      (o.csr ? d.touched_csr_runs : d.touched_csc_runs).push_back(narrow32(run, "run id exceeds 32 bits"));
      r.retained_entries_written = add(r.retained_entries_written, 1U);
      (o.csr ? r.csr_retained_entries_written : r.csc_retained_entries_written) += 1U;
      break;
    case BatchRunRepresentation::compact_nonzero_descriptors:
      (o.csr ? d.csr_run_descriptors : d.csc_run_descriptors)
          .push_back(RunLaneMaskDescriptor{narrow32(run, "run id exceeds 32 bits"), mask});
      r.descriptor_entries_written = add(r.descriptor_entries_written, 1U);
      (o.csr ? r.csr_descriptor_entries_written : r.csc_descriptor_entries_written) += 1U;
      break;
    case BatchRunRepresentation::device_materialized_run_masks:
      break;
  }
}

void cover_vertex(const Orientation& o, std::uint64_t edge_count, std::uint32_t v, LaneMask lanes,
                  BatchDeviceDescription& d, Coverage& c) {
  if (v + 1ULL >= o.vertex_runs.size() || o.vertex_runs[v] > o.vertex_runs[v + 1U] ||
      o.vertex_runs[v + 1U] > o.run_tiles.size())
    throw std::invalid_argument{"selected vertex has invalid tile-run offsets"};
  for (std::uint64_t run = o.vertex_runs[v]; run < o.vertex_runs[v + 1U]; ++run) {
    const std::uint32_t remote = o.run_tiles[run].value();
    if (remote >= d.tile_lane_masks.size())
      throw std::invalid_argument{"selected run references an invalid remote tile"};
    const EdgeOffset e0 = o.run_edges[run], e1 = o.run_edges[run + 1U];
    if (e0 >= e1 || e1 > edge_count) throw std::invalid_argument{"selected run has invalid edge offsets"};
    const LaneMask remote_lanes = d.tile_lane_masks[remote];
    const LaneMask mask = lanes & remote_lanes;
    const std::uint64_t edges = e1 - e0;
    c.visited = add(c.visited, 1U);
    if (remote_lanes) c.union_edges = add(c.union_edges, edges);
    if (mask) {
      c.active = add(c.active, 1U);
      c.lane_edge_pairs = add(c.lane_edge_pairs, mul(edges, static_cast<std::uint64_t>(std::popcount(mask))));
      for (std::uint32_t lane = 0; lane < maximum_batch_lanes; ++lane)
        if ((mask >> lane) & 1U) c.edges_by_lane[lane] = add(c.edges_by_lane[lane], edges);
    }
    record(o, run, mask, d);
  }
}

std::uint64_t vector_bytes(std::size_t n, std::size_t width) { return mul(n, width); }

// Device-shaped metadata of a description: its lane and union arrays, the
// three result masks and one controller and status record.
std::uint64_t metadata_bytes(const BatchDeviceDescription& d) {
  std::uint64_t b = 0;
  b = add(b, vector_bytes(d.query_ids_by_lane.size(), 4));
  b = add(b, vector_bytes(d.expansion_generations_by_lane.size(), 4));
  b = add(b, vector_bytes(d.source_offsets.size(), 4));
  b = add(b, vector_bytes(d.sources.size(), 4));
  b = add(b, vector_bytes(d.target_offsets.size(), 4));
  b = add(b, vector_bytes(d.targets.size(), 4));
  b = add(b, vector_bytes(d.selected_vertex_counts_by_lane.size(), 8));
  b = add(b, vector_bytes(d.selected_edge_estimates_by_lane.size(), 8));
  b = add(b, vector_bytes(d.union_tiles.size(), 4));
  b = add(b, vector_bytes(d.tile_lane_masks.size(), sizeof(LaneMask)));
  b = add(b, vector_bytes(d.selected_vertex_ranges.size(), sizeof(BatchVertexRange)));
  b = add(b, 3U * sizeof(LaneMask));
  b = add(b, sizeof(DeviceController));
  b = add(b, sizeof(DeviceRunStatus));
  return b;
}

// Doubles from the current capacity (or starts at the request) until the
// request fits; false when it already fits.
bool grow(std::uint64_t& capacity, std::uint64_t request) {
  if (request <= capacity) return false;
  std::uint64_t next = capacity ? capacity : request;
  while (next < request) next = next > kMax / 2U ? request : next * 2U;
  capacity = next;
  return true;
}

const char* storage_name(BatchVertexStorageStrategy s) {
  return s == BatchVertexStorageStrategy::full_graph_vertex_major ? "full_graph_vertex_major"
                                                                  : "compact_union_tiles";
}

const char* runs_name(BatchRunRepresentation r) {
  switch (r) {
    case BatchRunRepresentation::retained_per_run_masks: return "retained_per_run_masks";
    case BatchRunRepresentation::compact_nonzero_descriptors: return "compact_nonzero_descriptors";
    case BatchRunRepresentation::device_materialized_run_masks: return "device_materialized_run_masks";
  }
  return "invalid";
}

}  // namespace

// ---------------------------------------------------------------------------
// Run images and the compact mapping.
// ---------------------------------------------------------------------------
void prepare_batch_device_description(const WeightedGraph& graph, const TileRunLayout64& tile_runs,
                                      std::span<const RouteQuery> queries,
                                      std::span<const BatchQueryFeatures> features,
                                      const BatchPlanEntry& batch,
                                      BatchRunRepresentation run_representation,
                                      BatchDeviceDescription& out) {
  if (!known(run_representation)) throw std::invalid_argument{"unknown batch run representation"};
  if (out.run_representation_initialized && out.run_representation != run_representation)
    throw std::invalid_argument{"a reused batch description cannot switch run representations"};
  const std::size_t V = graph.vertex_count();
  if (tile_runs.csr_row_run_offsets.size() != V + 1U || tile_runs.csc_column_run_offsets.size() != V + 1U ||
      tile_runs.csr_run_edge_offsets.size() != tile_runs.csr_run_destination_tiles.size() + 1U ||
      tile_runs.csc_run_edge_offsets.size() != tile_runs.csc_run_source_tiles.size() + 1U)
    throw std::invalid_argument{"batch preparation needs the graph's tile-run layout"};
  (void)narrow32(tile_runs.csr_run_destination_tiles.size(), "CSR run count exceeds 32 bits");
  (void)narrow32(tile_runs.csc_run_source_tiles.size(), "CSC run count exceeds 32 bits");

  // The lane fields are rebuilt; the reusable images and their ledgers are
  // carried over.
  std::vector<LaneMask> csr_masks = std::move(out.csr_run_lane_masks);
  std::vector<LaneMask> csc_masks = std::move(out.csc_run_lane_masks);
  std::vector<std::uint32_t> csr_touched = std::move(out.touched_csr_runs);
  std::vector<std::uint32_t> csc_touched = std::move(out.touched_csc_runs);
  std::vector<std::uint32_t> biases = std::move(out.compact_vertex_biases_by_tile);
  std::vector<std::uint32_t> mapped = std::move(out.touched_compact_tiles);
  const bool initialized = out.run_representation_initialized;
  prepare_batch_device_description(graph, queries, features, batch, out);
  out.csr_run_lane_masks = std::move(csr_masks);
  out.csc_run_lane_masks = std::move(csc_masks);
  out.touched_csr_runs = std::move(csr_touched);
  out.touched_csc_runs = std::move(csc_touched);
  out.compact_vertex_biases_by_tile = std::move(biases);
  out.touched_compact_tiles = std::move(mapped);
  out.compact_vertex_mapping_valid = false;
  out.compact_mapping_report = {};

  BatchRunPreparationReport report{};
  clear_ledger(out.csr_run_lane_masks, out.touched_csr_runs, report.retained_entries_cleared,
               report.csr_retained_entries_cleared);
  clear_ledger(out.csc_run_lane_masks, out.touched_csc_runs, report.retained_entries_cleared,
               report.csc_retained_entries_cleared);
  out.run_representation = run_representation;
  out.run_report = report;
  out.csr_run_descriptors.clear();
  out.csc_run_descriptors.clear();
  out.csr_descriptor_offsets_by_union_vertex.clear();
  out.csc_descriptor_offsets_by_union_vertex.clear();
  const std::size_t R_csr = tile_runs.csr_run_destination_tiles.size();
  const std::size_t R_csc = tile_runs.csc_run_source_tiles.size();
  if (run_representation == BatchRunRepresentation::retained_per_run_masks) {
    BatchRunPreparationReport& r = out.run_report;
    if (!initialized || out.csr_run_lane_masks.size() != R_csr) {
      out.csr_run_lane_masks.assign(R_csr, 0U);
      r.csr_retained_entries_initialized = R_csr;
    }
    if (!initialized || out.csc_run_lane_masks.size() != R_csc) {
      out.csc_run_lane_masks.assign(R_csc, 0U);
      r.csc_retained_entries_initialized = R_csc;
    }
    r.retained_entries_initialized = add(r.csr_retained_entries_initialized, r.csc_retained_entries_initialized);
  } else {
    out.csr_run_lane_masks.clear();
    out.csc_run_lane_masks.clear();
    out.touched_csr_runs.clear();
    out.touched_csc_runs.clear();
    if (run_representation == BatchRunRepresentation::compact_nonzero_descriptors) {
      if (batch.union_vertex_count >= kMax) throw std::overflow_error{"union vertex count overflow"};
      out.csr_descriptor_offsets_by_union_vertex.reserve(batch.union_vertex_count + 1U);
      out.csc_descriptor_offsets_by_union_vertex.reserve(batch.union_vertex_count + 1U);
      out.csr_descriptor_offsets_by_union_vertex.push_back(0U);
      out.csc_descriptor_offsets_by_union_vertex.push_back(0U);
    }
  }

  if (run_representation != BatchRunRepresentation::device_materialized_run_masks) {
    const Orientation csr{true, tile_runs.csr_row_run_offsets, tile_runs.csr_run_edge_offsets,
                          tile_runs.csr_run_destination_tiles};
    const Orientation csc{false, tile_runs.csc_column_run_offsets, tile_runs.csc_run_edge_offsets,
                          tile_runs.csc_run_source_tiles};
    Coverage c_csr, c_csc;
    const bool descriptors = run_representation == BatchRunRepresentation::compact_nonzero_descriptors;
    for (const BatchVertexRange& range : out.selected_vertex_ranges) {
      if (range.begin > range.end || range.end > V)
        throw std::invalid_argument{"selected vertex range is outside the graph"};
      for (std::uint32_t v = range.begin; v < range.end; ++v) {
        cover_vertex(csr, graph.edge_count(), v, range.lane_mask, out, c_csr);
        if (descriptors)
          out.csr_descriptor_offsets_by_union_vertex.push_back(
              narrow32(out.csr_run_descriptors.size(), "CSR descriptor position exceeds 32 bits"));
        cover_vertex(csc, graph.edge_count(), v, range.lane_mask, out, c_csc);
        if (descriptors)
          out.csc_descriptor_offsets_by_union_vertex.push_back(
              narrow32(out.csc_run_descriptors.size(), "CSC descriptor position exceeds 32 bits"));
      }
    }
    BatchRunPreparationReport& r = out.run_report;
    r.csr_runs_visited = c_csr.visited;
    r.csc_runs_visited = c_csc.visited;
    r.active_csr_runs = c_csr.active;
    r.active_csc_runs = c_csc.active;
    r.csr_lane_edge_pairs = c_csr.lane_edge_pairs;
    r.csc_lane_edge_pairs = c_csc.lane_edge_pairs;
    if (c_csr.union_edges != batch.union_edge_estimate || c_csc.union_edges != batch.union_edge_estimate)
      throw std::invalid_argument{"the plan's union edge estimate disagrees with the admitted runs"};
    for (std::uint32_t lane = 0; lane < out.lane_width; ++lane)
      if (c_csr.edges_by_lane[lane] != out.selected_edge_estimates_by_lane[lane] ||
          c_csc.edges_by_lane[lane] != out.selected_edge_estimates_by_lane[lane])
        throw std::invalid_argument{"a lane's edge estimate disagrees with the admitted runs"};
  }
  out.run_representation_initialized = true;
}

void prepare_compact_vertex_mapping(const WeightedGraph& graph, const BatchPlanEntry& batch,
                                    BatchDeviceDescription& out) {
  out.compact_vertex_mapping_valid = false;
  out.compact_mapping_report = {};
  const auto offsets = graph.tile_vertex_offsets();
  const std::size_t T = graph.tile_coordinates().size();
  if (!graph.has_spatial_ordering() || T == 0U || offsets.size() != T + 1U ||
      out.lane_width != batch.lane_width || batch.union_tiles.empty() ||
      out.union_tiles.size() != batch.union_tiles.size() ||
      out.selected_vertex_ranges.size() != batch.union_tiles.size())
    throw std::invalid_argument{"compact vertex mapping needs the matching prepared batch"};
  std::uint64_t packed = 0;
  for (std::size_t i = 0; i < batch.union_tiles.size(); ++i) {
    const std::uint32_t t = batch.union_tiles[i].value();
    if (t >= T || (i && !(batch.union_tiles[i - 1] < batch.union_tiles[i])) || out.union_tiles[i] != t ||
        out.selected_vertex_ranges[i].begin != offsets[t] || out.selected_vertex_ranges[i].end != offsets[t + 1U])
      throw std::invalid_argument{"compact vertex mapping disagrees with the union tiles"};
    packed = add(packed, offsets[t + 1U] - offsets[t]);
  }
  if (packed != batch.union_vertex_count || packed > std::numeric_limits<std::uint32_t>::max())
    throw std::invalid_argument{"compact vertex mapping disagrees with the union vertex count"};
  BatchCompactMappingReport& r = out.compact_mapping_report;
  if (out.compact_vertex_biases_by_tile.size() != T) {
    out.compact_vertex_biases_by_tile.assign(T, 0U);
    out.touched_compact_tiles.clear();
    r.entries_initialized = T;
  } else {
    for (std::size_t i = 0; i < out.touched_compact_tiles.size(); ++i) {
      const std::uint32_t t = out.touched_compact_tiles[i];
      if (t >= T || (i && t <= out.touched_compact_tiles[i - 1]))
        throw std::invalid_argument{"reused compact mapping ledger is inconsistent"};
      out.compact_vertex_biases_by_tile[t] = 0U;
      r.entries_cleared = add(r.entries_cleared, 1U);
    }
    out.touched_compact_tiles.clear();
  }
  packed = 0;
  for (const TileId tile : batch.union_tiles) {
    const std::uint32_t t = tile.value();
    out.compact_vertex_biases_by_tile[t] = static_cast<std::uint32_t>(offsets[t] - packed);
    out.touched_compact_tiles.push_back(t);
    r.entries_written = add(r.entries_written, 1U);
    packed += offsets[t + 1U] - offsets[t];
  }
  out.compact_vertex_mapping_valid = true;
}

BatchLayoutValidationResult validate_compact_vertex_mapping(const WeightedGraph& graph,
                                                            const BatchPlanEntry& batch,
                                                            const BatchDeviceDescription& d) noexcept {
  const BatchLayoutValidationResult bad{BatchLayoutValidationErrorCode::invalid_compact_vertex_mapping,
                                        BatchLayoutValidationResult::no_position};
  const auto offsets = graph.tile_vertex_offsets();
  const std::size_t T = graph.tile_coordinates().size();
  if (!d.compact_vertex_mapping_valid || d.compact_vertex_biases_by_tile.size() != T ||
      offsets.size() != T + 1U || d.touched_compact_tiles.size() != batch.union_tiles.size())
    return bad;
  std::uint64_t packed = 0;
  std::size_t u = 0;
  for (std::size_t t = 0; t < T; ++t) {
    const bool in_union = u < batch.union_tiles.size() && batch.union_tiles[u].value() == t;
    if (!in_union) {
      if (d.compact_vertex_biases_by_tile[t] != 0U) return bad;
      continue;
    }
    if (d.touched_compact_tiles[u] != t || d.compact_vertex_biases_by_tile[t] != offsets[t] - packed) return bad;
    packed += offsets[t + 1U] - offsets[t];
    ++u;
  }
  if (u != batch.union_tiles.size() || packed != batch.union_vertex_count) return bad;
  return {};
}

// ---------------------------------------------------------------------------
// Byte model.
// ---------------------------------------------------------------------------
std::uint64_t estimate_vertex_major_distance_bytes(std::uint64_t storage_vertex_count,
                                                   std::uint32_t lane_width,
                                                   std::uint32_t distance_slot_count) {
  if (!supported_width(lane_width) || (distance_slot_count != 1U && distance_slot_count != 2U))
    throw std::invalid_argument{"distance estimate needs width 1/8/16/32 and one or two slots"};
  return mul(mul(storage_vertex_count, lane_width), mul(distance_slot_count, sizeof(float)));
}

BatchWorkspaceEstimate estimate_batch_workspace(const WeightedGraph& graph, const TileRunLayout64& tile_runs,
                                                const BatchPlanEntry& batch, const BatchDeviceDescription& d,
                                                BatchVertexStorageStrategy storage,
                                                BatchRunRepresentation runs, std::uint32_t slots,
                                                const BatchWorkspaceBudget& budget,
                                                const BatchWorkspaceStrategyTiming& timing) {
  if (!known(storage)) throw std::invalid_argument{"invalid batch vertex-storage strategy"};
  if (!known(runs)) throw std::invalid_argument{"invalid batch run representation"};
  if (slots != 1U && slots != 2U) throw std::invalid_argument{"distance slot count must be one or two"};
  const bool full = storage == BatchVertexStorageStrategy::full_graph_vertex_major;
  const bool device_runs = runs == BatchRunRepresentation::device_materialized_run_masks;
  if (device_runs && (!full || slots != 2U))
    throw std::invalid_argument{"device-materialised runs need full-graph two-slot storage"};
  if (batch.lane_width != d.lane_width || d.run_representation != runs || !supported_width(batch.lane_width))
    throw std::invalid_argument{"the workspace request does not match its prepared description"};
  const std::size_t R_csr = tile_runs.csr_run_destination_tiles.size();
  const std::size_t R_csc = tile_runs.csc_run_source_tiles.size();
  const std::size_t T = graph.tile_coordinates().size();
  const BatchRunPreparationReport& rr = d.run_report;
  if (rr.active_csr_runs > R_csr || rr.active_csc_runs > R_csc)
    throw std::invalid_argument{"batch run counts exceed the tile-run layout"};
  if (full && timing.mapping_build_nanoseconds != 0U)
    throw std::invalid_argument{"full-graph rows record no compact mapping work"};
  if (!full && (!d.compact_vertex_mapping_valid || d.compact_vertex_biases_by_tile.size() != T ||
                d.touched_compact_tiles.size() != batch.union_tiles.size() ||
                d.compact_mapping_report.entries_written != batch.union_tiles.size() ||
                d.compact_mapping_report.entries_initialized > T || d.compact_mapping_report.entries_cleared > T))
    throw std::invalid_argument{"compact storage needs a prepared compact vertex mapping"};
  if (runs == BatchRunRepresentation::compact_nonzero_descriptors &&
      (batch.union_vertex_count >= kMax ||
       d.csr_descriptor_offsets_by_union_vertex.size() != batch.union_vertex_count + 1U ||
       d.csc_descriptor_offsets_by_union_vertex.size() != batch.union_vertex_count + 1U ||
       d.csr_descriptor_offsets_by_union_vertex.back() != d.csr_run_descriptors.size() ||
       d.csc_descriptor_offsets_by_union_vertex.back() != d.csc_run_descriptors.size()))
    throw std::invalid_argument{"descriptor storage needs traversable per-vertex offsets"};
  if (device_runs &&
      (d.tile_lane_masks.size() != T || d.selected_vertex_ranges.empty() || !d.csr_run_lane_masks.empty() ||
       !d.csc_run_lane_masks.empty() || !d.touched_csr_runs.empty() || !d.touched_csc_runs.empty() ||
       !d.csr_run_descriptors.empty() || !d.csc_run_descriptors.empty() ||
       !d.csr_descriptor_offsets_by_union_vertex.empty() || !d.csc_descriptor_offsets_by_union_vertex.empty() ||
       rr != BatchRunPreparationReport{} || timing.run_build_nanoseconds != 0U))
    throw std::invalid_argument{"device-materialised runs need an empty host run image"};
  const bool counters_ok =
      d.run_representation_initialized && rr.csr_retained_entries_initialized <= R_csr &&
      rr.csc_retained_entries_initialized <= R_csc && rr.csr_retained_entries_cleared <= R_csr &&
      rr.csc_retained_entries_cleared <= R_csc && rr.csr_retained_entries_written <= R_csr &&
      rr.csc_retained_entries_written <= R_csc && rr.csr_descriptor_entries_written <= R_csr &&
      rr.csc_descriptor_entries_written <= R_csc &&
      !(runs == BatchRunRepresentation::retained_per_run_masks &&
        (rr.csr_descriptor_entries_written || rr.csc_descriptor_entries_written)) &&
      !(runs == BatchRunRepresentation::compact_nonzero_descriptors &&
        (rr.retained_entries_initialized || rr.csr_retained_entries_written || rr.csc_retained_entries_written)) &&
      !(device_runs && rr != BatchRunPreparationReport{});
  if (!counters_ok) throw std::invalid_argument{"batch run preparation counters are inconsistent"};

  BatchWorkspaceEstimate e;
  e.vertex_storage = storage;
  e.run_representation = runs;
  e.lane_width = batch.lane_width;
  e.distance_slot_count = slots;
  e.mapping_build_nanoseconds = timing.mapping_build_nanoseconds;
  e.run_build_nanoseconds = timing.run_build_nanoseconds;
  e.union_vertex_count = batch.union_vertex_count;
  e.selected_lane_vertex_count = batch.selected_lane_vertex_count;
  e.storage_vertex_count = full ? graph.vertex_count() : batch.union_vertex_count;
  e.allocated_lane_vertex_count = mul(e.storage_vertex_count, e.lane_width);
  if (e.selected_lane_vertex_count > e.allocated_lane_vertex_count)
    throw std::invalid_argument{"selected lane vertices exceed the workspace storage"};
  e.wasted_lane_vertex_count = e.allocated_lane_vertex_count - e.selected_lane_vertex_count;
  e.distance_bytes = estimate_vertex_major_distance_bytes(e.storage_vertex_count, e.lane_width, slots);
  // One u32 bias per tile maps a global vertex to its packed label.
  e.tile_mapping_bytes = full ? 0U : mul(T, sizeof(std::uint32_t));

  std::uint64_t device_csc_writes = 0;
  e.active_csr_runs = rr.active_csr_runs;
  e.active_csc_runs = rr.active_csc_runs;
  if (device_runs) {
    // An offline count over the selected CSC columns: every selected run is
    // written (zero masks included); the nonzero ones are admitted.
    e.active_csr_runs = 0;
    e.active_csc_runs = 0;
    for (const BatchVertexRange& range : d.selected_vertex_ranges) {
      if (range.begin > range.end || range.end > graph.vertex_count())
        throw std::invalid_argument{"device-materialised selected vertex range is invalid"};
      for (std::uint32_t v = range.begin; v < range.end; ++v) {
        const EdgeOffset r0 = tile_runs.csc_column_run_offsets[v], r1 = tile_runs.csc_column_run_offsets[v + 1U];
        device_csc_writes = add(device_csc_writes, r1 - r0);
        for (EdgeOffset run = r0; run < r1; ++run)
          if (range.lane_mask & d.tile_lane_masks[tile_runs.csc_run_source_tiles[run].value()])
            e.active_csc_runs = add(e.active_csc_runs, 1U);
      }
    }
  }
  e.zero_csr_runs = device_runs ? 0U : R_csr - e.active_csr_runs;
  e.zero_csc_runs = R_csc - e.active_csc_runs;
  switch (runs) {
    case BatchRunRepresentation::retained_per_run_masks:
      e.run_storage_bytes = mul(std::max<std::uint64_t>(R_csr, R_csc), sizeof(LaneMask));
      break;
    case BatchRunRepresentation::compact_nonzero_descriptors:
      e.run_storage_bytes = mul(std::max(e.active_csr_runs, e.active_csc_runs), sizeof(RunLaneMaskDescriptor));
      e.descriptor_offset_bytes = mul(add(e.union_vertex_count, 1U), sizeof(std::uint32_t));
      break;
    case BatchRunRepresentation::device_materialized_run_masks:
      e.run_storage_bytes = mul(R_csc, sizeof(LaneMask));
      break;
  }
  e.batch_metadata_bytes = metadata_bytes(d);
  for (std::uint64_t part : {e.distance_bytes, e.tile_mapping_bytes, e.run_storage_bytes,
                             e.descriptor_offset_bytes, e.batch_metadata_bytes})
    e.total_workspace_bytes = add(e.total_workspace_bytes, part);

  e.distance_reset_bytes = mul(mul(e.selected_lane_vertex_count, slots), sizeof(float));
  if (!full) {
    const BatchCompactMappingReport& m = d.compact_mapping_report;
    e.tile_mapping_write_bytes =
        mul(add(m.entries_initialized, add(m.entries_cleared, m.entries_written)), sizeof(std::uint32_t));
  }
  switch (runs) {
    case BatchRunRepresentation::retained_per_run_masks: {
      const std::uint64_t csr = add(rr.csr_retained_entries_initialized,
                                    add(rr.csr_retained_entries_cleared, rr.csr_retained_entries_written));
      const std::uint64_t csc = add(rr.csc_retained_entries_initialized,
                                    add(rr.csc_retained_entries_cleared, rr.csc_retained_entries_written));
      e.run_preparation_write_bytes = mul(std::max(csr, csc), sizeof(LaneMask));
      break;
    }
    case BatchRunRepresentation::compact_nonzero_descriptors: {
      // The host builds both orientations; the device holds one at a time.
      const std::uint64_t csr = add(mul(rr.csr_descriptor_entries_written, sizeof(RunLaneMaskDescriptor)),
                                    vector_bytes(d.csr_descriptor_offsets_by_union_vertex.size(), 4));
      const std::uint64_t csc = add(mul(rr.csc_descriptor_entries_written, sizeof(RunLaneMaskDescriptor)),
                                    vector_bytes(d.csc_descriptor_offsets_by_union_vertex.size(), 4));
      e.run_preparation_write_bytes = std::max(csr, csc);
      break;
    }
    case BatchRunRepresentation::device_materialized_run_masks:
      e.run_preparation_write_bytes = mul(device_csc_writes, sizeof(LaneMask));
      break;
  }
  e.total_preparation_write_bytes =
      add(e.distance_reset_bytes, add(e.tile_mapping_write_bytes, e.run_preparation_write_bytes));
  e.reusable_allocation = true;
  const std::uint64_t committed = add(budget.resident_graph_bytes, budget.explicit_reserve_bytes);
  if (committed <= budget.device_capacity_bytes && e.total_workspace_bytes)
    e.maximum_concurrent_workspaces = (budget.device_capacity_bytes - committed) / e.total_workspace_bytes;
  return e;
}

std::vector<BatchWorkspaceEstimate> compare_batch_workspace_strategies(
    const WeightedGraph& graph, const TileRunLayout64& tile_runs, const BatchPlanEntry& batch,
    const BatchDeviceDescription& retained, const BatchDeviceDescription& descriptors, std::uint32_t slots,
    const BatchWorkspaceBudget& budget, const BatchWorkspaceBuildMeasurements& m) {
  if (retained.run_representation != BatchRunRepresentation::retained_per_run_masks ||
      descriptors.run_representation != BatchRunRepresentation::compact_nonzero_descriptors)
    throw std::invalid_argument{"the strategy comparison needs both run representations"};
  std::vector<BatchWorkspaceEstimate> rows;
  rows.reserve(4);
  for (const BatchVertexStorageStrategy storage :
       {BatchVertexStorageStrategy::full_graph_vertex_major, BatchVertexStorageStrategy::compact_union_tiles}) {
    const std::uint64_t mapping =
        storage == BatchVertexStorageStrategy::compact_union_tiles ? m.compact_mapping_build_nanoseconds : 0U;
    rows.push_back(estimate_batch_workspace(graph, tile_runs, batch, retained, storage,
                                            BatchRunRepresentation::retained_per_run_masks, slots, budget,
                                            {mapping, m.retained_run_build_nanoseconds}));
    rows.push_back(estimate_batch_workspace(graph, tile_runs, batch, descriptors, storage,
                                            BatchRunRepresentation::compact_nonzero_descriptors, slots, budget,
                                            {mapping, m.descriptor_run_build_nanoseconds}));
  }
  return rows;
}

std::uint64_t BatchWorkspaceCapacity::total_bytes() const {
  std::uint64_t t = 0;
  for (std::uint64_t part : {distance_bytes, tile_mapping_bytes, run_storage_bytes, descriptor_offset_bytes,
                             batch_metadata_bytes})
    t = add(t, part);
  return t;
}

BatchWorkspaceReservationResult ReusableBatchWorkspaceReservation::reserve(const BatchWorkspaceEstimate& e) {
  if (!e.reusable_allocation || e.total_workspace_bytes == 0U)
    throw std::invalid_argument{"a reservation needs a nonempty estimate"};
  std::uint64_t sum = 0;
  for (std::uint64_t part : {e.distance_bytes, e.tile_mapping_bytes, e.run_storage_bytes,
                             e.descriptor_offset_bytes, e.batch_metadata_bytes})
    sum = add(sum, part);
  if (sum != e.total_workspace_bytes) throw std::invalid_argument{"estimate components do not sum to its total"};
  bool grew = false;
  grew |= grow(capacity_.distance_bytes, e.distance_bytes);
  grew |= grow(capacity_.tile_mapping_bytes, e.tile_mapping_bytes);
  grew |= grow(capacity_.run_storage_bytes, e.run_storage_bytes);
  grew |= grow(capacity_.descriptor_offset_bytes, e.descriptor_offset_bytes);
  grew |= grow(capacity_.batch_metadata_bytes, e.batch_metadata_bytes);
  if (grew) ++growth_events_;
  if (generation_ == kMax) throw std::overflow_error{"workspace generation overflow"};
  return BatchWorkspaceReservationResult{grew, ++generation_};
}

// ---------------------------------------------------------------------------
// Decisions.
// ---------------------------------------------------------------------------
namespace {

bool checked_mul(std::uint64_t a, std::uint64_t b, std::uint64_t& out) noexcept {
  if (b != 0U && a > kMax / b) return false;
  out = a * b;
  return true;
}

bool checked_add(std::uint64_t a, std::uint64_t b, std::uint64_t& out) noexcept {
  if (a > kMax - b) return false;
  out = a + b;
  return true;
}

// One row's own formulas (distance, reset, totals, run storage, mapping).
bool row_consistent(const BatchWorkspaceDecision& dec, const BatchWorkspaceEstimate& e,
                    std::uint64_t committed) noexcept {
  if (!known(e.vertex_storage) ||
      (e.run_representation != BatchRunRepresentation::retained_per_run_masks &&
       e.run_representation != BatchRunRepresentation::compact_nonzero_descriptors) ||
      e.lane_width != dec.lane_width || e.distance_slot_count != dec.distance_slot_count ||
      e.union_vertex_count == 0U || e.selected_lane_vertex_count == 0U || e.total_workspace_bytes == 0U ||
      !e.reusable_allocation)
    return false;
  const bool full = e.vertex_storage == BatchVertexStorageStrategy::full_graph_vertex_major;
  std::uint64_t sum = 0, lanes = 0, dist = 0, reset = 0, prep = 0;
  for (std::uint64_t part : {e.distance_bytes, e.tile_mapping_bytes, e.run_storage_bytes,
                             e.descriptor_offset_bytes, e.batch_metadata_bytes})
    if (!checked_add(sum, part, sum)) return false;
  if (!checked_mul(e.storage_vertex_count, e.lane_width, lanes) ||
      !checked_mul(lanes, e.distance_slot_count, dist) || !checked_mul(dist, 4U, dist) ||
      !checked_mul(e.selected_lane_vertex_count, e.distance_slot_count, reset) || !checked_mul(reset, 4U, reset) ||
      !checked_add(e.distance_reset_bytes, e.tile_mapping_write_bytes, prep) ||
      !checked_add(prep, e.run_preparation_write_bytes, prep))
    return false;
  if (sum != e.total_workspace_bytes || lanes != e.allocated_lane_vertex_count || dist != e.distance_bytes ||
      reset != e.distance_reset_bytes || prep != e.total_preparation_write_bytes ||
      e.selected_lane_vertex_count > e.allocated_lane_vertex_count ||
      e.wasted_lane_vertex_count != e.allocated_lane_vertex_count - e.selected_lane_vertex_count)
    return false;
  std::uint64_t csr_runs = 0, csc_runs = 0;
  if (!checked_add(e.active_csr_runs, e.zero_csr_runs, csr_runs) ||
      !checked_add(e.active_csc_runs, e.zero_csc_runs, csc_runs))
    return false;
  if (full) {
    if (e.storage_vertex_count != dec.graph_vertex_count || e.tile_mapping_bytes || e.tile_mapping_write_bytes ||
        e.mapping_build_nanoseconds)
      return false;
  } else {
    std::uint64_t table = 0;
    if (!checked_mul(dec.graph_tile_count, 4U, table)) return false;
    const std::uint64_t entries = e.tile_mapping_bytes / 4U, writes = e.tile_mapping_write_bytes / 4U;
    // A build initialises or clears at most one table image and writes at
    // most one more.
    if (e.storage_vertex_count != e.union_vertex_count || e.tile_mapping_bytes != table ||
        e.tile_mapping_write_bytes == 0U || e.tile_mapping_write_bytes % 4U || entries < (writes + 1U) / 2U)
      return false;
  }
  if (e.run_representation == BatchRunRepresentation::retained_per_run_masks) {
    std::uint64_t storage = 0, minimum = 0;
    if (!checked_mul(std::max(csr_runs, csc_runs), 4U, storage) ||
        !checked_mul(std::max(e.active_csr_runs, e.active_csc_runs), 4U, minimum))
      return false;
    if (e.run_storage_bytes != storage || e.descriptor_offset_bytes || e.run_preparation_write_bytes < minimum ||
        e.run_preparation_write_bytes % 4U || e.run_preparation_write_bytes / 2U > e.run_storage_bytes)
      return false;
  } else {
    std::uint64_t storage = 0, offsets = 0, writes = 0;
    if (!checked_mul(std::max(e.active_csr_runs, e.active_csc_runs), 8U, storage) ||
        !checked_add(e.union_vertex_count, 1U, offsets) || !checked_mul(offsets, 4U, offsets) ||
        !checked_add(storage, offsets, writes))
      return false;
    if (e.run_storage_bytes != storage || e.descriptor_offset_bytes != offsets ||
        e.run_preparation_write_bytes != writes)
      return false;
  }
  const std::uint64_t concurrent =
      committed <= dec.budget.device_capacity_bytes
          ? (dec.budget.device_capacity_bytes - committed) / e.total_workspace_bytes
          : 0U;
  return e.maximum_concurrent_workspaces == concurrent;
}

}  // namespace

bool validate_batch_workspace_decision(const BatchWorkspaceDecision& dec) noexcept {
  if ((dec.measurement_scope != BatchMeasurementScope::bounded_synthetic &&
       dec.measurement_scope != BatchMeasurementScope::real_query_corpus) ||
      !known(dec.selected_vertex_storage) ||
      (dec.selected_run_representation != BatchRunRepresentation::retained_per_run_masks &&
       dec.selected_run_representation != BatchRunRepresentation::compact_nonzero_descriptors) ||
      !supported_width(dec.lane_width) || (dec.distance_slot_count != 1U && dec.distance_slot_count != 2U) ||
      dec.plan_fingerprint == 0U || dec.graph_vertex_count == 0U || dec.graph_tile_count == 0U ||
      dec.graph_vertex_count > std::numeric_limits<std::uint32_t>::max() ||
      dec.graph_tile_count > std::numeric_limits<std::uint32_t>::max() || dec.compared_strategies.size() != 4U ||
      dec.quantitative_reason.empty() || dec.quantitative_reason.find_first_of("\r\n\t") != std::string::npos)
    return false;
  std::uint64_t committed = 0;
  if (!checked_add(dec.budget.resident_graph_bytes, dec.budget.explicit_reserve_bytes, committed)) return false;
  // rows[storage * 2 + runs]: each combination exactly once
  std::array<const BatchWorkspaceEstimate*, 4> rows{};
  for (const BatchWorkspaceEstimate& e : dec.compared_strategies) {
    if (!row_consistent(dec, e, committed)) return false;
    const std::size_t i =
        (e.vertex_storage == BatchVertexStorageStrategy::compact_union_tiles ? 2U : 0U) +
        (e.run_representation == BatchRunRepresentation::compact_nonzero_descriptors ? 1U : 0U);
    if (rows[i]) return false;
    rows[i] = &e;
  }
  const std::size_t selected =
      (dec.selected_vertex_storage == BatchVertexStorageStrategy::compact_union_tiles ? 2U : 0U) +
      (dec.selected_run_representation == BatchRunRepresentation::compact_nonzero_descriptors ? 1U : 0U);
  const BatchWorkspaceEstimate& s = *rows[selected];
  if (s.maximum_concurrent_workspaces == 0U || dec.selected_mapping_build_nanoseconds != s.mapping_build_nanoseconds ||
      dec.selected_run_build_nanoseconds != s.run_build_nanoseconds)
    return false;
  // The four rows describe one batch: the same counts everywhere, the same
  // vertex components per storage and the same run components per
  // representation.
  auto same_batch = [](const BatchWorkspaceEstimate& a, const BatchWorkspaceEstimate& b) {
    return a.union_vertex_count == b.union_vertex_count &&
           a.selected_lane_vertex_count == b.selected_lane_vertex_count &&
           a.batch_metadata_bytes == b.batch_metadata_bytes && a.active_csr_runs == b.active_csr_runs &&
           a.active_csc_runs == b.active_csc_runs && a.zero_csr_runs == b.zero_csr_runs &&
           a.zero_csc_runs == b.zero_csc_runs;
  };
  auto same_vertices = [](const BatchWorkspaceEstimate& a, const BatchWorkspaceEstimate& b) {
    return a.storage_vertex_count == b.storage_vertex_count &&
           a.allocated_lane_vertex_count == b.allocated_lane_vertex_count &&
           a.wasted_lane_vertex_count == b.wasted_lane_vertex_count && a.distance_bytes == b.distance_bytes &&
           a.distance_reset_bytes == b.distance_reset_bytes && a.tile_mapping_bytes == b.tile_mapping_bytes &&
           a.tile_mapping_write_bytes == b.tile_mapping_write_bytes &&
           a.mapping_build_nanoseconds == b.mapping_build_nanoseconds;
  };
  auto same_runs = [](const BatchWorkspaceEstimate& a, const BatchWorkspaceEstimate& b) {
    return a.run_storage_bytes == b.run_storage_bytes && a.descriptor_offset_bytes == b.descriptor_offset_bytes &&
           a.run_preparation_write_bytes == b.run_preparation_write_bytes &&
           a.run_build_nanoseconds == b.run_build_nanoseconds;
  };
  const auto& fr = *rows[0];
  const auto& fd = *rows[1];
  const auto& cr = *rows[2];
  const auto& cd = *rows[3];
  return same_batch(fr, fd) && same_batch(fr, cr) && same_batch(fr, cd) && same_vertices(fr, fd) &&
         same_vertices(cr, cd) && same_runs(fr, cr) && same_runs(fd, cd) &&
         fr.storage_vertex_count >= fr.union_vertex_count;
}

std::string serialize_batch_workspace_decision(const BatchWorkspaceDecision& dec) {
  if (!validate_batch_workspace_decision(dec))
    throw std::invalid_argument{"cannot serialize an invalid batch workspace decision"};
  auto line = [](std::string& out, const char* key, const std::string& value) {
    out += key;
    out += '=';
    out += value;
    out += '\n';
  };
  using std::to_string;
  std::string out = "bfnew_phase13_workspace_decision_v1\n";
  line(out, "plan_fingerprint", to_string(dec.plan_fingerprint));
  line(out, "measurement_scope", to_string(static_cast<std::uint32_t>(dec.measurement_scope)));
  line(out, "device_capacity_bytes", to_string(dec.budget.device_capacity_bytes));
  line(out, "resident_graph_bytes", to_string(dec.budget.resident_graph_bytes));
  line(out, "explicit_reserve_bytes", to_string(dec.budget.explicit_reserve_bytes));
  line(out, "graph_vertex_count", to_string(dec.graph_vertex_count));
  line(out, "graph_tile_count", to_string(dec.graph_tile_count));
  line(out, "lane_width", to_string(dec.lane_width));
  line(out, "distance_slot_count", to_string(dec.distance_slot_count));
  line(out, "selected_vertex_storage", storage_name(dec.selected_vertex_storage));
  line(out, "selected_run_representation", runs_name(dec.selected_run_representation));
  line(out, "selected_mapping_build_nanoseconds", to_string(dec.selected_mapping_build_nanoseconds));
  line(out, "selected_run_build_nanoseconds", to_string(dec.selected_run_build_nanoseconds));
  line(out, "run_build_scope", "dual_orientation_host_proof_image");
  line(out, "reason", dec.quantitative_reason);
  out += "vertex_storage\trun_representation\ttotal_bytes\ttotal_preparation_write_bytes\t"
         "wasted_lane_vertices\tactive_csr_runs\tactive_csc_runs\tmapping_build_ns\t"
         "dual_orientation_host_run_build_ns\tmax_concurrent\n";
  for (const BatchWorkspaceEstimate& e : dec.compared_strategies) {
    out += storage_name(e.vertex_storage);
    for (const std::string& field :
         {std::string{runs_name(e.run_representation)}, to_string(e.total_workspace_bytes),
          to_string(e.total_preparation_write_bytes), to_string(e.wasted_lane_vertex_count),
          to_string(e.active_csr_runs), to_string(e.active_csc_runs), to_string(e.mapping_build_nanoseconds),
          to_string(e.run_build_nanoseconds), to_string(e.maximum_concurrent_workspaces)}) {
      out += '\t';
      out += field;
    }
    out += '\n';
  }
  return out;
}

}  // namespace bfnew
