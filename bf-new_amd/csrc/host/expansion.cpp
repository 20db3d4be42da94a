// expansion.cpp -- reachability expansion over batched queries: plan the
// active queries, run each lane batch, and restart every bounding-box miss
// in a larger region (reference behaviour: src/batched_expansion.cpp:134-151
// schedule rules, :153-214 the located tile box, :223-245 per-query state,
// :247-303 comparison fingerprint, :305-444 margin growth and region
// growth, :667-790 the restart decision, :831-1680 options and the pass
// loop, :1825-1920 schedule evidence).
//
// Each query keeps an anchor: the box of located tiles its first region
// covers.  An expansion widens the box by the schedule's x/y margins and
// selects every located tile inside it (plus the spill tile once a selected
// located tile is adjacent to it in the directory).  A query that gets no new
// tile stalls; after the scheduled expansions the terminal policy either
// restarts it once on every tile or classifies it.
#include "bfnew/batched_expansion.hpp"

#include "bfnew/compact_paths.hpp"
#include "bfnew/device_layout.hpp"
#include "bfnew/graph.hpp"

#include <algorithm>
#include <bit>
#include <chrono>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string_view>

namespace bfnew {
namespace {

using Clock = std::chrono::steady_clock;
constexpr std::uint64_t kU64Max = std::numeric_limits<std::uint64_t>::max();
constexpr std::uint32_t kU32Max = std::numeric_limits<std::uint32_t>::max();

std::uint64_t since(Clock::time_point t0) {
  const auto n = std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
  return n > 0 ? static_cast<std::uint64_t>(n) : 0U;
}

// Aggregates must not wrap: a wrapped counter would silently misreport.
void add(std::uint64_t& total, std::uint64_t value) {
  if (value > kU64Max - total) throw std::overflow_error{"batched expansion counter overflow"};
  total += value;
}

std::uint64_t lanes_in(LaneMask mask) { return static_cast<std::uint64_t>(std::popcount(mask)); }

// 64-bit FNV-1a over little-endian 8-byte words.
struct Fnv {
  std::uint64_t h{14695981039346656037ULL};
  void word(std::uint64_t v) {
    for (int i = 0; i < 8; ++i, v >>= 8) {
      h ^= v & 0xFFU;
      h *= 1099511628211ULL;
    }
  }
  template <class T, class F>
  void list(std::span<const T> values, F value_of) {
    word(values.size());
    for (const T& v : values) word(value_of(v));
  }
};

bool schedule_ok(const ExpansionSchedulePolicy& s) noexcept {
  switch (s.kind) {
    case ExpansionScheduleKind::one_geometric_ring:
    case ExpansionScheduleKind::doubling_xy_margins:
      return s.fixed_ring_size == 0U && s.hybrid_small_expansion_count == 0U;
    case ExpansionScheduleKind::fixed_larger_ring:
      return s.fixed_ring_size >= 2U && s.hybrid_small_expansion_count == 0U;
    case ExpansionScheduleKind::hybrid_small_then_doubling:
      return s.fixed_ring_size == 0U && s.hybrid_small_expansion_count != 0U;
    default:
      return false;
  }
}

// Inclusive box of located tile coordinates.
struct TileBox {
  bool present{};
  std::int64_t x0{}, y0{}, x1{}, y1{};
};

TileBox located_box(const WeightedGraph& graph, std::span<const TileId> tiles) {
  TileBox b;
  for (const TileId t : tiles) {
    const TileCoordinate c = graph.tile_coordinates()[t.value()];
    if (!c.has_location) continue;
    if (!b.present) {
      b = TileBox{true, c.tile_x, c.tile_y, c.tile_x, c.tile_y};
      continue;
    }
    b.x0 = std::min(b.x0, c.tile_x);
    b.y0 = std::min(b.y0, c.tile_y);
    b.x1 = std::max(b.x1, c.tile_x);
    b.y1 = std::max(b.y1, c.tile_y);
  }
  return b;
}

// value -/+ margin, clamped to the int64 range.
std::int64_t lower_clamped(std::int64_t value, std::uint64_t margin) {
  const __int128 r = static_cast<__int128>(value) - static_cast<__int128>(margin);
  return r < std::numeric_limits<std::int64_t>::min() ? std::numeric_limits<std::int64_t>::min()
                                                      : static_cast<std::int64_t>(r);
}
std::int64_t upper_clamped(std::int64_t value, std::uint64_t margin) {
  const __int128 r = static_cast<__int128>(value) + static_cast<__int128>(margin);
  return r > std::numeric_limits<std::int64_t>::max() ? std::numeric_limits<std::int64_t>::max()
                                                      : static_cast<std::int64_t>(r);
}

std::span<const TileId> neighbours(const TileDirectory& dir, TileId t) {
  const auto off = dir.neighbor_tile_offsets();
  return dir.neighbor_tiles().subspan(off[t.value()], off[t.value() + 1U] - off[t.value()]);
}

void sort_unique(std::vector<TileId>& tiles) {
  std::sort(tiles.begin(), tiles.end());
  tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
}

struct QueryState {
  RouteQuery query;
  TileBox anchor;
  std::uint64_t margin_x{};
  std::uint64_t margin_y{};
  std::uint32_t attempts{};
  std::uint32_t scheduled{};
  std::uint32_t total{};
  bool fallback_used{};
  bool done{};
  ExpansionQueryDisposition disposition{ExpansionQueryDisposition::engine_failure};
  std::uint32_t stop_reason{};
  std::uint32_t error_bits{};
  std::vector<float> distances;
  std::optional<CompactPathPayload> paths;
};

void finish(QueryState& q, ExpansionQueryDisposition d, const DeviceRunStatus& status) {
  q.done = true;
  q.disposition = d;
  q.stop_reason = status.stop_reason;
  q.error_bits = status.error_bits;
}

// Next margins of the schedule; false when they would not fit 64 bits.
bool grow_margins(QueryState& q, const ExpansionSchedulePolicy& s) {
  const auto doubled = [](std::uint64_t m, std::uint64_t& out) {
    if (m == 0U) return out = 1U, true;
    if (m > kU64Max / 2U) return false;
    return out = 2U * m, true;
  };
  const auto plus = [](std::uint64_t m, std::uint64_t step, std::uint64_t& out) {
    if (m > kU64Max - step) return false;
    return out = m + step, true;
  };
  std::uint64_t x = 0, y = 0;
  bool ok = false;
  switch (s.kind) {
    case ExpansionScheduleKind::one_geometric_ring:
      ok = plus(q.margin_x, 1U, x) && plus(q.margin_y, 1U, y);
      break;
    case ExpansionScheduleKind::fixed_larger_ring:
      ok = plus(q.margin_x, s.fixed_ring_size, x) && plus(q.margin_y, s.fixed_ring_size, y);
      break;
    case ExpansionScheduleKind::doubling_xy_margins:
      ok = doubled(q.margin_x, x) && doubled(q.margin_y, y);
      break;
    case ExpansionScheduleKind::hybrid_small_then_doubling:
      ok = q.scheduled < s.hybrid_small_expansion_count
               ? plus(q.margin_x, 1U, x) && plus(q.margin_y, 1U, y)
               : doubled(q.margin_x, x) && doubled(q.margin_y, y);
      break;
    default:
      ok = false;
  }
  if (ok) q.margin_x = x, q.margin_y = y;
  return ok;
}

// The region of the query's next expansion.  A query with no located tile
// (spill only) takes the spill tile's located neighbours first, which then
// become its anchor with zero margins.
std::vector<TileId> grown_region(const WeightedGraph& graph, const TileDirectory& dir, QueryState& q) {
  std::vector<TileId> region = q.query.selected_tiles;
  const TileId spill = dir.spill_tile();
  const auto coords = graph.tile_coordinates();
  const bool has_spill = std::binary_search(region.begin(), region.end(), spill);
  if (!q.anchor.present) {
    if (has_spill) {
      for (const TileId n : neighbours(dir, spill))
        if (coords[n.value()].has_location) region.push_back(n);
      sort_unique(region);
      q.anchor = located_box(graph, region);
      q.margin_x = q.margin_y = 0U;
    }
    return region;
  }
  const std::int64_t x0 = lower_clamped(q.anchor.x0, q.margin_x);
  const std::int64_t y0 = lower_clamped(q.anchor.y0, q.margin_y);
  const std::int64_t x1 = upper_clamped(q.anchor.x1, q.margin_x);
  const std::int64_t y1 = upper_clamped(q.anchor.y1, q.margin_y);
  for (std::size_t t = 0; t < coords.size(); ++t) {
    const TileCoordinate c = coords[t];
    if (c.has_location && c.tile_x >= x0 && c.tile_x <= x1 && c.tile_y >= y0 && c.tile_y <= y1)
      region.push_back(TileId{static_cast<std::uint32_t>(t)});
  }
  sort_unique(region);
  if (!has_spill) {
    for (const TileId t : region) {
      if (!coords[t.value()].has_location) continue;
      const auto n = neighbours(dir, t);
      if (std::binary_search(n.begin(), n.end(), spill)) {
        region.push_back(spill);
        sort_unique(region);
        break;
      }
    }
  }
  return region;
}

enum class Next : std::uint8_t { retry, done };

// The decision for a query whose clean batch missed its targets.
Next after_miss(const WeightedGraph& graph, const TileDirectory& dir, const SelectedRegionIndex& regions,
                std::span<const TileId> every_tile, const BatchedExpansionOptions& options,
                const DeviceRunStatus& status, QueryState& q, BatchedExpansionMetrics& m) {
  const bool whole_graph = std::ranges::equal(q.query.selected_tiles, every_tile);
  if (q.fallback_used || whole_graph) {
    finish(q, ExpansionQueryDisposition::unreachable_in_full_region, status);
    return Next::done;
  }
  // Installs the next region; false when a counter would wrap.
  const auto restart = [&](std::vector<TileId> region, bool fallback) {
    if (q.query.expansion_generation == kU32Max || q.total == kU32Max) return false;
    std::uint64_t repeated = m.repeated_selected_edge_estimate;
    std::uint64_t fallbacks = m.full_region_fallbacks;
    add(repeated, regions.selected_edge_count(q.query.selected_tiles));
    if (fallback) add(fallbacks, 1U);
    m.repeated_selected_edge_estimate = repeated;
    m.full_region_fallbacks = fallbacks;
    ++q.query.expansion_generation;
    ++q.total;
    q.query.selected_tiles = std::move(region);
    q.distances.clear();
    q.paths.reset();
    q.fallback_used = q.fallback_used || fallback;
    return true;
  };
  const bool explicit_failure = options.terminal_policy == ExpansionTerminalPolicy::explicit_failure;
  if (q.scheduled < options.maximum_expansions) {
    if (!grow_margins(q, options.schedule)) {
      finish(q, ExpansionQueryDisposition::identity_or_count_overflow, status);
      return Next::done;
    }
    std::vector<TileId> region = grown_region(graph, dir, q);
    if (region != q.query.selected_tiles) {
      if (q.scheduled == kU32Max || !restart(std::move(region), false)) {
        finish(q, ExpansionQueryDisposition::identity_or_count_overflow, status);
        return Next::done;
      }
      ++q.scheduled;
      add(m.scheduled_expansions, 1U);
      return Next::retry;
    }
    if (explicit_failure) {
      finish(q, ExpansionQueryDisposition::region_stalled, status);
      return Next::done;
    }
  } else if (explicit_failure) {
    finish(q, ExpansionQueryDisposition::expansion_limit, status);
    return Next::done;
  }
  if (!restart(std::vector<TileId>(every_tile.begin(), every_tile.end()), true)) {
    finish(q, ExpansionQueryDisposition::identity_or_count_overflow, status);
    return Next::done;
  }
  return Next::retry;
}

bool clean(const DeviceRunStatus& s) {
  return s.converged == 1U && s.error_bits == device_error::none &&
         s.stop_reason == static_cast<std::uint32_t>(DeviceStopReason::converged);
}

bool work_is_zero(const DeviceWorkStatistics& w) {
  static const DeviceWorkStatistics zero{};
  return std::memcmp(&w, &zero, sizeof w) == 0;
}

void add_work(DeviceWorkStatistics& into, const DeviceWorkStatistics& w) {
  add(into.edges_examined, w.edges_examined);
  add(into.successful_decreases, w.successful_decreases);
  add(into.active_vertices, w.active_vertices);
  add(into.active_lane_rounds, w.active_lane_rounds);
  into.maximum_queue_size = std::max(into.maximum_queue_size, w.maximum_queue_size);
  add(into.host_checks, w.host_checks);
  add(into.host_synchronizations, w.host_synchronizations);
  add(into.controller_copies, w.controller_copies);
  add(into.kernel_dispatches, w.kernel_dispatches);
  add(into.expansion_count, w.expansion_count);
  add(into.atomic_attempts, w.atomic_attempts);
  add(into.successful_atomic_updates, w.successful_atomic_updates);
  add(into.queue_claims, w.queue_claims);
  add(into.duplicate_suppressions, w.duplicate_suppressions);
  add(into.mask_operations, w.mask_operations);
  add(into.overflow_events, w.overflow_events);
  add(into.high_contention_destinations, w.high_contention_destinations);
  add(into.changed_flag_updates, w.changed_flag_updates);
  add(into.full_edge_rounds, w.full_edge_rounds);
  add(into.empty_frontier_rounds, w.empty_frontier_rounds);
  add(into.small_frontier_rounds, w.small_frontier_rounds);
}

void add_transfer(CompactTransferAccounting& into, const CompactTransferAccounting& a) {
  add(into.summary_bytes, a.summary_bytes);
  add(into.vertex_bytes, a.vertex_bytes);
  add(into.distance_label_bytes, a.distance_label_bytes);
  add(into.edge_id_bytes, a.edge_id_bytes);
  add(into.total_bytes, a.total_bytes);
}

bool sums_to(std::uint64_t total, std::initializer_list<std::uint64_t> parts) {
  std::uint64_t s = 0;
  for (const std::uint64_t p : parts) {
    if (p > kU64Max - s) return false;
    s += p;
  }
  return s == total;
}

bool transfer_is_zero(const CompactTransferAccounting& a) { return a == CompactTransferAccounting{}; }

// The byte and timing envelope every execution must satisfy.
bool envelope_ok(const ExpansionBatchExecution& e) {
  const CompactTransferAccounting& a = e.compact_transfer;
  if (!sums_to(a.total_bytes, {a.summary_bytes, a.vertex_bytes, a.distance_label_bytes, a.edge_id_bytes}))
    return false;
  if (e.compact_total_device_to_host_bytes == 0U) {
    if (e.compact_status_bytes || e.compact_error_bytes || e.compact_controller_poll_count ||
        e.compact_controller_poll_bytes || e.compact_overall_device_to_host_bytes)
      return false;
  } else {
    if (!sums_to(e.compact_total_device_to_host_bytes,
                 {a.total_bytes, e.compact_status_bytes, e.compact_error_bytes}))
      return false;
    // A runner's control record size is its own; polls and their bytes come
    // together, and a persistent kernel is never polled.
    if ((e.compact_controller_poll_count == 0U) != (e.compact_controller_poll_bytes == 0U)) return false;
    const auto control = static_cast<ControlMode>(e.result.control_mode);
    if (control == ControlMode::persistent_cooperative && e.compact_controller_poll_count != 0U) return false;
    if (!sums_to(e.compact_overall_device_to_host_bytes,
                 {e.compact_total_device_to_host_bytes, e.compact_controller_poll_bytes}))
      return false;
  }
  const CompactPathExecutionMetrics& t = e.compact_execution;
  const auto known = [](CompactStageTimingEvidence v) {
    return v == CompactStageTimingEvidence::unavailable || v == CompactStageTimingEvidence::measured;
  };
  if (!known(t.device_timing) || !known(t.host_timing)) return false;
  if (t.device_timing == CompactStageTimingEvidence::unavailable &&
      (t.sssp_device_nanoseconds || t.reconstruction_device_nanoseconds || t.result_transfer_device_nanoseconds))
    return false;
  if (t.host_timing == CompactStageTimingEvidence::unavailable)
    return !(t.sssp_host_nanoseconds || t.reconstruction_host_nanoseconds ||
             t.result_transfer_host_nanoseconds || t.end_to_end_host_nanoseconds);
  const unsigned __int128 stages = static_cast<unsigned __int128>(t.sssp_host_nanoseconds) +
                                   t.reconstruction_host_nanoseconds + t.result_transfer_host_nanoseconds;
  return stages <= t.end_to_end_host_nanoseconds;
}

const RouteQuery& submitted(std::span<const RouteQuery> queries, QueryId id) {
  const auto it = std::lower_bound(queries.begin(), queries.end(), id,
                                   [](const RouteQuery& q, QueryId v) { return q.query_id < v; });
  if (it == queries.end() || it->query_id != id) throw std::logic_error{"batch names a query that was not submitted"};
  return *it;
}

// Compact-mode rules: payloads exactly for the valid lanes of a clean batch
// (sorted, unique, lane identity, valid), bytes equal to the payloads'
// measure; a nonclean batch carries summaries only.
void check_compact(const WeightedGraph& graph, std::span<const RouteQuery> queries, const BatchPlanEntry& batch,
                   const ExpansionBatchExecution& e) {
  if (e.compact_execution.host_timing != CompactStageTimingEvidence::measured)
    throw std::logic_error{"compact paths need measured host stage timing"};
  const bool ok = clean(e.result.status);
  const std::size_t expected = ok ? static_cast<std::size_t>(lanes_in(batch.valid_lane_mask)) : 0U;
  const auto by_id = [](const CompactPathPayload& a, const CompactPathPayload& b) { return a.query_id < b.query_id; };
  if (e.compact_paths.size() != expected || !std::is_sorted(e.compact_paths.begin(), e.compact_paths.end(), by_id) ||
      std::adjacent_find(e.compact_paths.begin(), e.compact_paths.end(),
                         [](const CompactPathPayload& a, const CompactPathPayload& b) {
                           return a.query_id == b.query_id;
                         }) != e.compact_paths.end())
    throw std::logic_error{"compact payloads do not match the batch's valid lanes"};
  for (const CompactPathPayload& p : e.compact_paths) {
    bool lane_found = false;
    for (std::uint32_t l = 0; l < batch.lane_width && !lane_found; ++l)
      lane_found = (batch.valid_lane_mask >> l & 1U) && batch.query_ids_by_lane[l] == p.query_id &&
                   batch.expansion_generations_by_lane[l] == p.expansion_generation;
    if (!lane_found || !detail::validate_compact_payload_trusted_graph(graph, submitted(queries, p.query_id), p).ok())
      throw std::logic_error{"a compact payload names no lane of the batch or fails validation"};
  }
  if (ok && e.compact_transfer != measure_compact_transfer(e.compact_paths))
    throw std::logic_error{"compact transfer bytes disagree with the returned paths"};
  if (!ok && e.compact_total_device_to_host_bytes == 0U && !transfer_is_zero(e.compact_transfer))
    throw std::logic_error{"an unmeasured nonclean batch carried transfer bytes"};
  if (!ok && e.compact_total_device_to_host_bytes != 0U) {
    std::uint64_t targets = 0;
    for (std::uint32_t l = 0; l < batch.lane_width; ++l)
      if (batch.valid_lane_mask >> l & 1U) add(targets, submitted(queries, batch.query_ids_by_lane[l]).targets.size());
    if (targets > kU64Max / sizeof(CompactTargetSummary)) throw std::overflow_error{"summary bytes overflow"};
    const CompactTransferAccounting& a = e.compact_transfer;
    if (a.summary_bytes != targets * sizeof(CompactTargetSummary) || a.vertex_bytes || a.distance_label_bytes ||
        a.edge_id_bytes)
      throw std::logic_error{"a nonclean batch's transfer is not summaries only"};
  }
  if (e.compact_total_device_to_host_bytes != 0U && e.compact_status_bytes != sizeof(DeviceRunStatus))
    throw std::logic_error{"compact transfer does not carry exactly one status record"};
}

const CompactPathPayload* payload_of(const std::vector<CompactPathPayload>& payloads, QueryId id) {
  const auto it = std::lower_bound(payloads.begin(), payloads.end(), id,
                                   [](const CompactPathPayload& p, QueryId v) { return p.query_id < v; });
  return it != payloads.end() && it->query_id == id ? &*it : nullptr;
}

struct TargetClasses {
  bool any_unreachable{};
  bool all_classified{true};  // every target complete or unreachable
  bool any_reconstruction_failure{};
};

TargetClasses classify(const CompactPathPayload& p) {
  TargetClasses c;
  for (const CompactTargetPath& t : p.targets) {
    const CompactPathStatus s = t.summary.reconstruction;
    c.any_unreachable = c.any_unreachable || s == CompactPathStatus::unreachable;
    c.all_classified = c.all_classified && (s == CompactPathStatus::complete || s == CompactPathStatus::unreachable);
    c.any_reconstruction_failure = c.any_reconstruction_failure || s == CompactPathStatus::no_tight_path ||
                                   s == CompactPathStatus::path_length_overflow ||
                                   s == CompactPathStatus::query_terminal_failure;
  }
  return c;
}

bool directory_matches(const TileDirectory& a, const TileDirectory& b) {
  return a.spill_tile() == b.spill_tile() && std::ranges::equal(a.neighbor_tile_offsets(), b.neighbor_tile_offsets()) &&
         std::ranges::equal(a.neighbor_tiles(), b.neighbor_tiles()) &&
         std::ranges::equal(a.internal_edge_offsets(), b.internal_edge_offsets()) &&
         std::ranges::equal(a.internal_edge_ids(), b.internal_edge_ids()) &&
         std::ranges::equal(a.outgoing_cross_edge_offsets(), b.outgoing_cross_edge_offsets()) &&
         std::ranges::equal(a.outgoing_cross_edge_ids(), b.outgoing_cross_edge_ids()) &&
         std::ranges::equal(a.incoming_cross_edge_offsets(), b.incoming_cross_edge_offsets()) &&
         std::ranges::equal(a.incoming_cross_edge_ids(), b.incoming_cross_edge_ids()) &&
         std::ranges::equal(a.halo_vertex_offsets(), b.halo_vertex_offsets()) &&
         std::ranges::equal(a.halo_vertices(), b.halo_vertices());
}

std::uint64_t comparison_fingerprint(const DeviceGraphFingerprint& g, std::span<const QueryState> states,
                                     const BatchedExpansionOptions& o) {
  Fnv f;
  f.word(g.first);
  f.word(g.second);
  const GpuRunOptions& r = o.run_options;
  for (const std::uint64_t w :
       {std::uint64_t{static_cast<std::uint32_t>(r.engine)}, std::uint64_t{static_cast<std::uint32_t>(r.control_mode)},
        std::uint64_t{r.rounds_per_chunk}, std::uint64_t{r.block_size},
        std::uint64_t{static_cast<std::uint32_t>(r.grid_policy)}, std::uint64_t{r.blocks_per_wgp},
        std::uint64_t{static_cast<std::uint32_t>(r.instrumentation)}, r.maximum_rounds,
        std::uint64_t{r.enable_per_lane_convergence}, std::uint64_t{o.planner_policy.lane_width},
        std::uint64_t{o.planner_policy.minimum_jaccard_numerator},
        std::uint64_t{o.planner_policy.minimum_jaccard_denominator},
        std::uint64_t{o.planner_policy.maximum_union_inflation_numerator},
        std::uint64_t{o.planner_policy.maximum_union_inflation_denominator}, o.execution_configuration_fingerprint,
        std::uint64_t{o.maximum_expansions}, std::uint64_t{static_cast<std::uint32_t>(o.terminal_policy)},
        std::uint64_t{o.enable_compact_paths}, std::uint64_t{states.size()}})
    f.word(w);
  const auto vid = [](VertexId v) -> std::uint64_t { return v.value(); };
  const auto u32 = [](std::uint32_t v) -> std::uint64_t { return v; };
  for (const QueryState& s : states) {
    const RouteQuery& q = s.query;
    f.word(q.query_id.value());
    f.word(q.expansion_generation);
    f.list<VertexId>(q.source_terminals, vid);
    f.list<VertexId>(q.target_terminals, vid);
    f.list<VertexId>(q.sources, vid);
    f.list<VertexId>(q.targets, vid);
    f.list<std::uint32_t>(q.source_terminal_to_source, u32);
    f.list<std::uint32_t>(q.target_terminal_to_target, u32);
    f.list<TileId>(q.selected_tiles, [](TileId t) -> std::uint64_t { return t.value(); });
  }
  return f.h == 0U ? 1U : f.h;
}

std::uint64_t milliqueries_per_second(std::uint64_t queries, std::uint64_t ns) {
  if (queries == 0U || ns == 0U) return 0U;
  const unsigned __int128 r = static_cast<unsigned __int128>(queries) * 1000000000000ULL / ns;
  return r > kU64Max ? kU64Max : static_cast<std::uint64_t>(r);
}

// Orders evidence records: true when `a` ranks strictly before `b`.
bool ranks_before(const ExpansionScheduleEvidence& a, const ExpansionScheduleEvidence& b, bool work_measured) {
  if (a.reached_queries != b.reached_queries) return a.reached_queries > b.reached_queries;
  if (a.terminal_failures != b.terminal_failures) return a.terminal_failures < b.terminal_failures;
  if (work_measured && a.logical_lane_edge_work != b.logical_lane_edge_work)
    return a.logical_lane_edge_work < b.logical_lane_edge_work;
  if (work_measured && a.shared_edge_work != b.shared_edge_work) return a.shared_edge_work < b.shared_edge_work;
  if (a.attempted_selected_edge_estimate != b.attempted_selected_edge_estimate)
    return a.attempted_selected_edge_estimate < b.attempted_selected_edge_estimate;
  if (a.total_expansions != b.total_expansions) return a.total_expansions < b.total_expansions;
  return a.batches_executed < b.batches_executed;
}

}  // namespace

bool BatchedExpansionOptions::operator==(const BatchedExpansionOptions& o) const noexcept {
  const GpuRunOptions& a = run_options;
  const GpuRunOptions& b = o.run_options;
  return a.engine == b.engine && a.control_mode == b.control_mode && a.rounds_per_chunk == b.rounds_per_chunk &&
         a.block_size == b.block_size && a.grid_policy == b.grid_policy && a.blocks_per_wgp == b.blocks_per_wgp &&
         a.instrumentation == b.instrumentation && a.maximum_rounds == b.maximum_rounds &&
         a.enable_per_lane_convergence == b.enable_per_lane_convergence && planner_policy == o.planner_policy &&
         execution_configuration_fingerprint == o.execution_configuration_fingerprint && schedule == o.schedule &&
         maximum_expansions == o.maximum_expansions && terminal_policy == o.terminal_policy &&
         enable_compact_paths == o.enable_compact_paths;
}

BatchedExpansionOptionsError validate_batched_expansion_options(const BatchedExpansionOptions& o) noexcept {
  using E = BatchedExpansionOptionsError;
  if (validate_gpu_run_options(o.run_options) != GpuRunOptionsError::none) return E::invalid_run_options;
  if (validate_batch_planner_policy(o.planner_policy) != BatchPlannerPolicyError::none)
    return E::invalid_planner_policy;
  if (o.execution_configuration_fingerprint == 0U) return E::missing_execution_configuration_fingerprint;
  const ExpansionSchedulePolicy& s = o.schedule;
  switch (s.kind) {
    case ExpansionScheduleKind::unspecified:
      return E::unspecified_schedule;
    case ExpansionScheduleKind::one_geometric_ring:
    case ExpansionScheduleKind::doubling_xy_margins:
      if (s.fixed_ring_size != 0U) return E::unexpected_fixed_ring_size;
      if (s.hybrid_small_expansion_count != 0U) return E::unexpected_hybrid_small_expansion_count;
      break;
    case ExpansionScheduleKind::fixed_larger_ring:
      if (s.fixed_ring_size < 2U) return E::invalid_fixed_ring_size;
      if (s.hybrid_small_expansion_count != 0U) return E::unexpected_hybrid_small_expansion_count;
      break;
    case ExpansionScheduleKind::hybrid_small_then_doubling:
      if (s.fixed_ring_size != 0U) return E::unexpected_fixed_ring_size;
      if (s.hybrid_small_expansion_count == 0U) return E::invalid_hybrid_small_expansion_count;
      break;
    default:
      return E::invalid_schedule;
  }
  if (o.terminal_policy != ExpansionTerminalPolicy::full_region_fallback &&
      o.terminal_policy != ExpansionTerminalPolicy::explicit_failure)
    return E::invalid_terminal_policy;
  if (o.enable_compact_paths > 1U) return E::invalid_compact_paths_flag;
  return E::none;
}

BatchedExpansionRunResult run_batched_expansion(const WeightedGraph& graph, const TileDirectory& directory,
                                                const TileRunLayout64& tile_runs,
                                                std::span<const RouteQuery> queries,
                                                const BatchedExpansionOptions& options,
                                                const ExpansionBatchRunner& runner) {
  if (validate_batched_expansion_options(options) != BatchedExpansionOptionsError::none)
    throw std::invalid_argument{"batched expansion: invalid options"};
  if (!runner) throw std::invalid_argument{"batched expansion: no batch runner"};
  if (!graph.has_spatial_ordering()) throw std::invalid_argument{"batched expansion: graph is not spatially ordered"};
  if (const std::string& why = weighted_graph_validation(graph); !why.empty())
    throw std::invalid_argument{"batched expansion: invalid graph: " + why};
  if (!directory.built_from(graph) && !directory_matches(directory, build_tile_directory(graph)))
    throw std::invalid_argument{"batched expansion: the tile directory is not the graph's"};
  // Also checks that the tile runs are the graph's (std::invalid_argument
  // otherwise); the graph keeps the value, so later runs do not pay again.
  const DeviceGraphFingerprint graph_identity = fingerprint_device_graph_source32(graph, tile_runs);

  const auto t_total = Clock::now();
  BatchedExpansionRunResult out;
  BatchedExpansionMetrics& m = out.metrics;
  const bool compact = options.enable_compact_paths != 0U;
  m.input_queries = queries.size();
  if (queries.empty()) {
    if (compact) m.compact_host_timing = CompactStageTimingEvidence::measured;
    m.total_nanoseconds = since(t_total);
    m.expansion_count_histogram.assign(1U, 0U);
    return out;
  }

  // tile_runs equal the graph's own (checked above), so the graph's memoised
  // index serves.
  const SelectedRegionIndex& regions = graph_region_index(graph);
  std::vector<TileId> every_tile(graph.tile_coordinates().size());
  for (std::size_t t = 0; t < every_tile.size(); ++t) every_tile[t] = TileId{static_cast<std::uint32_t>(t)};

  std::vector<QueryState> states(queries.size());
  for (std::size_t i = 0; i < queries.size(); ++i) {
    if (!validate_route_query(graph, queries[i]).ok())
      throw std::invalid_argument{"batched expansion: invalid route query"};
    states[i].query = queries[i];
    states[i].anchor = located_box(graph, queries[i].selected_tiles);
  }
  std::sort(states.begin(), states.end(),
            [](const QueryState& a, const QueryState& b) { return a.query.query_id < b.query.query_id; });
  for (std::size_t i = 1; i < states.size(); ++i)
    if (states[i - 1].query.query_id == states[i].query.query_id)
      throw std::invalid_argument{"batched expansion: query ids must be unique"};
  m.schedule_comparison_fingerprint = comparison_fingerprint(graph_identity, states, options);

  const auto state_of = [&states](QueryId id) -> QueryState& {
    const auto it = std::lower_bound(states.begin(), states.end(), id,
                                     [](const QueryState& s, QueryId v) { return s.query.query_id < v; });
    if (it == states.end() || it->query.query_id != id) throw std::logic_error{"batch names an unknown query"};
    return *it;
  };

  // Indices into `states`, ascending (so QueryId order).
  std::vector<std::size_t> active(states.size());
  for (std::size_t i = 0; i < active.size(); ++i) active[i] = i;
  std::uint64_t ordinal = 0;
  std::uint32_t pass = 0;
  while (!active.empty()) {
    std::vector<RouteQuery> pass_queries;
    pass_queries.reserve(active.size());
    for (const std::size_t i : active) pass_queries.push_back(states[i].query);

    const auto t_plan = Clock::now();
    const std::vector<BatchQueryFeatures> features = make_batch_query_features(graph, regions, pass_queries);
    const BatchPlan plan = make_overlapping_batch_plan(regions, features, options.planner_policy);
    if (!validate_batch_plan(regions, features, plan).ok())
      throw std::logic_error{"batched expansion: the pass's plan failed validation"};
    for (std::size_t i = 0; i < features.size(); ++i)
      if (features[i].query_id != pass_queries[i].query_id)
        throw std::logic_error{"batched expansion: features and queries are not aligned"};
    (pass == 0 ? m.initial_planning_nanoseconds : m.replanning_nanoseconds) += since(t_plan);
    add(m.planning_passes, 1U);

    std::vector<std::size_t> missed;
    for (std::size_t bi = 0; bi < plan.batches.size(); ++bi) {
      const BatchPlanEntry& batch = plan.batches[bi];
      const ExpansionBatchContext context{ordinal, pass, static_cast<std::uint32_t>(bi), pass != 0U};
      const auto t_run = Clock::now();
      ExpansionBatchExecution e = runner(pass_queries, features, batch, context);
      add(m.execution_nanoseconds, since(t_run));
      add(ordinal, 1U);

      // The result must describe this batch.
      if (e.result.engine_kind != static_cast<std::uint32_t>(options.run_options.engine) ||
          e.result.control_mode != static_cast<std::uint32_t>(options.run_options.control_mode))
        throw std::logic_error{"batch runner returned a result of another engine or control mode"};
      if (const DeviceRunStatusError why = validate_device_run_status(e.result.status);
          why != DeviceRunStatusError::none)
        throw std::logic_error{"batch runner returned an invalid status (error " +
                               std::to_string(static_cast<int>(why)) + ")"};
      if (e.result.status.valid_lane_mask != batch.valid_lane_mask)
        throw std::logic_error{"batch runner returned the status of other lanes"};
      const bool measured = e.work_evidence == ExpansionWorkEvidence::measured;
      if (!measured && e.work_evidence != ExpansionWorkEvidence::unavailable)
        throw std::logic_error{"batch runner returned an unknown work evidence state"};
      if (!measured && (e.shared_edge_work || e.logical_lane_edge_work || !work_is_zero(e.result.work)))
        throw std::logic_error{"batch runner reported work it did not measure"};
      if (measured && e.logical_lane_edge_work < e.shared_edge_work)
        throw std::logic_error{"logical lane-edge work below shared edge work"};
      if (!e.final_distances.empty() &&
          e.final_distances.size() != static_cast<std::uint64_t>(graph.vertex_count()) * batch.lane_width)
        throw std::logic_error{"batch runner returned a label image of the wrong size"};
      if (!envelope_ok(e)) throw std::logic_error{"batch runner returned inconsistent transfer or timing records"};
      if (compact) {
        check_compact(graph, pass_queries, batch, e);
      } else if (!e.compact_paths.empty() || !transfer_is_zero(e.compact_transfer) || e.compact_status_bytes ||
                 e.compact_error_bytes || e.compact_total_device_to_host_bytes || e.compact_controller_poll_count ||
                 e.compact_controller_poll_bytes || e.compact_overall_device_to_host_bytes ||
                 e.compact_execution.device_timing != CompactStageTimingEvidence::unavailable ||
                 e.compact_execution.host_timing != CompactStageTimingEvidence::unavailable) {
        throw std::logic_error{"compact results returned without compact mode"};
      }

      // Aggregates.
      add(m.batches_executed, 1U);
      if (pass == 0) {
        add(m.initial_batches_executed, 1U);
      } else {
        add(m.retry_batches_executed, 1U);
        add(m.retry_valid_lane_observations, lanes_in(batch.valid_lane_mask));
        add(m.retry_lane_capacity, batch.lane_width);
      }
      add(m.attempted_selected_vertex_count, batch.selected_lane_vertex_count);
      add(m.attempted_selected_edge_estimate, batch.selected_lane_edge_estimate);
      if (measured) {
        add(m.work_measured_batches, 1U);
        add(m.shared_edge_work, e.shared_edge_work);
        add(m.logical_lane_edge_work, e.logical_lane_edge_work);
        if (pass != 0) {
          add(m.retry_work_measured_batches, 1U);
          add(m.retry_shared_edge_work, e.shared_edge_work);
          add(m.retry_logical_lane_edge_work, e.logical_lane_edge_work);
        }
      } else {
        m.work_evidence = ExpansionWorkEvidence::unavailable;
      }
      add_work(m.device_work, e.result.work);
      add_transfer(m.compact_transfer, e.compact_transfer);
      add(m.compact_status_bytes, e.compact_status_bytes);
      add(m.compact_error_bytes, e.compact_error_bytes);
      add(m.compact_total_device_to_host_bytes, e.compact_total_device_to_host_bytes);
      add(m.compact_controller_poll_count, e.compact_controller_poll_count);
      add(m.compact_controller_poll_bytes, e.compact_controller_poll_bytes);
      add(m.compact_overall_device_to_host_bytes, e.compact_overall_device_to_host_bytes);
      const CompactPathExecutionMetrics& ct = e.compact_execution;
      if (ct.device_timing == CompactStageTimingEvidence::measured) {
        add(m.compact_device_timing_measured_batches, 1U);
        add(m.sssp_device_nanoseconds, ct.sssp_device_nanoseconds);
        add(m.reconstruction_device_nanoseconds, ct.reconstruction_device_nanoseconds);
        add(m.result_transfer_device_nanoseconds, ct.result_transfer_device_nanoseconds);
      }
      if (ct.host_timing == CompactStageTimingEvidence::measured) {
        add(m.compact_host_timing_measured_batches, 1U);
        add(m.sssp_host_nanoseconds, ct.sssp_host_nanoseconds);
        add(m.reconstruction_host_nanoseconds, ct.reconstruction_host_nanoseconds);
        add(m.result_transfer_host_nanoseconds, ct.result_transfer_host_nanoseconds);
        add(m.compact_end_to_end_host_nanoseconds, ct.end_to_end_host_nanoseconds);
      }

      // A clean batch splits its valid lanes into reached and missed; any
      // other stop classifies none.
      const DeviceRunStatus& st = e.result.status;
      const bool ok = clean(st);
      if (ok && ((st.reached_target_mask | st.bounding_box_miss_mask) != batch.valid_lane_mask ||
                 (st.reached_target_mask & st.bounding_box_miss_mask) != 0U))
        throw std::logic_error{"a clean batch's reached and missed lanes do not split its valid lanes"};
      if (!ok && (st.reached_target_mask || st.bounding_box_miss_mask))
        throw std::logic_error{"a nonclean batch classified lanes"};

      ExpansionBatchTrace tr;
      tr.context = context;
      tr.lane_width = batch.lane_width;
      tr.valid_lane_mask = batch.valid_lane_mask;
      tr.reached_lane_mask = st.reached_target_mask;
      tr.miss_lane_mask = st.bounding_box_miss_mask;
      tr.query_ids_by_lane = batch.query_ids_by_lane;
      tr.expansion_generations_by_lane = batch.expansion_generations_by_lane;
      tr.union_vertex_count = batch.union_vertex_count;
      tr.union_edge_estimate = batch.union_edge_estimate;
      tr.selected_lane_vertex_count = batch.selected_lane_vertex_count;
      tr.selected_lane_edge_estimate = batch.selected_lane_edge_estimate;
      tr.shared_edge_work = e.shared_edge_work;
      tr.logical_lane_edge_work = e.logical_lane_edge_work;
      tr.work_evidence = e.work_evidence;
      tr.stop_reason = st.stop_reason;
      tr.error_bits = st.error_bits;
      out.trace.push_back(std::move(tr));

      if (ok && st.bounding_box_miss_mask != 0U) {
        add(m.failed_lane_observations, lanes_in(st.bounding_box_miss_mask));
        add(m.failed_origin_valid_lane_observations, lanes_in(batch.valid_lane_mask));
        if (measured) {
          add(m.failed_batch_work_measured_batches, 1U);
          add(m.failed_batch_shared_edge_work, e.shared_edge_work);
          add(m.failed_batch_logical_lane_edge_work, e.logical_lane_edge_work);
        }
      }

      // Per lane: record the attempt, then retire or queue for expansion.
      for (std::uint32_t lane = 0; lane < batch.lane_width; ++lane) {
        const LaneMask bit = LaneMask{1U} << lane;
        if (!(batch.valid_lane_mask & bit)) continue;
        QueryState& q = state_of(batch.query_ids_by_lane[lane]);
        if (q.done || q.query.expansion_generation != batch.expansion_generations_by_lane[lane])
          throw std::logic_error{"batch runner crossed a query's generation"};
        if (q.attempts == kU32Max) {
          q.distances.clear();
          q.paths.reset();
          finish(q, ExpansionQueryDisposition::identity_or_count_overflow, st);
          continue;
        }
        ++q.attempts;
        q.distances.clear();
        if (!e.final_distances.empty()) {
          const std::uint32_t V = graph.vertex_count();
          q.distances.resize(V);
          for (std::uint32_t v = 0; v < V; ++v)
            q.distances[v] = e.final_distances[static_cast<std::size_t>(v) * batch.lane_width + lane];
        }
        q.paths.reset();
        if (!ok) {
          q.distances.clear();
          finish(q, ExpansionQueryDisposition::engine_failure, st);
          continue;
        }
        const CompactPathPayload* payload = compact ? payload_of(e.compact_paths, q.query.query_id) : nullptr;
        if (compact && !payload) throw std::logic_error{"a valid lane of a clean batch has no compact payload"};
        if (st.reached_target_mask & bit) {
          if (compact) {
            q.paths = *payload;
            q.distances.clear();
            if (!compact_path_payload_complete(*q.paths)) {
              // Reached, but a path could not be rebuilt: keep the payload
              // only when it names the reconstruction failure.
              if (!classify(*q.paths).any_reconstruction_failure) q.paths.reset();
              finish(q, ExpansionQueryDisposition::engine_failure, st);
              continue;
            }
          }
          finish(q, ExpansionQueryDisposition::reached, st);
          if (pass == 0) add(m.initial_reached_queries, 1U);
          continue;
        }
        if (compact) {
          q.paths = *payload;
          const TargetClasses c = classify(*q.paths);
          if (!c.any_unreachable || !c.all_classified) {
            // A miss must name an unreachable target and classify the rest.
            if (c.all_classified) q.paths.reset();
            q.distances.clear();
            finish(q, ExpansionQueryDisposition::engine_failure, st);
            continue;
          }
        }
        missed.push_back(static_cast<std::size_t>(&q - states.data()));
      }
    }

    std::sort(missed.begin(), missed.end());
    if (std::adjacent_find(missed.begin(), missed.end()) != missed.end())
      throw std::logic_error{"a query missed twice in one pass"};
    active.clear();
    const auto t_grow = Clock::now();
    for (const std::size_t i : missed) {
      QueryState& q = states[i];
      DeviceRunStatus miss{};
      miss.valid_lane_mask = 1U;
      miss.converged_lane_mask = 1U;
      miss.converged = 1U;
      miss.bounding_box_miss_mask = 1U;
      miss.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::converged);
      if (after_miss(graph, directory, regions, every_tile, options, miss, q, m) == Next::retry) {
        active.push_back(i);
      } else if (q.disposition != ExpansionQueryDisposition::unreachable_in_full_region) {
        q.paths.reset();
      } else if (compact) {
        if (!q.paths) throw std::logic_error{"a full-region miss lost its compact payload"};
        const TargetClasses c = classify(*q.paths);
        if (!c.any_unreachable) throw std::logic_error{"a full-region miss names no unreachable target"};
        if (!c.all_classified) q.disposition = ExpansionQueryDisposition::engine_failure;
      }
    }
    if (!missed.empty()) add(m.geometric_expansion_host_nanoseconds, since(t_grow));
    if (!active.empty()) {
      if (pass == kU32Max) throw std::overflow_error{"batched expansion: planning pass overflow"};
      ++pass;
    }
  }

  // Outcomes in QueryId order.
  out.queries.reserve(states.size());
  std::uint32_t most = 0;
  for (QueryState& q : states) {
    if (!q.done) throw std::logic_error{"batched expansion left a query without an outcome"};
    ExpansionQueryOutcome o;
    o.final_query = std::move(q.query);
    o.disposition = q.disposition;
    o.attempts = q.attempts;
    o.scheduled_expansions = q.scheduled;
    o.total_expansions = q.total;
    o.used_full_region_fallback = q.fallback_used;
    o.selected_vertex_count = regions.selected_vertex_count(o.final_query.selected_tiles);
    o.selected_edge_count = regions.selected_edge_count(o.final_query.selected_tiles);
    o.terminal_stop_reason = q.stop_reason;
    o.terminal_error_bits = q.error_bits;
    o.final_distances = std::move(q.distances);
    o.compact_paths = std::move(q.paths);
    most = std::max(most, o.total_expansions);
    switch (o.disposition) {
      case ExpansionQueryDisposition::reached: add(m.reached_queries, 1U); break;
      case ExpansionQueryDisposition::unreachable_in_full_region: add(m.unreachable_full_region_queries, 1U); break;
      case ExpansionQueryDisposition::expansion_limit: add(m.expansion_limit_queries, 1U); break;
      case ExpansionQueryDisposition::region_stalled: add(m.stalled_region_queries, 1U); break;
      case ExpansionQueryDisposition::identity_or_count_overflow: add(m.identity_or_count_overflow_queries, 1U); break;
      case ExpansionQueryDisposition::engine_failure: add(m.engine_failure_queries, 1U); break;
    }
    add(m.final_selected_tile_count, o.final_query.selected_tiles.size());
    add(m.final_selected_vertex_count, o.selected_vertex_count);
    add(m.final_selected_edge_count, o.selected_edge_count);
    out.queries.push_back(std::move(o));
  }
  m.expansion_count_histogram.assign(static_cast<std::size_t>(most) + 1U, 0U);
  for (const ExpansionQueryOutcome& o : out.queries) add(m.expansion_count_histogram[o.total_expansions], 1U);
  m.device_work.expansion_count = m.scheduled_expansions;
  add(m.device_work.expansion_count, m.full_region_fallbacks);
  if (m.batches_executed != 0U && m.compact_device_timing_measured_batches == m.batches_executed)
    m.compact_device_timing = CompactStageTimingEvidence::measured;
  if (m.batches_executed != 0U && m.compact_host_timing_measured_batches == m.batches_executed)
    m.compact_host_timing = CompactStageTimingEvidence::measured;
  m.total_nanoseconds = since(t_total);
  m.all_query_throughput_milliqueries_per_second = milliqueries_per_second(m.input_queries, m.total_nanoseconds);
  return out;
}

ExpansionScheduleEvidence make_expansion_schedule_evidence(const ExpansionSchedulePolicy& schedule,
                                                           const BatchedExpansionRunResult& result) noexcept {
  const BatchedExpansionMetrics& m = result.metrics;
  ExpansionScheduleEvidence e;
  e.schedule = schedule;
  e.input_queries = m.input_queries;
  e.reached_queries = m.reached_queries;
  e.terminal_failures = m.input_queries >= m.reached_queries ? m.input_queries - m.reached_queries : kU64Max;
  e.batches_executed = m.batches_executed;
  const bool total_fits = m.scheduled_expansions <= kU64Max - m.full_region_fallbacks;
  e.total_expansions = total_fits ? m.scheduled_expansions + m.full_region_fallbacks : kU64Max;
  e.attempted_selected_edge_estimate = m.attempted_selected_edge_estimate;
  e.work_evidence = m.work_evidence;
  e.shared_edge_work = m.shared_edge_work;
  e.logical_lane_edge_work = m.logical_lane_edge_work;
  e.comparison_fingerprint = m.schedule_comparison_fingerprint;
  e.campaign_valid = m.engine_failure_queries == 0U && m.identity_or_count_overflow_queries == 0U && total_fits;
  return e;
}

std::optional<ExpansionSchedulePolicy> select_expansion_schedule_from_evidence(
    std::span<const ExpansionScheduleEvidence> evidence) noexcept {
  constexpr std::size_t kinds = 4U;
  if (evidence.size() != kinds) return std::nullopt;
  const std::uint64_t n = evidence.front().input_queries;
  const std::uint64_t identity = evidence.front().comparison_fingerprint;
  if (n == 0U || identity == 0U) return std::nullopt;
  std::uint32_t seen = 0;
  bool work_measured = true;
  for (const ExpansionScheduleEvidence& r : evidence) {
    const bool work_known =
        r.work_evidence == ExpansionWorkEvidence::measured || r.work_evidence == ExpansionWorkEvidence::unavailable;
    if (!schedule_ok(r.schedule) || r.input_queries != n || r.comparison_fingerprint != identity ||
        r.reached_queries > r.input_queries || r.batches_executed == 0U || !r.campaign_valid ||
        r.terminal_failures != r.input_queries - r.reached_queries || !work_known ||
        (r.work_evidence == ExpansionWorkEvidence::unavailable && (r.shared_edge_work || r.logical_lane_edge_work)) ||
        (r.work_evidence == ExpansionWorkEvidence::measured && r.logical_lane_edge_work < r.shared_edge_work))
      return std::nullopt;
    const auto kind = static_cast<std::uint32_t>(r.schedule.kind);
    const std::uint32_t bit = 1U << (kind - 1U);
    if (seen & bit) return std::nullopt;
    seen |= bit;
    work_measured = work_measured && r.work_evidence == ExpansionWorkEvidence::measured;
  }
  if (seen != (1U << kinds) - 1U) return std::nullopt;
  const ExpansionScheduleEvidence* best = &evidence.front();
  for (const ExpansionScheduleEvidence& r : evidence)
    if (ranks_before(r, *best, work_measured)) best = &r;
  for (const ExpansionScheduleEvidence& r : evidence)
    if (&r != best && !ranks_before(*best, r, work_measured)) return std::nullopt;
  return best->schedule;
}

}  // namespace bfnew
