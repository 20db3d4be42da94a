// gpu_api.cpp -- option, controller and status rules of the GPU engines
// (reference rules: src/gpu_api.cpp:70-276).  Each validator returns the
// first rule a record breaks, in the order the error enums list them.
#include "bfnew/gpu_api.hpp"

#include <stdexcept>

namespace bfnew {
namespace {

constexpr bool known(EngineKind e) noexcept { return static_cast<std::uint32_t>(e) <= 2U; }
constexpr bool known(ControlMode m) noexcept { return static_cast<std::uint32_t>(m) <= 2U; }
constexpr bool known(GridPolicy p) noexcept { return static_cast<std::uint32_t>(p) <= 1U; }
constexpr bool known(InstrumentationLevel i) noexcept {
  return static_cast<std::uint32_t>(i) <= 2U;
}
constexpr bool known_stop(std::uint32_t reason) noexcept {
  return reason <= static_cast<std::uint32_t>(DeviceStopReason::device_failure);
}

// The error bit that a failure stop carries; zero for the normal stops.
constexpr std::uint32_t error_bit_of(std::uint32_t reason) noexcept {
  switch (static_cast<DeviceStopReason>(reason)) {
    case DeviceStopReason::queue_overflow: return device_error::queue_overflow;
    case DeviceStopReason::invalid_controller_state: return device_error::invalid_controller_state;
    case DeviceStopReason::device_failure: return device_error::device_failure;
    default: return device_error::none;
  }
}

}  // namespace

GpuRunOptionsError validate_gpu_run_options(const GpuRunOptions& o) noexcept {
  using E = GpuRunOptionsError;
  if (!known(o.engine)) return E::invalid_engine;
  if (!known(o.control_mode)) return E::invalid_control_mode;
  if (o.rounds_per_chunk == 0U) return E::invalid_rounds_per_chunk;
  if (o.block_size == 0U) return E::invalid_block_size;
  if (!known(o.grid_policy)) return E::invalid_grid_policy;
  // A fixed grid needs a positive count; a derived grid takes none.
  const bool fixed = o.grid_policy == GridPolicy::fixed_blocks_per_wgp;
  if (fixed == (o.blocks_per_wgp == 0U)) return E::invalid_blocks_per_wgp;
  if (!known(o.instrumentation)) return E::invalid_instrumentation;
  if (o.maximum_rounds == 0U) return E::invalid_maximum_rounds;
  if (o.enable_per_lane_convergence > 1U) return E::invalid_per_lane_flag;
  return E::none;
}

DeviceController initialize_device_controller(const GpuRunOptions& options,
                                              LaneMask valid_lane_mask,
                                              std::uint64_t initial_frontier_size) {
  if (validate_gpu_run_options(options) != GpuRunOptionsError::none)
    throw std::invalid_argument{"device controller: invalid run options"};
  if (valid_lane_mask == 0U)
    throw std::invalid_argument{"device controller: no valid lane"};
  const bool frontier = options.engine == EngineKind::frontier_push;
  if (frontier && initial_frontier_size == 0U)
    throw std::invalid_argument{"device controller: frontier engine with an empty frontier"};
  DeviceController c{};
  c.valid_lane_mask = c.active_lane_mask = c.execute_lane_mask = valid_lane_mask;
  c.engine_kind = static_cast<std::uint32_t>(options.engine);
  c.enable_per_lane_convergence = options.enable_per_lane_convergence;
  c.maximum_rounds = options.maximum_rounds;
  // Jacobi reads one column and writes the other; the push engines work in
  // place.  Frontier slots alternate, the first read slot holding the seeds.
  c.distance_read_slot = 0U;
  c.distance_write_slot = options.engine == EngineKind::jacobi_pull ? 1U : 0U;
  c.frontier_read_slot = 0U;
  c.frontier_write_slot = 1U;
  c.frontier_size[0] = frontier ? initial_frontier_size : 0U;
  return c;
}

DeviceControllerError validate_device_controller(const DeviceController& c) noexcept {
  using E = DeviceControllerError;
  const auto engine = static_cast<EngineKind>(c.engine_kind);
  if (!known(engine)) return E::invalid_engine;
  if (c.valid_lane_mask == 0U) return E::empty_valid_lanes;
  const LaneMask any = c.active_lane_mask | c.changed_lane_mask | c.converged_lane_mask |
                       c.execute_lane_mask | c.next_frontier_lane_mask;
  if (any & ~c.valid_lane_mask) return E::mask_outside_valid_lanes;
  if (c.active_lane_mask & c.converged_lane_mask) return E::active_converged_overlap;
  // A running controller executes exactly its active lanes; a stopped one none.
  const LaneMask expected_execute = c.done ? 0U : c.active_lane_mask;
  if (c.execute_lane_mask != expected_execute) return E::execute_mask_mismatch;
  if (c.enable_per_lane_convergence > 1U) return E::invalid_per_lane_flag;
  if (c.maximum_rounds == 0U) return E::invalid_round_limit;
  if (c.rounds_completed > c.maximum_rounds) return E::completed_round_overflow;
  const bool jacobi = engine == EngineKind::jacobi_pull;
  const bool slots_in_range = c.distance_read_slot <= 1U && c.distance_write_slot <= 1U;
  const bool slots_shaped = jacobi ? c.distance_read_slot != c.distance_write_slot
                                   : (c.distance_read_slot == 0U && c.distance_write_slot == 0U);
  if (!slots_in_range || !slots_shaped) return E::invalid_distance_slots;
  if (c.frontier_read_slot > 1U || c.frontier_write_slot > 1U ||
      c.frontier_read_slot == c.frontier_write_slot)
    return E::invalid_frontier_slots;
  if (engine == EngineKind::frontier_push && !c.done && c.frontier_size[c.frontier_read_slot] == 0U)
    return E::invalid_frontier_slots;
  if (c.done > 1U) return E::invalid_done_flag;
  if (!known_stop(c.stop_reason)) return E::invalid_stop_reason;
  const bool stopped = c.stop_reason != static_cast<std::uint32_t>(DeviceStopReason::none);
  if ((c.done != 0U) != stopped) return E::inconsistent_done_state;
  if (c.error_bits & ~device_error::known_mask) return E::unknown_error_bits;
  if (c.error_bits != error_bit_of(c.stop_reason)) return E::inconsistent_error_state;
  if (c.done && (c.changed_lane_mask | c.next_frontier_lane_mask)) return E::inconsistent_terminal_state;
  const auto reason = static_cast<DeviceStopReason>(c.stop_reason);
  if (reason == DeviceStopReason::converged &&
      (c.active_lane_mask != 0U || c.converged_lane_mask != c.valid_lane_mask))
    return E::inconsistent_terminal_state;
  if (reason == DeviceStopReason::maximum_rounds && c.rounds_completed != c.maximum_rounds)
    return E::inconsistent_terminal_state;
  return E::none;
}

DeviceRunStatusError validate_device_run_status(const DeviceRunStatus& s) noexcept {
  using E = DeviceRunStatusError;
  if (s.final_distance_slot > 1U) return E::invalid_distance_slot;
  if (s.converged > 1U) return E::invalid_converged_flag;
  if (s.valid_lane_mask == 0U) return E::empty_valid_lanes;
  const LaneMask any =
      s.reached_target_mask | s.bounding_box_miss_mask | s.active_lane_mask | s.converged_lane_mask;
  if (any & ~s.valid_lane_mask) return E::mask_outside_valid_lanes;
  if (s.active_lane_mask & s.converged_lane_mask) return E::active_converged_overlap;
  if (s.reached_target_mask & s.bounding_box_miss_mask) return E::reached_miss_overlap;
  if (!known_stop(s.stop_reason)) return E::invalid_stop_reason;
  const bool converged_stop = s.stop_reason == static_cast<std::uint32_t>(DeviceStopReason::converged);
  if ((s.converged != 0U) != converged_stop) return E::inconsistent_convergence_state;
  if (converged_stop && (s.active_lane_mask != 0U || s.converged_lane_mask != s.valid_lane_mask))
    return E::inconsistent_terminal_state;
  if (s.error_bits & ~device_error::known_mask) return E::unknown_error_bits;
  if (s.error_bits != error_bit_of(s.stop_reason)) return E::inconsistent_error_state;
  if (s.error_bits != device_error::none && s.bounding_box_miss_mask != 0U)
    return E::error_reported_as_bounding_box_miss;
  return E::none;
}

}  // namespace bfnew
