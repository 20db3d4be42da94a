// runtime.hpp -- device, resident graph and workspace owners of the MI355X
// engine (reference: include/bfnew/hip/runtime.hpp:19-327).
//
// Like the reference, no HIP header is included: every object wraps a handle
// of the C ABI in bfnew_amd.h, and every failure of that ABI is rethrown as
// the reference's exception type (HipRuntimeError for HIP failures,
// std::invalid_argument for bad arguments, std::runtime_error for an
// inconsistent device result, std::logic_error for a call out of order).
#pragma once

#include "bfnew/device_layout.hpp"
#include "bfnew_amd.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace bfnew::hip {

class HipRuntimeError final : public std::runtime_error {
 public:
  HipRuntimeError(std::int32_t status, std::string message);
  [[nodiscard]] std::int32_t status() const noexcept { return status_; }

 private:
  std::int32_t status_{};
};

// Throws the exception that matches a C ABI status (no-op for BFN_OK); the
// message is `what` followed by bfn_last_error().
void throw_if_error(std::int32_t status, std::string_view what);

// A device and its streams.  Work of an engine runs on the stream its
// workspace owns on this device; synchronize() waits for all of it.
class HipStream final {
 public:
  explicit HipStream(bool nonblocking = true, std::int32_t device_ordinal = 0);
  ~HipStream() noexcept;
  HipStream(const HipStream&) = delete;
  HipStream& operator=(const HipStream&) = delete;
  HipStream(HipStream&& other) noexcept;
  HipStream& operator=(HipStream&& other) noexcept;

  void synchronize() const;
  [[nodiscard]] void* native_handle() const noexcept { return device_; }
  [[nodiscard]] bfn_device device() const noexcept { return device_; }

 private:
  bfn_device device_{};
};

// Staging image plus the derived identity of one graph upload.  The tile
// vertex offsets are derived from owner_tiles (vertices are tile-major).
struct ResidentGraphPlan {
  DeviceGraphLayout32 layout;
  DeviceGraphMemoryReport memory;
  DeviceGraphFingerprint fingerprint;
  std::vector<std::uint32_t> tile_vertex_offsets;
};

// Throws std::invalid_argument when owner_tiles is not tile-major.
[[nodiscard]] ResidentGraphPlan make_resident_graph_plan(DeviceGraphLayout32 layout);

class ResidentDeviceGraph final {
 public:
  ResidentDeviceGraph() = default;
  ~ResidentDeviceGraph() noexcept;
  ResidentDeviceGraph(const ResidentDeviceGraph&) = delete;
  ResidentDeviceGraph& operator=(const ResidentDeviceGraph&) = delete;
  ResidentDeviceGraph(ResidentDeviceGraph&& other) noexcept;
  ResidentDeviceGraph& operator=(ResidentDeviceGraph&& other) noexcept;

  // Immutable after the first upload attempt (a second call throws
  // std::logic_error).  The upload completes before the call returns, so the
  // plan may be released at once.
  void upload_once_async(ResidentGraphPlan plan, const HipStream& stream);
  void wait_until_ready(const HipStream& stream) const;
  void synchronize_upload();

  [[nodiscard]] bool has_upload() const noexcept { return graph_ != nullptr; }
  [[nodiscard]] bool upload_complete() const noexcept { return graph_ != nullptr; }
  [[nodiscard]] const DeviceGraphMemoryReport& memory_report() const;
  [[nodiscard]] const DeviceGraphFingerprint& fingerprint() const;
  [[nodiscard]] std::uint32_t vertex_count() const noexcept { return vertex_count_; }
  [[nodiscard]] std::uint32_t tile_count() const noexcept { return tile_count_; }
  [[nodiscard]] bfn_graph handle() const;

  // Negotiated congestion: every edge u->v gets base_cost[v] * (1 + 0.5 *
  // occupancy[v]) (occupancy empty = all zero).  The graph is mutable only
  // through this call, between batches.
  void update_node_costs(const std::vector<float>& base_cost,
                         const std::vector<std::int32_t>& occupancy) const;

 private:
  bfn_graph graph_{};
  std::uint32_t vertex_count_{};
  std::uint32_t tile_count_{};
  DeviceGraphMemoryReport memory_{};
  DeviceGraphFingerprint fingerprint_{};
};

// Device scratch of one stream: labels, work lists, control, path arenas.
// Capacities grow geometrically and are reused by later batches; a nonzero
// budget_bytes refuses batches whose estimate exceeds it.
class ReusableBatchedWorkspace final {
 public:
  explicit ReusableBatchedWorkspace(const HipStream& stream, std::uint64_t budget_bytes = 0U);
  ~ReusableBatchedWorkspace() noexcept;
  ReusableBatchedWorkspace(const ReusableBatchedWorkspace&) = delete;
  ReusableBatchedWorkspace& operator=(const ReusableBatchedWorkspace&) = delete;

  [[nodiscard]] bfn_workspace handle() const noexcept { return workspace_; }

 private:
  bfn_workspace workspace_{};
};

// The reference keeps one workspace type per engine and one for the
// single-query engines; here every engine shares the one layout.
using ReusableBatchedJacobiWorkspace = ReusableBatchedWorkspace;
using ReusableBatchedDenseWorkspace = ReusableBatchedWorkspace;
using ReusableBatchedFrontierWorkspace = ReusableBatchedWorkspace;
using ReusableDeviceWorkspace = ReusableBatchedWorkspace;

// Path reconstruction scratch lives in the batch workspace; this token keeps
// the reference's run_compact_paths signature.
class ReusableCompactPathWorkspace final {};

}  // namespace bfnew::hip
