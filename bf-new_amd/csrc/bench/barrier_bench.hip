// barrier_bench.hip -- cost of one grid-wide barrier on this device, for the
// persistent relaxation kernel's control loop (the reference measures
// cooperative_groups grid.sync the same way, tools/gpu_barrier_benchmark.hip.cpp).
//
// For workgroups of 256 threads at 1, 2, 4, 5 and the occupancy limit per
// CU, one cooperative launch runs `rounds` barriers with a little arithmetic
// between them, timed with HIP events; reported per barrier:
//   flat        -- every workgroup arrives on one counter (grid_barrier_flat)
//   grouped_G   -- arrivals gathered on G group counters first
//                  (grid_barrier_grouped<G>, G = 8, 16, 32, 64)
//   cg          -- cooperative_groups::this_grid().sync()
// One JSON object per line on stdout.
#include "bfn_internal.hpp"

#include <hip/hip_cooperative_groups.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

using bfn::DevControl;

template <int Kind>
__global__ __launch_bounds__(256) void barrier_kernel(DevControl* c, std::uint32_t rounds,
                                                      unsigned long long* sink) {
  unsigned long long acc = blockIdx.x + 1ull;
  for (std::uint32_t r = 0; r < rounds; ++r) {
    acc = acc * 6364136223846793005ull + r + threadIdx.x;
    if constexpr (Kind == 0) {
      if (!bfn::grid_barrier_flat(c, gridDim.x)) break;
    } else if constexpr (Kind > 0) {
      if (!bfn::grid_barrier_grouped<static_cast<std::uint32_t>(Kind)>(c, gridDim.x)) break;
    } else {
      cooperative_groups::this_grid().sync();
    }
  }
  if (threadIdx.x == 0) sink[blockIdx.x] = acc;
}

#define CHECK_HIP(x)                                                                 \
  do {                                                                               \
    const hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

template <int Kind>
float run(DevControl* c, unsigned long long* sink, std::uint32_t grid, std::uint32_t rounds) {
  CHECK_HIP(hipMemset(c, 0, sizeof(DevControl)));
  void* args[] = {&c, &rounds, &sink};
  const void* k = reinterpret_cast<const void*>(&barrier_kernel<Kind>);
  std::uint32_t warm = 10;
  void* warm_args[] = {&c, &warm, &sink};
  CHECK_HIP(hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), warm_args, 0, nullptr));
  CHECK_HIP(hipDeviceSynchronize());
  hipEvent_t a, b;
  CHECK_HIP(hipEventCreate(&a));
  CHECK_HIP(hipEventCreate(&b));
  CHECK_HIP(hipEventRecord(a, nullptr));
  CHECK_HIP(hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), args, 0, nullptr));
  CHECK_HIP(hipEventRecord(b, nullptr));
  CHECK_HIP(hipEventSynchronize(b));
  float ms = 0;
  CHECK_HIP(hipEventElapsedTime(&ms, a, b));
  CHECK_HIP(hipEventDestroy(a));
  CHECK_HIP(hipEventDestroy(b));
  return ms;
}

}  // namespace

int main(int argc, char** argv) {
  const std::uint32_t rounds = argc > 1 ? static_cast<std::uint32_t>(std::atoi(argv[1])) : 1000u;
  hipDeviceProp_t p;
  CHECK_HIP(hipGetDeviceProperties(&p, 0));
  int occ = 0;
  CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &occ, reinterpret_cast<const void*>(&barrier_kernel<0>), 256, 0));
  DevControl* c = nullptr;
  unsigned long long* sink = nullptr;
  const std::uint32_t cus = static_cast<std::uint32_t>(p.multiProcessorCount);
  CHECK_HIP(hipMalloc(&c, sizeof(DevControl)));
  CHECK_HIP(hipMalloc(&sink, sizeof(unsigned long long) * cus * static_cast<std::uint32_t>(occ)));
  std::vector<std::uint32_t> per_cu{1, 2, 4, 5};
  if (occ > 5) per_cu.push_back(static_cast<std::uint32_t>(occ));
  for (const std::uint32_t k : per_cu) {
    if (k > static_cast<std::uint32_t>(occ)) continue;
    const std::uint32_t grid = k * cus;
    const float us = 1e3f / static_cast<float>(rounds);
    const float flat = run<0>(c, sink, grid, rounds) * us;
    const float g8 = run<8>(c, sink, grid, rounds) * us;
    const float g16 = run<16>(c, sink, grid, rounds) * us;
    const float g32 = run<32>(c, sink, grid, rounds) * us;
    const float g64 = run<64>(c, sink, grid, rounds) * us;
    const float cg = run<-1>(c, sink, grid, rounds) * us;
    std::printf("{\"workgroups\": %u, \"per_cu\": %u, \"rounds\": %u, \"us_per_barrier\": "
                "{\"flat\": %.3f, \"grouped_8\": %.3f, \"grouped_16\": %.3f, \"grouped_32\": %.3f, "
                "\"grouped_64\": %.3f, \"cg\": %.3f}}\n",
                grid, k, rounds, flat, g8, g16, g32, g64, cg);
    std::fflush(stdout);
  }
  CHECK_HIP(hipFree(sink));
  CHECK_HIP(hipFree(c));
  return 0;
}
