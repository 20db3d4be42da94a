// emu_rt.cpp -- CPU emulation of the HIP runtime subset that
// bf-new_amd/csrc/hip uses (test infrastructure: tests/test_emulated_kernels.py).
//
// Every lane of a workgroup is a fiber on one host thread; lanes switch at
// barriers, wave exchanges, sleeps and (EMU_YIELD_PCT percent of the time)
// at atomics, so lanes of concurrent workgroups interleave.  Ordinary
// launches run their workgroups one after another; cooperative (persistent)
// launches run all workgroups together.  Streams are synchronous, hipGraph
// capture records the launches and copies and replays them.
//
// Environment: EMU_CUS (compute units, default 2), EMU_OCC (resident
// workgroups per unit, default 2), EMU_YIELD_PCT (default 0), EMU_STACK_KB
// (fiber stack, default 256), EMU_TRACE (print every launch).
#include "hip/hip_runtime.h"

struct emu_event { std::chrono::steady_clock::time_point t; };
struct emu_graph { std::vector<std::function<void()>> ops; };

namespace emu {
dim3 g_tid, g_bid, g_bdim, g_gdim;
Block* g_blk = nullptr;

namespace {
bool g_capturing = false;
emu_graph* g_capture = nullptr;

// Minimal x86-64 context switch: callee-saved registers on the stack.
extern "C" void emu_switch(void** save_sp, void* new_sp);
asm(R"(
.text
.globl emu_switch
.type emu_switch,@function
emu_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
)");

const std::size_t kStack = std::getenv("EMU_STACK_KB") ? std::atoi(std::getenv("EMU_STACK_KB")) * 1024ull : 256 * 1024;

struct Lane {
  void* sp = nullptr;
  bool done = false;
  dim3 tid, bid;
  Block* blk = nullptr;
  char* stack = nullptr;
};

struct Sched {
  void* main_sp = nullptr;
  std::vector<Lane> lanes;
  std::size_t cur = 0;
  const std::function<void()>* body = nullptr;
  dim3 grid, block;
  std::vector<char*> stacks;
  std::uint64_t switches = 0;
};
Sched g_s;

void trampoline() {
  Lane& me = g_s.lanes[g_s.cur];
  (*g_s.body)();
  // lane exit: leave the block's and the wave's barrier counts
  Block* b = me.blk;
  const std::uint32_t w = me.tid.x / 64;
  --b->live;
  if (b->count && b->count == b->live) { b->count = 0; ++b->gen; }
  --b->wave_live[w];
  if (b->wave_count[w] && b->wave_count[w] == b->wave_live[w]) { b->wave_count[w] = 0; ++b->wave_gen[w]; }
  me.done = true;
  void* dummy;
  emu_switch(&dummy, g_s.main_sp);
  std::abort();
}

void prepare(Lane& l) {
  char* top = l.stack + kStack;
  top = reinterpret_cast<char*>(reinterpret_cast<std::uintptr_t>(top) & ~std::uintptr_t{15});
  void** sp = reinterpret_cast<void**>(top);
  *--sp = nullptr;                                   // fake return address of trampoline
  *--sp = reinterpret_cast<void*>(&trampoline);      // popped by ret
  for (int i = 0; i < 6; ++i) *--sp = nullptr;       // rbp rbx r12 r13 r14 r15
  l.sp = sp;
}

void run_lanes() {
  std::size_t live = g_s.lanes.size();
  while (live) {
    live = 0;
    for (std::size_t i = 0; i < g_s.lanes.size(); ++i) {
      Lane& l = g_s.lanes[i];
      if (l.done) continue;
      g_s.cur = i;
      g_tid = l.tid;
      g_bid = l.bid;
      g_blk = l.blk;
      emu_switch(&g_s.main_sp, l.sp);
      if (!l.done) ++live;
      if (++g_s.switches > 20000000000ull) {
        std::fprintf(stderr, "[emu] no progress: giving up\n");
        std::abort();
      }
    }
  }
}

void setup_block(Block& b, dim3 block, std::size_t lds) {
  const std::uint32_t B = block.x;
  b.live = B;
  b.lds.assign(lds / 4, 0xA5A5A5A5u);  // garbage, as LDS starts
  for (std::uint32_t w = 0; w * 64 < B; ++w) b.wave_live[w] = std::min<std::uint32_t>(64, B - w * 64);
}

char* stack_for(std::size_t i) {
  while (g_s.stacks.size() <= i) g_s.stacks.push_back(static_cast<char*>(std::malloc(kStack)));
  return g_s.stacks[i];
}
}  // namespace

void maybe_yield() {
  static const int pct = std::getenv("EMU_YIELD_PCT") ? std::atoi(std::getenv("EMU_YIELD_PCT")) : 0;
  static std::uint64_t x = 88172645463325252ull;
  if (!pct || g_s.lanes.empty()) return;
  x ^= x << 13; x ^= x >> 7; x ^= x << 17;
  if (static_cast<int>(x % 100) < pct) yield();
}

void yield() {
  Lane& me = g_s.lanes[g_s.cur];
  emu_switch(&me.sp, g_s.main_sp);
}

void barrier() {
  Block* b = g_blk;
  const std::uint64_t gen = b->gen;
  if (++b->count == b->live) {
    b->count = 0;
    ++b->gen;
    return;
  }
  while (b->gen == gen) yield();
}

void wave_barrier() {
  Block* b = g_blk;
  const std::uint32_t w = threadIdx.x / 64;
  const std::uint64_t gen = b->wave_gen[w];
  if (++b->wave_count[w] == b->wave_live[w]) {
    b->wave_count[w] = 0;
    ++b->wave_gen[w];
    return;
  }
  while (b->wave_gen[w] == gen) yield();
}

bool capturing() { return g_capturing; }
void capture(std::function<void()> f) { g_capture->ops.push_back(std::move(f)); }
void begin_capture() {
  g_capture = new emu_graph;
  g_capturing = true;
}
emu_graph* end_capture() {
  g_capturing = false;
  emu_graph* g = g_capture;
  g_capture = nullptr;
  return g;
}

void launch(dim3 grid, dim3 block, std::size_t lds, bool concurrent, std::function<void()> body) {
  if (g_capturing) {
    g_capture->ops.push_back([=] { launch(grid, block, lds, concurrent, body); });
    return;
  }
  static const bool trace = std::getenv("EMU_TRACE") != nullptr;
  static std::uint64_t launches = 0;
  ++launches;
  if (trace)
    std::fprintf(stderr, "[emu] launch %llu grid %u block %u lds %zu %s %s\n",
                 static_cast<unsigned long long>(launches), grid.x, block.x, lds,
                 concurrent ? "concurrent" : "sequential", body.target_type().name());
  g_s.body = &body;
  g_s.grid = grid;
  g_s.block = block;
  g_gdim = grid;
  g_bdim = block;
  const std::uint32_t per_run = concurrent ? grid.x : 1;
  for (std::uint32_t first = 0; first < grid.x; first += per_run) {
    std::vector<std::unique_ptr<Block>> blocks;
    g_s.lanes.clear();
    for (std::uint32_t x = first; x < first + per_run; ++x) {
      blocks.emplace_back(new Block);
      setup_block(*blocks.back(), block, lds);
      for (std::uint32_t i = 0; i < block.x; ++i) {
        Lane l;
        l.tid = dim3(i);
        l.bid = dim3(x);
        l.blk = blocks.back().get();
        l.stack = stack_for(g_s.lanes.size());
        prepare(l);
        g_s.lanes.push_back(l);
      }
    }
    run_lanes();
  }
}
}  // namespace emu

const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  std::memset(p, 0, sizeof *p);
  std::strcpy(p->name, "cpu emulation");
  std::strcpy(p->gcnArchName, "gfx950");
  p->totalGlobalMem = 16ull << 30;
  p->sharedMemPerBlock = 160 * 1024;
  p->maxSharedMemoryPerMultiProcessor = 160 * 1024;
  p->multiProcessorCount = std::getenv("EMU_CUS") ? std::atoi(std::getenv("EMU_CUS")) : 2;
  p->warpSize = 64;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipMalloc(void** p, std::size_t n) {
  *p = std::aligned_alloc(256, (std::max<std::size_t>(n, 1) + 255) / 256 * 256);
  if (*p) std::memset(*p, 0xCD, n);  // garbage, as device memory starts
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipHostMalloc(void** p, std::size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, std::size_t n, hipMemcpyKind, hipStream_t) {
  if (emu::capturing()) {
    emu::capture([=] { std::memcpy(d, s, n); });
    return hipSuccess;
  }
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, std::size_t n, hipStream_t) {
  if (emu::capturing()) {
    emu::capture([=] { std::memset(d, v, n); });
    return hipSuccess;
  }
  std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = reinterpret_cast<void*>(1); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  emu::begin_capture();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = emu::end_capture();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, std::size_t) {
  *e = g;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& op : e->ops) op();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }  // owned by the graph
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, std::size_t) {
  *n = std::getenv("EMU_OCC") ? std::atoi(std::getenv("EMU_OCC")) : 2;
  return hipSuccess;
}
