// Host emulation of the HIP device model, for checking kernel index
// arithmetic on the CPU under AddressSanitizer / UBSan.
//
// Test infrastructure only: the kernels of the library are compiled with the
// host compiler against this header (in place of <hip/hip_runtime.h>) and
// executed one workgroup at a time.  Every GPU thread is a user-level fiber on
// one OS thread; __syncthreads, the wave barrier and every cross-lane
// operation (shuffles, readfirstlane) are collective points at which the
// scheduler switches fibers.  Lanes run in any order between collective
// points, so code that silently relies on lock-step execution between them is
// exposed as a data race in the results, and a collective reached by only
// part of a wave or workgroup is reported as divergence (abort).
//
// LDS arrays (__shared__) become static arrays of the exact declared size,
// so ASan reports accesses past them; private arrays live on the fiber stack
// (stack-buffer-overflow); device buffers are separate heap allocations.
// The runtime API subset the library's batch solver uses is implemented on
// host memory: launches run synchronously (or are recorded into a "graph"
// under stream capture and replayed by hipGraphLaunch).
#pragma once

#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <utility>
#include <vector>

// ---------------------------------------------------------------------------
// Qualifiers
// ---------------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __constant__
#define __shared__ static
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {

[[noreturn]] inline void fail(const char* what, int a = -1, int b = -1) {
  std::fprintf(stderr, "EMU ERROR: %s (%d, %d)\n", what, a, b);
  std::fflush(stderr);
  std::abort();
}

// ---------------------------------------------------------------------------
// Fibers: a minimal x86-64 SysV context switch (callee-saved registers,
// MXCSR and the x87 control word), no signal-mask syscalls.
// ---------------------------------------------------------------------------
extern "C" void emu_ctx_switch(void** save_sp, void* new_sp);
asm(R"(
.text
.globl emu_ctx_switch
.type emu_ctx_switch,@function
emu_ctx_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
.size emu_ctx_switch, .-emu_ctx_switch
)");

enum State { RUN = 0, WAIT_BLOCK = 1, WAIT_WAVE = 2, DONE = 3 };

struct Fiber {
  void* sp = nullptr;
  char* stack = nullptr;
  size_t stack_size = 0;
  void* fake_stack = nullptr;
  int tid = 0;
  int state = RUN;
};

struct Emu {
  std::vector<Fiber> fibers;
  int cur = -1;
  void* sched_sp = nullptr;
  const void* sched_bottom = nullptr;
  size_t sched_size = 0;
  dim3 block_idx{0, 0, 0}, block_dim{1, 1, 1}, grid_dim{1, 1, 1};
  std::function<void()> body;
  std::vector<uint64_t> xbuf;  // [waves][64] exchange slots
  std::vector<int> xtag;       // [waves] kind of the pending collective
  uint64_t switches = 0;
};

inline Emu& E() {
  static Emu e;
  return e;
}

// Frees the fiber stacks (called once the emulated launches are over).
inline void release_fibers() {
  for (auto& f : E().fibers) std::free(f.stack);
  E().fibers.clear();
}

constexpr size_t kStack = 256 * 1024;

inline void yield_to_scheduler() {
  Emu& e = E();
  Fiber& f = e.fibers[e.cur];
  __sanitizer_start_switch_fiber(&f.fake_stack, e.sched_bottom, e.sched_size);
  emu_ctx_switch(&f.sp, e.sched_sp);
  __sanitizer_finish_switch_fiber(f.fake_stack, &e.sched_bottom, &e.sched_size);
}

[[noreturn]] inline void trampoline() {
  Emu& e = E();
  __sanitizer_finish_switch_fiber(nullptr, &e.sched_bottom, &e.sched_size);
  e.body();
  Fiber& f = e.fibers[e.cur];
  f.state = DONE;
  __sanitizer_start_switch_fiber(nullptr, e.sched_bottom, e.sched_size);
  emu_ctx_switch(&f.sp, e.sched_sp);
  __builtin_unreachable();
}

inline void init_fiber(Fiber& f) {
  if (!f.stack) {
    f.stack_size = kStack;
    f.stack = static_cast<char*>(std::aligned_alloc(64, kStack));
  }
  uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kStack) & ~uintptr_t(15);
  uint64_t* s = reinterpret_cast<uint64_t*>(top);
  s[-1] = 0;                                             // padding
  s[-2] = reinterpret_cast<uint64_t>(&trampoline);       // return address (16-aligned slot)
  for (int r = 3; r <= 8; ++r) s[-r] = 0;                // rbp rbx r12 r13 r14 r15
  uint32_t csr[2];
  asm volatile("stmxcsr %0" : "=m"(csr[0]));
  asm volatile("fnstcw %0" : "=m"(csr[1]));
  std::memcpy(&s[-9], csr, 8);
  f.sp = &s[-9];
  f.state = RUN;
  f.fake_stack = nullptr;
}

// Runs body() once per thread of one workgroup, to completion.
inline void run_block(dim3 grid, dim3 blk, dim3 bidx, std::function<void()> body) {
  Emu& e = E();
  const int nt = (int)(blk.x * blk.y * blk.z);
  if (blk.y != 1 || blk.z != 1) fail("only 1-D workgroups are emulated");
  e.grid_dim = grid;
  e.block_dim = blk;
  e.block_idx = bidx;
  e.body = std::move(body);
  if ((int)e.fibers.size() < nt) e.fibers.resize(nt);
  const int nw = (nt + 63) / 64;
  e.xbuf.assign((size_t)nw * 64, 0);
  e.xtag.assign(nw, 0);
  for (int t = 0; t < nt; ++t) {
    e.fibers[t].tid = t;
    init_fiber(e.fibers[t]);
  }
  int done = 0;
  while (done < nt) {
    bool ran = false;
    for (int t = 0; t < nt; ++t) {
      Fiber& f = e.fibers[t];
      if (f.state != RUN) continue;
      e.cur = t;
      void* fake = nullptr;
      __sanitizer_start_switch_fiber(&fake, f.stack, f.stack_size);
      emu_ctx_switch(&e.sched_sp, f.sp);
      __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
      ++e.switches;
      ran = true;
      if (f.state == DONE) ++done;
    }
    if (done == nt) break;
    if (ran) continue;
    // nobody runnable: release a wave whose live lanes all wait at a wave
    // collective, else the workgroup if every live thread waits at the barrier
    bool released = false;
    for (int w = 0; w < nw; ++w) {
      int live = 0, waiting = 0;
      for (int l = 0; l < 64 && w * 64 + l < nt; ++l) {
        const int s = e.fibers[w * 64 + l].state;
        if (s != DONE) ++live;
        if (s == WAIT_WAVE) ++waiting;
      }
      if (waiting > 0 && waiting == live) {
        for (int l = 0; l < 64 && w * 64 + l < nt; ++l)
          if (e.fibers[w * 64 + l].state == WAIT_WAVE) e.fibers[w * 64 + l].state = RUN;
        released = true;
      }
    }
    if (released) continue;
    int live = 0, at_bar = 0;
    for (int t = 0; t < nt; ++t) {
      if (e.fibers[t].state != DONE) ++live;
      if (e.fibers[t].state == WAIT_BLOCK) ++at_bar;
    }
    if (at_bar == live) {
      for (int t = 0; t < nt; ++t)
        if (e.fibers[t].state == WAIT_BLOCK) e.fibers[t].state = RUN;
      continue;
    }
    fail("divergent barrier: threads wait at different collective points", live, at_bar);
  }
  e.cur = -1;
}

inline int tid() {
  const int c = E().cur;
  if (c < 0) fail("device code outside a launch");
  return E().fibers[c].tid;
}

inline void block_barrier() {
  Emu& e = E();
  e.fibers[e.cur].state = WAIT_BLOCK;
  yield_to_scheduler();
}

inline void wave_wait() {
  Emu& e = E();
  e.fibers[e.cur].state = WAIT_WAVE;
  yield_to_scheduler();
}

// Cross-lane exchange: every live lane of the wave posts its value, then
// reads the slot of lane src(lane); reading a lane that has exited is an error.
template <class T, class F>
inline T exchange(T v, F src_of) {
  static_assert(sizeof(T) <= 8, "exchange of values up to 8 bytes");
  Emu& e = E();
  const int t = tid(), w = t / 64, l = t % 64;
  uint64_t bits = 0;
  std::memcpy(&bits, &v, sizeof(T));
  e.xbuf[(size_t)w * 64 + l] = bits;
  wave_wait();
  const int s = src_of(l);
  T r = v;
  if (s >= 0 && s < 64) {
    const int st = w * 64 + s;
    if (st >= (int)(e.block_dim.x) || e.fibers[st].state == DONE) fail("cross-lane read of an exited lane", t, st);
    std::memcpy(&r, &e.xbuf[(size_t)w * 64 + s], sizeof(T));
  }
  wave_wait();  // the slots may be reused only after every lane has read
  return r;
}

}  // namespace emu

// ---------------------------------------------------------------------------
// Built-in variables and device functions
// ---------------------------------------------------------------------------
namespace emu {
inline dim3 thread_idx() { return dim3((unsigned)tid(), 0, 0); }
}  // namespace emu
#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::E().block_idx)
#define blockDim (::emu::E().block_dim)
#define gridDim (::emu::E().grid_dim)

inline void __syncthreads() { emu::block_barrier(); }
#define __builtin_amdgcn_wave_barrier() ::emu::wave_wait()
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_s_barrier() ::emu::block_barrier()

template <class T>
inline T emu_readfirstlane(T v) {
  // every lane of the wave must take part with the same value: a value
  // that differs between lanes means the code's uniformity assumption fails
  const T first = emu::exchange(v, [](int) { return 0; });
  if (std::memcmp(&first, &v, sizeof(T)) != 0) emu::fail("readfirstlane of a non-uniform value", emu::tid());
  return first;
}
#define __builtin_amdgcn_readfirstlane(v) emu_readfirstlane(v)

template <class T>
inline T emu_readlane(T v, int l) {
  // the lane index must be the same on every lane and name a lane of the wave
  const int l0 = emu::exchange(l, [](int) { return 0; });
  if (l0 != l) emu::fail("readlane with a non-uniform lane index", emu::tid());
  if (l < 0 || l >= 64) emu::fail("readlane lane index out of range", emu::tid());
  return emu::exchange(v, [=](int) { return l; });
}
#define __builtin_amdgcn_readlane(v, l) emu_readlane(v, l)

template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  return emu::exchange(v, [=](int l) { return (l & ~(width - 1)) | ((l ^ mask) & (width - 1)); });
}
template <class T>
inline T __shfl_down(T v, unsigned off, int width = 64) {
  return emu::exchange(v, [=](int l) {
    const int s = (l & (width - 1)) + (int)off;
    return s < width ? (l & ~(width - 1)) + s : l;
  });
}
template <class T>
inline T __shfl(T v, int src, int width = 64) {
  return emu::exchange(v, [=](int l) { return (l & ~(width - 1)) + (src & (width - 1)); });
}

// FP64 matrix-core product D = A B + C over one wave, 16 x 4 times 4 x 16:
// lane l holds A[l % 16][l / 16] and B[l / 16][l % 16], and its four results
// are D[l / 16 + 4 r][l % 16] (the layout the library's kernels rely on).
template <class V>
inline V emu_mfma_f64_16x16x4(double a, double b, V c, int cbsz, int abid, int blgp) {
  if (cbsz || abid || blgp) emu::fail("mfma modifiers are not emulated");
  emu::Emu& e = emu::E();
  const int t = emu::tid(), w = t / 64, l = t % 64;
  static std::vector<double> xa, xb;
  const size_t nw = e.xbuf.size() / 64;
  if (xa.size() < nw * 64) {
    xa.resize(nw * 64);
    xb.resize(nw * 64);
  }
  xa[(size_t)w * 64 + l] = a;
  xb[(size_t)w * 64 + l] = b;
  emu::wave_wait();
  for (int s = 0; s < 64; ++s)
    if (w * 64 + s >= (int)e.block_dim.x || e.fibers[w * 64 + s].state == emu::DONE)
      emu::fail("mfma with exited lanes", t, w * 64 + s);
  V d = c;
  const int j = l % 16;
  for (int r = 0; r < 4; ++r) {
    const int i = l / 16 + 4 * r;
    double acc = c[r];
    for (int k = 0; k < 4; ++k) acc += xa[(size_t)w * 64 + i + 16 * k] * xb[(size_t)w * 64 + j + 16 * k];
    d[r] = acc;
  }
  emu::wave_wait();
  return d;
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu_mfma_f64_16x16x4(a, b, c, x, y, z)

// Integer min/max overloads of the device headers.
inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }

template <class T>
inline T atomicAdd(T* p, T v) {
  const T old = *p;
  *p = old + v;
  return old;
}

// ---------------------------------------------------------------------------
// Runtime API subset
// ---------------------------------------------------------------------------
enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotSupported = 801 };
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 63, hipDeviceAttributeMaxSharedMemoryPerBlock = 74 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
constexpr unsigned hipStreamNonBlocking = 1;

struct emu_stream {};
struct emu_event {
  std::chrono::steady_clock::time_point t;
};
struct emu_graph {
  std::vector<std::function<void()>> nodes;
};
using hipStream_t = emu_stream*;
using hipEvent_t = emu_event*;
using hipGraph_t = emu_graph*;
using hipGraphExec_t = emu_graph*;

namespace emu {
inline emu_graph*& capturing() {
  static emu_graph* g = nullptr;
  return g;
}
inline int& device_cus() {
  static int n = 4;
  return n;
}
}  // namespace emu

inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  *v = a == hipDeviceAttributeMultiprocessorCount ? emu::device_cus() : 65536;
  return hipSuccess;
}
// Free device memory as the emulated device reports it: EMU_FREE_BYTES, or 8 GiB.
inline hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
  const char* v = std::getenv("EMU_FREE_BYTES");
  *free_b = v && *v ? (size_t)std::strtoull(v, nullptr, 10) : ((size_t)8 << 30);
  *total_b = (size_t)16 << 30;
  return hipSuccess;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) {
  // device memory starts as NaN bit patterns (0xFF bytes), so reads of
  // never-written entries show up in the results
  void* q = std::malloc(n ? n : 1);
  if (!q) return hipErrorOutOfMemory;
  std::memset(q, 0xFF, n);
  *p = static_cast<T*>(q);
  return hipSuccess;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr) {
  if (n) std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) { return hipMemcpyAsync(d, s, n, k); }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) {
  std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new emu_stream;
  return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new emu_event;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  emu::capturing() = new emu_graph;
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = emu::capturing();
  emu::capturing() = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t) {
  *x = new emu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
  for (auto& n : x->nodes) n();
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t g) {
  delete g;
  return hipSuccess;
}

namespace emu {
// Runs (or records, under capture) a launch: workgroups one after another.
template <class... P, class... A>
inline void launch(void (*k)(P...), dim3 grid, dim3 blk, size_t shmem, hipStream_t, A&&... args) {
  if (shmem != 0) fail("dynamic LDS is not emulated", (int)shmem);
  std::tuple<std::decay_t<P>...> a(std::forward<A>(args)...);
  auto run = [k, grid, blk, a]() {
    for (unsigned b = 0; b < grid.x * grid.y * grid.z; ++b) {
      const dim3 bi(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
      run_block(grid, blk, bi, [&]() { std::apply(k, a); });
    }
  };
  if (capturing())
    capturing()->nodes.push_back(run);
  else
    run();
}
}  // namespace emu

#define hipLaunchKernelGGL(k, g, b, shm, s, ...) ::emu::launch(k, dim3(g), dim3(b), (size_t)(shm), s, __VA_ARGS__)
