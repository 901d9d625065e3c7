// Host stand-in for the HIP device API that emd-cuda_amd/csrc/emd_sift_dev.h uses, so the
// sift kernels' own source compiles with g++ and runs on the CPU (tests/native/
// sift_emu.cpp).  Test infrastructure only: found through -I tests/native/emu in place of
// ROCm's header, never by a GPU build.
//
// Execution model: a workgroup runs on one OS thread, its lanes as fibers (own stacks,
// switched cooperatively by a small x86-64 context switch, annotated for
// AddressSanitizer).  A lane runs until it reaches a barrier: __syncthreads waits for every
// lane of the workgroup, and every wave-level operation (ballot, compare masks, shuffles,
// DPP moves, readlane, wave barriers) is an exchange among the wave's 64 lanes through a
// slot array between two 64-lane barriers.  That assumes what the kernels guarantee: all
// lanes of a workgroup reach the same barriers and all 64 lanes of a wave the same wave
// operations.  The scheduler checks it: a barrier completed by lanes standing at different
// source lines, or a point where no lane can run while some have not finished (a barrier
// never completed), is reported with every lane's position and the run exits with
// status 3.  Workgroups of a persistent launch run on concurrent OS threads and share the
// queue counter, as on the GPU.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#endif

typedef struct emu_stream* hipStream_t;
typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
};

// emu_switch(from, to): save the callee-saved registers on the current stack, store its
// pointer in *from, load *to and restore the registers saved there (System V x86-64).
extern "C" void emu_switch(void** from, void* to);
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
.size emu_switch, .-emu_switch
)");

namespace emu {

constexpr uint32_t kPoison32 = 0xDEADBEEFu;  // lanes a DPP move leaves undefined
constexpr size_t kStack = 1 << 20;           // bytes of stack per lane

struct Group;

struct Lane {
  dim3 tid, bid, gdim, bdim;
  int lane = 0, wave = 0;
  Group* g = nullptr;
  // fiber
  void* sp = nullptr;
  std::vector<char> stack;
  std::function<void()> body;
  bool done = false;
  int waiting = -1;       // barrier id waited on (-1: runnable), gen at arrival below
  uint64_t wait_gen = 0;
  const char* site = "start";
  int line = 0;
};

struct Barrier {
  int arrived = 0;
  uint64_t gen = 0;
  const char* site = nullptr;  // source position of the first lane to arrive
  int line = 0;
};

struct Group {
  explicit Group(int wg) : lanes(wg), bars(1 + wg / 64), slots((wg / 64) * 64) {}
  std::vector<Lane> lanes;
  std::vector<Barrier> bars;  // [0] the workgroup, [1 + w] wave w
  std::vector<uint64_t> slots;
  void* sched_sp = nullptr;
#if defined(__SANITIZE_ADDRESS__)
  const void* sched_stack_bottom = nullptr;
  size_t sched_stack_size = 0;
#endif
};

inline thread_local Lane* cur = nullptr;

[[noreturn]] inline void report(Group* g, const char* what) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  std::fprintf(stderr, "EMU_DEADLOCK: %s (workgroup %u); position of each lane:\n", what,
               g->lanes[0].bid.x);
  for (size_t i = 0; i < g->lanes.size(); ++i) {
    const Lane& L = g->lanes[i];
    if (i == 0 || L.line != g->lanes[i - 1].line || L.site != g->lanes[i - 1].site ||
        L.done != g->lanes[i - 1].done)
      std::fprintf(stderr, "  lane %zu%s: %s:%d\n", i, L.done ? " (finished)" : "", L.site,
                   L.line);
  }
  std::fflush(stderr);
  std::_Exit(3);
}

// Back to the workgroup's scheduler; returns when the scheduler resumes this lane.
inline void yield() {
  Lane* L = cur;
  Group* g = L->g;
#if defined(__SANITIZE_ADDRESS__)
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, g->sched_stack_bottom, g->sched_stack_size);
#endif
  emu_switch(&L->sp, g->sched_sp);
#if defined(__SANITIZE_ADDRESS__)
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
  cur = L;
}

// Barrier `id` of the lane's workgroup (0) or wave (1 + wave) over `count` lanes.
inline void barrier(int id, int count, const char* file, int line, const char* what) {
  Lane* L = cur;
  Group* g = L->g;
  L->site = file;
  L->line = line;
  Barrier& b = g->bars[id];
  if (b.arrived == 0) {
    b.site = file;
    b.line = line;
  } else if (b.site != file || b.line != line) {
    char msg[256];
    std::snprintf(msg, sizeof msg, "%s reached at different source lines by one %s", what,
                  id == 0 ? "workgroup" : "wave");
    report(g, msg);
  }
  if (++b.arrived == count) {
    b.arrived = 0;
    ++b.gen;
    return;
  }
  L->waiting = id;
  L->wait_gen = b.gen;
  yield();
}

inline void note(const char* file, int line) {
  cur->site = file;
  cur->line = line;
}

inline void syncthreads(const char* file, int line) {
  barrier(0, (int)cur->g->lanes.size(), file, line, "__syncthreads");
}

inline void fiber_entry() {
  Lane* L = cur;
#if defined(__SANITIZE_ADDRESS__)
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
  L->body();
  L->done = true;
  L->site = "exit";
  L->line = 0;
  Group* g = L->g;
#if defined(__SANITIZE_ADDRESS__)
  __sanitizer_start_switch_fiber(nullptr, g->sched_stack_bottom, g->sched_stack_size);
#endif
  void* dead = nullptr;
  emu_switch(&dead, g->sched_sp);
  std::abort();  // a finished lane is never resumed
}

// Runs every lane of g to completion (body(lane) on each), on the calling thread.
inline void run_group(Group& g, const std::function<void(Lane&)>& body) {
  for (auto& L : g.lanes) {
    L.g = &g;
    L.stack.resize(kStack);
    L.body = [&body, &L] { body(L); };
    // initial frame: six register slots, then fiber_entry as the return address, then a
    // pad so that fiber_entry starts with the stack aligned as after a call
    uintptr_t top = (reinterpret_cast<uintptr_t>(L.stack.data()) + kStack) & ~uintptr_t(15);
    void** sp = reinterpret_cast<void**>(top);
    *--sp = nullptr;
    *--sp = reinterpret_cast<void*>(&fiber_entry);
    for (int r = 0; r < 6; ++r) *--sp = nullptr;
    L.sp = sp;
  }
#if defined(__SANITIZE_ADDRESS__)
  {
    // the scheduler runs on this thread's own stack
    pthread_attr_t attr;
    pthread_getattr_np(pthread_self(), &attr);
    void* addr = nullptr;
    size_t size = 0;
    pthread_attr_getstack(&attr, &addr, &size);
    pthread_attr_destroy(&attr);
    g.sched_stack_bottom = addr;
    g.sched_stack_size = size;
  }
#endif
  const int n = (int)g.lanes.size();
  int finished = 0;
  while (finished < n) {
    bool ran = false;
    for (auto& L : g.lanes) {
      if (L.done) continue;
      if (L.waiting >= 0) {
        if (g.bars[L.waiting].gen == L.wait_gen) continue;  // barrier not yet complete
        L.waiting = -1;
      }
      cur = &L;
#if defined(__SANITIZE_ADDRESS__)
      void* fake = nullptr;
      __sanitizer_start_switch_fiber(&fake, L.stack.data(), kStack);
#endif
      emu_switch(&g.sched_sp, L.sp);
#if defined(__SANITIZE_ADDRESS__)
      __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
      ran = true;
      if (L.done) ++finished;
    }
    if (!ran && finished < n) report(&g, "no lane can run: a barrier is never completed");
  }
  cur = nullptr;
}

// One wave-level exchange: every lane publishes v, then f(slots, lane) gives its result.
template <class F>
inline uint64_t exchange(uint64_t v, F&& f, const char* what) {
  Lane* L = cur;
  uint64_t* slot = &L->g->slots[L->wave * 64];
  slot[L->lane] = v;
  barrier(1 + L->wave, 64, what, 0, what);
  const uint64_t r = f(slot, L->lane);
  barrier(1 + L->wave, 64, what, 1, what);
  return r;
}

inline uint64_t ballot(bool p) {
  return exchange(p ? 1 : 0, [](const uint64_t* s, int) {
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i) m |= (s[i] & 1) << i;
    return m;
  }, "ballot");
}

template <class T>
inline uint64_t bits(T v) {
  static_assert(sizeof(T) <= 8, "64-bit values at most");
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T from_bits(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}

template <class T>
inline T shfl(T v, int src) {
  return from_bits<T>(exchange(bits(v), [src](const uint64_t* s, int) { return s[src & 63]; },
                               "__shfl"));
}
template <class T>
inline T shfl_up(T v, unsigned d) {
  return from_bits<T>(exchange(bits(v), [d](const uint64_t* s, int l) {
    return (int)(l - (int)d) >= 0 ? s[l - d] : s[l];
  }, "__shfl_up"));
}
template <class T>
inline T shfl_down(T v, unsigned d) {
  return from_bits<T>(exchange(bits(v), [d](const uint64_t* s, int l) {
    return l + (int)d < 64 ? s[l + d] : s[l];
  }, "__shfl_down"));
}

template <class T>
inline T shfl_xor(T v, int m) {
  return from_bits<T>(exchange(bits(v), [m](const uint64_t* s, int l) { return s[(l ^ m) & 63]; },
                               "__shfl_xor"));
}

inline int readlane(int v, int l) {
  return (int)(uint32_t)exchange((uint32_t)v, [l](const uint64_t* s, int) { return s[l & 63]; },
                                 "readlane");
}
inline int readfirstlane(int v) { return readlane(v, 0); }

// Source lane of a DPP control for lane l; false when the source does not exist.
inline bool dpp_source(int ctrl, int l, int* src) {
  const int row = l >> 4, r = l & 15;
  if (ctrl >= 0 && ctrl <= 0xFF) {  // quad_perm
    *src = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
    return true;
  }
  if (ctrl >= 0x101 && ctrl <= 0x10F) {  // row_shl:n, lane i <- lane i+n of its row
    const int s = r + (ctrl & 15);
    *src = row * 16 + s;
    return s < 16;
  }
  if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr:n, lane i <- lane i-n of its row
    const int s = r - (ctrl & 15);
    *src = row * 16 + s;
    return s >= 0;
  }
  if (ctrl >= 0x121 && ctrl <= 0x12F) {  // row_ror:n
    *src = row * 16 + ((r - (ctrl & 15)) & 15);
    return true;
  }
  switch (ctrl) {
    case 0x130: *src = l + 1; return l + 1 < 64;        // wave_shl:1
    case 0x134: *src = (l + 1) & 63; return true;       // wave_rol:1
    case 0x138: *src = l - 1; return l >= 1;            // wave_shr:1
    case 0x13C: *src = (l - 1) & 63; return true;       // wave_ror:1
    case 0x140: *src = row * 16 + (15 - r); return true;                   // row_mirror
    case 0x141: *src = row * 16 + (r & 8) + (7 - (r & 7)); return true;    // row_half_mirror
    case 0x142: *src = row * 16 - 1; return row >= 1;   // row_bcast:15
    case 0x143: *src = 31; return row >= 2;             // row_bcast:31
  }
  std::fprintf(stderr, "EMU: DPP control 0x%x not emulated\n", ctrl);
  std::abort();
}

// update_dpp(old, src, ...): lanes of disabled rows or banks keep old; a lane whose source
// does not exist reads 0 with bound_ctrl, else keeps old.  mov_dpp is update_dpp with an
// undefined old value (a poison pattern here).
inline int update_dpp(int old, int v, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  return (int)(uint32_t)exchange((uint32_t)v, [=](const uint64_t* s, int l) -> uint64_t {
    if (!((row_mask >> (l >> 4)) & 1) || !((bank_mask >> ((l >> 2) & 3)) & 1))
      return (uint32_t)old;
    int src = 0;
    if (!dpp_source(ctrl, l, &src)) return bound_ctrl ? 0u : (uint32_t)old;
    return s[src] & 0xFFFFFFFFu;
  }, "dpp");
}

inline bool fcmp_pred(double a, double b, int pred) {
  const bool un = std::isnan(a) || std::isnan(b);
  switch (pred) {
    case 0: return false;
    case 1: return !un && a == b;
    case 2: return !un && a > b;
    case 3: return !un && a >= b;
    case 4: return !un && a < b;
    case 5: return !un && a <= b;
    case 6: return !un && a != b;
    case 7: return !un;
    case 8: return un;
    case 9: return un || a == b;
    case 10: return un || a > b;
    case 11: return un || a >= b;
    case 12: return un || a < b;
    case 13: return un || a <= b;
    case 14: return un || a != b;
    case 15: return true;
  }
  std::abort();
}
inline bool icmp_pred(int a, int b, int pred) {
  switch (pred) {
    case 32: return a == b;
    case 33: return a != b;
    case 34: return (unsigned)a > (unsigned)b;
    case 35: return (unsigned)a >= (unsigned)b;
    case 36: return (unsigned)a < (unsigned)b;
    case 37: return (unsigned)a <= (unsigned)b;
    case 38: return a > b;
    case 39: return a >= b;
    case 40: return a < b;
    case 41: return a <= b;
  }
  std::abort();
}

inline unsigned mbcnt_lo(unsigned m, unsigned acc) {
  const int l = cur->lane;
  const unsigned below = l >= 32 ? 0xFFFFFFFFu : ((1u << l) - 1u);
  return acc + (unsigned)__builtin_popcount(m & below);
}
inline unsigned mbcnt_hi(unsigned m, unsigned acc) {
  const int l = cur->lane;
  const unsigned below = l < 32 ? 0u : ((l == 63 ? 0x80000000u : (1u << (l - 32))) - 1u);
  return acc + (unsigned)__builtin_popcount(m & below);
}

inline void wave_barrier(const char* file, int line) {
  barrier(1 + cur->wave, 64, file, line, "wave barrier");
}

}  // namespace emu

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
// a kernel's __shared__ variables: one instance per OS thread, i.e. per workgroup
#define __shared__ thread_local

#define threadIdx (::emu::cur->tid)
#define blockIdx (::emu::cur->bid)
#define gridDim (::emu::cur->gdim)
#define blockDim (::emu::cur->bdim)

#define __syncthreads() ::emu::syncthreads(__FILE__, __LINE__)
#define __ballot(p) ::emu::ballot((p) != 0)
// (an optional width argument must be the wave size, 64)
#define __shfl(v, s, ...) ::emu::shfl((v), (s))
#define __shfl_up(v, d, ...) ::emu::shfl_up((v), (d))
#define __shfl_down(v, d, ...) ::emu::shfl_down((v), (d))
#define __shfl_xor(v, m, ...) ::emu::shfl_xor((v), (m))
#define __popcll(v) __builtin_popcountll((unsigned long long)(v))
#define __builtin_amdgcn_readfirstlane(v) ::emu::readfirstlane((int)(v))
#define __builtin_amdgcn_readlane(v, l) ::emu::readlane((int)(v), (int)(l))
#define __builtin_amdgcn_mov_dpp(v, ctrl, rm, bm, bc) \
  ::emu::update_dpp((int)::emu::kPoison32, (int)(v), (ctrl), (rm), (bm), (bc))
#define __builtin_amdgcn_update_dpp(old, v, ctrl, rm, bm, bc) \
  ::emu::update_dpp((int)(old), (int)(v), (ctrl), (rm), (bm), (bc))
#define __builtin_amdgcn_fcmp(a, b, pred) ::emu::ballot(::emu::fcmp_pred((a), (b), (pred)))
#define __builtin_amdgcn_sicmp(a, b, pred) ::emu::ballot(::emu::icmp_pred((a), (b), (pred)))
#define __builtin_amdgcn_mbcnt_lo(m, acc) ::emu::mbcnt_lo((m), (acc))
#define __builtin_amdgcn_mbcnt_hi(m, acc) ::emu::mbcnt_hi((m), (acc))
#define __builtin_amdgcn_wave_barrier() ::emu::wave_barrier(__FILE__, __LINE__)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_s_setprio(n) ((void)0)
#define __builtin_amdgcn_rcp(x) (1.0 / (x))

inline long long __double_as_longlong(double d) { return ::emu::from_bits<long long>(::emu::bits(d)); }
inline double __longlong_as_double(long long v) { return ::emu::from_bits<double>(::emu::bits(v)); }

inline unsigned atomicAdd(unsigned* p, unsigned v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}

template <class A, class B>
inline std::common_type_t<A, B> min(A a, B b) {
  return a < b ? a : b;
}
template <class A, class B>
inline std::common_type_t<A, B> max(A a, B b) {
  return a < b ? b : a;
}

// The kernels' empty `asm volatile("" : "+v"(x))` (an optimisation barrier on the GPU)
// becomes a call of an empty function.
inline void emu_asm_stmt() {}
#define asm emu_asm_stmt
#define volatile(...) ()
