// emd_capi.hip -- C ABI of libemd_gfx950.so (declared in include/emd_gfx950.h).
#include "../../include/emd_gfx950.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "emd_sift.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace emdk {

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(EMD_ERR_RUNTIME, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace emdk

namespace {

using emdk::fail;
using emdk::hip_fail;

constexpr int kMaxPlainBlocks = 32768;  // blocks per plain sub-launch (bounds spill memory)

}  // namespace

struct emd_handle {
  emd_config cfg;
  int device = 0;
  int num_cus = 0;
  int lds_max = 0;
  char* spill = nullptr;       // per-workgroup global knot tiles
  size_t spill_size = 0;
  unsigned* counters = nullptr;  // [0] persistent-kernel queue head
  int timing = 0;                // record ev[0]/ev[1] around the sift launches
  int timed = 0;                 // the last run recorded them
  hipEvent_t ev[2] = {nullptr, nullptr};
  // The scratch above is shared by every launch of the handle.  `busy` is recorded after
  // each launch on its stream; a launch on another stream first waits for it, so launches
  // of one handle never overlap each other's queue counter or global knot tiles.
  hipEvent_t busy = nullptr;
  hipStream_t busy_stream = nullptr;
  int busy_set = 0;
  // this device's copy of the static scratch (emd_run_ragged_noalloc)
  char* static_scratch = nullptr;
};

// ---- Static scratch of emd_run_ragged_noalloc --------------------------------------------
// One array per GPU, held until the process ends: the queue counter in its first 256
// bytes, then the workgroups' global knot tiles.  Every call on a GPU
// shares it (calls that overlap on one GPU race); each call clears the counter first.
namespace {
constexpr size_t kStaticScratchBytes = (size_t)EMD_STATIC_SCRATCH_BYTES;
constexpr size_t kCounterBytes = 256;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

// Device address of the current device's copy of `scratch` (the array's name is only a
// host-side stand-in: kernels and copies must use this address).  Looked up once per
// device, with that device current.
static int static_scratch_address(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices)
    return fail(EMD_ERR_RUNTIME, "device %d outside the %d devices the library tracks", device,
                kMaxDevices);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(static scratch)");
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return EMD_OK;
}

namespace {

int ensure_spill(emd_handle* h, size_t bytes) {
  if (bytes <= h->spill_size) return EMD_OK;
  if (h->spill) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize before growing scratch");
    e = hipFree(h->spill);
    h->spill = nullptr;
    h->spill_size = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree(scratch)");
  }
  size_t sz = std::max(bytes, (size_t)1 << 20);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&h->spill), sz);
  if (e != hipSuccess) {
    h->spill = nullptr;
    return fail(EMD_ERR_MEMORY, "hipMalloc(%zu bytes of spill scratch): %s", sz,
                hipGetErrorString(e));
  }
  h->spill_size = sz;
  return EMD_OK;
}

// knot slots that can never be exceeded for traces of length <= n
int cap_slots(const emd_config& c, int n) {
  return 2 * std::min(c.max_extrema, std::max(n, 1)) + 4 * c.nbsym;
}

struct Shape {
  int wg, tile, spill_slots, max_n;
  int slot;  // bytes per knot slot: kSlotBytes, or kSlotBytesTime for the time-axis kernels
};

int check_wg(int wg) {
  if (wg != 64 && wg != 128 && wg != 256 && wg != 512 && wg != 1024)
    return fail(EMD_ERR_VALUE, "threads per workgroup must be 64, 128, 256, 512 or 1024, got %d",
                wg);
  return EMD_OK;
}

// LDS knot slots of slot_bytes that fit next to the work area of traces up to max_n samples
int lds_fit_slots(const emd_handle* h, int wg, int max_n, int slot_bytes = emdk::kSlotBytes) {
  const long avail =
      (long)h->lds_max - emdk::static_lds_bytes(wg) - (long)emdk::dyn_lds_bytes(0, max_n);
  return avail > 0 ? (int)(avail / slot_bytes) : 0;
}

// tile_req LDS knot slots (0 = the capacity rule), clamped to what LDS holds; a global
// tile of the full capacity backs sifts whose knots do not fit.  timed: the time-axis
// kernels (wider knot slots).
int make_shape(const emd_handle* h, int wg, int tile_req, int max_n, Shape* s,
               int timed = 0) {
  int rc = check_wg(wg);
  if (rc) return rc;
  if (max_n > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "trace length %d exceeds the supported maximum %d", max_n,
                emdk::kMaxN);
  if (emdk::dyn_lds_bytes(0, max_n) + emdk::static_lds_bytes(wg) > (size_t)h->lds_max)
    return fail(EMD_ERR_VALUE, "trace length %d does not fit the LDS work area", max_n);
  const int cap = cap_slots(h->cfg, max_n);
  const int slot = timed ? emdk::kSlotBytesTime : emdk::kSlotBytes;
  int tile = tile_req > 0 ? std::min(tile_req, cap) : cap;
  tile = std::max(0, std::min(tile, lds_fit_slots(h, wg, max_n, slot)));
  s->wg = wg;
  s->slot = slot;
  s->tile = tile;
  // a global knot tile of the full capacity backs sifts whose knots outgrow the LDS tile
  s->spill_slots = tile < cap ? cap : 0;
  s->max_n = std::max(max_n, 1);
  return EMD_OK;
}

// zero_fill: write zeros into each trace's rows at and past n_imfs (the padded layout,
// whose reference counterpart is a zero-initialised array); the ragged layout leaves them.
emdk::SiftParams params(const emd_handle* h, const Shape& s, int zero_fill) {
  emdk::SiftParams P;
  P.max_extrema = h->cfg.max_extrema;
  P.nbsym = h->cfg.nbsym;
  P.max_iteration = h->cfg.max_iteration;
  P.max_imfs = h->cfg.max_imfs;
  P.tile_slots = s.tile;
  P.spill_slots = s.spill_slots;
  P.max_n = s.max_n;
  P.zero_fill = zero_fill;
  return P;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return EMD_OK;
}

}  // namespace

extern "C" {

const char* emd_last_error(void) { return g_last_error.c_str(); }

int emd_version(void) { return 1; }

int emd_static_lds_bytes(int threads_per_block) {
  return emdk::static_lds_bytes(threads_per_block);
}

int emd_lds_slot_bytes(void) { return emdk::kSlotBytes; }

int emd_device_info(int* lds_bytes_per_workgroup, int* num_cus, int* warp_size) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (lds_bytes_per_workgroup) *lds_bytes_per_workgroup = (int)p.sharedMemPerBlock;
  if (num_cus) *num_cus = p.multiProcessorCount;
  if (warp_size) *warp_size = p.warpSize;
  return EMD_OK;
}

int emd_create(const emd_config* cfg, emd_handle** out) {
  if (!cfg || !out) return fail(EMD_ERR_VALUE, "emd_create: null argument");
  *out = nullptr;
  if (cfg->max_extrema < 1 || cfg->max_extrema > 32767)
    return fail(EMD_ERR_VALUE, "max_extrema must be in [1, 32767], got %d", cfg->max_extrema);
  if (cfg->nbsym < 1 || cfg->nbsym > emdk::kMaxNbsym)
    return fail(EMD_ERR_VALUE, "nbsym must be in [1, %d], got %d", emdk::kMaxNbsym,
                cfg->nbsym);
  if (cfg->max_iteration < 1)
    return fail(EMD_ERR_VALUE, "max_iteration must be >= 1, got %d", cfg->max_iteration);
  if (cfg->max_imfs < 1) return fail(EMD_ERR_VALUE, "max_imfs must be >= 1, got %d", cfg->max_imfs);
  const int tpb = cfg->threads_per_block;
  if (tpb != 64 && tpb != 128 && tpb != 256 && tpb != 512 && tpb != 1024)
    return fail(EMD_ERR_VALUE,
                "threads_per_block must be 64, 128, 256, 512 or 1024 on gfx950, got %d", tpb);
  if (cfg->tile_slots < 0) return fail(EMD_ERR_VALUE, "tile_slots must be >= 0");
  emd_handle* h = new emd_handle();
  h->cfg = *cfg;
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipGetDevice");
  }
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipGetDeviceProperties");
  }
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    std::string arch = p.gcnArchName;
    delete h;
    return fail(EMD_ERR_RUNTIME, "libemd_gfx950 needs a gfx950 (MI355X) device, found %s",
                arch.c_str());
  }
  h->num_cus = p.multiProcessorCount;
  h->lds_max = (int)p.sharedMemPerBlock;
  {
    const int rc = emdk::prepare_kernels(h->lds_max);
    if (rc != (int)hipSuccess) {
      delete h;
      return hip_fail((hipError_t)rc, "hipFuncSetAttribute(sift kernels)");
    }
  }
  e = hipMalloc(reinterpret_cast<void**>(&h->counters), 256);
  if (e != hipSuccess) {
    delete h;
    return fail(EMD_ERR_MEMORY, "hipMalloc(counters): %s", hipGetErrorString(e));
  }
  e = hipEventCreateWithFlags(&h->busy, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFree(h->counters);
    delete h;
    return hip_fail(e, "hipEventCreateWithFlags");
  }
  {
    const int rc = static_scratch_address(h->device, &h->static_scratch);
    if (rc) {
      (void)hipEventDestroy(h->busy);
      (void)hipFree(h->counters);
      delete h;
      return rc;
    }
  }
  *out = h;
  return EMD_OK;
}

int emd_destroy(emd_handle* h) {
  if (!h) return EMD_OK;
  hipError_t e = hipDeviceSynchronize();
  for (hipEvent_t ev : h->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (h->busy) (void)hipEventDestroy(h->busy);
  if (h->spill) (void)hipFree(h->spill);
  if (h->counters) (void)hipFree(h->counters);
  delete h;
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize in emd_destroy");
  return EMD_OK;
}

__global__ void zero_counters_kernel(unsigned* c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0u;
}

static int mark(emd_handle* h, int which, hipStream_t st) {
  if (!h->timing) return EMD_OK;
  hipError_t e = hipEventRecord(h->ev[which], st);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
  if (which == 1) h->timed = 1;
  return EMD_OK;
}

// Launch plan of one run: shapes, parameters, grid and the global-tile scratch it needs.
struct Plan {
  Shape s;
  emdk::SiftParams P;
  int persistent = 0;
  int grid = 0;            // persistent: resident workgroups; plain: traces per sub-launch
  int64_t stride = 0;      // bytes per workgroup global region (0: none needed)
  size_t spill_bytes = 0;  // stride x resident workgroups
};

static int make_plan(const emd_handle* h, int B, int max_n, const emd_launch& L, int zero_fill,
                     Plan* pl, int timed = 0) {
  int rc = make_shape(h, L.threads, L.tile_slots > 0 ? L.tile_slots : h->cfg.tile_slots, max_n,
                      &pl->s, timed);
  if (rc) return rc;
  pl->P = params(h, pl->s, zero_fill);
  pl->stride = pl->s.spill_slots ? emdk::tile_bytes(pl->s.spill_slots, pl->s.slot) : 0;
  pl->persistent = L.persistent ? 1 : 0;
  if (pl->persistent) {
    const size_t lds = emdk::dyn_lds_bytes(pl->s.tile, pl->P.max_n, pl->s.slot);
    const int fit = emdk::max_resident_wgs(1, pl->s.wg, lds, timed);
    if (fit <= 0) return fail(EMD_ERR_RUNTIME, "persistent kernel shape does not fit a CU");
    int per_cu = L.wgs_per_cu;
    if (per_cu <= 0 || per_cu > fit) per_cu = fit;
    pl->grid = std::min(B, h->num_cus * per_cu);
  } else {
    pl->grid = std::min(B, kMaxPlainBlocks);
  }
  pl->spill_bytes = (size_t)pl->stride * (size_t)pl->grid;
  return EMD_OK;
}

// Enqueue the sift launch(es) of a plan on st.  spill must hold plan.spill_bytes;
// counters is the queue head of a persistent launch.
static int enqueue(emd_handle* h, const Plan& pl, emdk::Batch b, int B, char* spill,
                   unsigned* counters, bool timed, hipStream_t st) {
  b.spill = pl.stride ? spill : nullptr;
  b.spill_stride = pl.stride;
  b.queue = counters;
  int rc;
  if (pl.persistent) {
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, st, counters);
    if ((rc = check_launch("queue reset"))) return rc;
    b.B = B;
    b.base = 0;
    if (timed && (rc = mark(h, 0, st))) return rc;
    rc = emdk::launch_sift(1, pl.s.wg, pl.grid, b, pl.P, st);
    if (rc) return hip_fail((hipError_t)rc, "persistent sift launch");
    if ((rc = check_launch("persistent sift kernel"))) return rc;
    return timed ? mark(h, 1, st) : EMD_OK;
  }
  if (timed && (rc = mark(h, 0, st))) return rc;
  for (int base = 0; base < B; base += pl.grid) {
    const int nb = std::min(pl.grid, B - base);
    b.B = nb;
    b.base = base;
    rc = emdk::launch_sift(0, pl.s.wg, nb, b, pl.P, st);
    if (rc) return hip_fail((hipError_t)rc, "sift launch");
    if ((rc = check_launch("sift kernel"))) return rc;
  }
  return timed ? mark(h, 1, st) : EMD_OK;
}

// Launches that use the handle's scratch are ordered: one on a stream other than the
// previous launch's first waits for that launch (skipped under stream capture, where the
// caller's graph orders its own nodes).
static int scratch_acquire(emd_handle* h, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) return hip_fail(e, "hipStreamIsCapturing");
  if (cs != hipStreamCaptureStatusNone) return EMD_OK;
  if (h->busy_set && h->busy_stream != st) {
    e = hipStreamWaitEvent(st, h->busy, 0);
    if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent(previous launch)");
  }
  return EMD_OK;
}

static int scratch_release(emd_handle* h, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) return hip_fail(e, "hipStreamIsCapturing");
  if (cs != hipStreamCaptureStatusNone) return EMD_OK;
  e = hipEventRecord(h->busy, st);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord(launch done)");
  h->busy_stream = st;
  h->busy_set = 1;
  return EMD_OK;
}

static int check_io(int B, int max_n, const double* values, double* imfs, int* n_imfs,
                    int* overflow, int* debug_n, int* debug_extno, int* debug_nzm) {
  if (B < 0) return fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (max_n < 0) return fail(EMD_ERR_VALUE, "trace length must be >= 0, got %d", max_n);
  if (B > 0 && ((max_n > 0 && (!values || !imfs)) || !n_imfs || !overflow || !debug_n ||
                !debug_extno || !debug_nzm))
    return fail(EMD_ERR_VALUE, "null input or output pointer");
  return EMD_OK;
}

static emdk::Batch make_batch(const double* values, const int64_t* offsets, int N,
                              const int* order, int total, double* imfs,
                              const int64_t* out_offsets, int* n_imfs, int* overflow,
                              int* debug_n, int* debug_extno, int* debug_nzm, int* stats) {
  emdk::Batch b{};
  b.values = values;
  b.offsets = offsets;
  b.N = N;
  b.order = order;
  b.total = total;
  b.imfs = imfs;
  b.out_offsets = out_offsets;
  b.n_imfs = n_imfs;
  b.overflow = overflow;
  b.dbg_n = debug_n;
  b.dbg_e = debug_extno;
  b.dbg_z = debug_nzm;
  b.stats = stats;
  return b;
}

// times: null (positions = sample indices), or the time-axis kernels' normalised times,
// laid out like values, or one shared row when times_shared.
static int run_common(emd_handle* h, const double* values, const int64_t* offsets, int N,
                      int B, int max_n, const int* order, const emd_launch& L, double* imfs,
                      const int64_t* out_offsets, int* n_imfs, int* overflow, int* debug_n,
                      int* debug_extno, int* debug_nzm, int* stats, hipStream_t st,
                      const double* times = nullptr, int times_shared = 0) {
  int rc = check_io(B, max_n, values, imfs, n_imfs, overflow, debug_n, debug_extno, debug_nzm);
  if (rc) return rc;
  h->timed = 0;
  if (B == 0) return EMD_OK;
  if (L.traces < 0) return fail(EMD_ERR_VALUE, "launch.traces must be >= 0, got %d", L.traces);
  Plan pl;
  if ((rc = make_plan(h, B, max_n, L, offsets == nullptr, &pl, times != nullptr))) return rc;
  if (pl.spill_bytes && (rc = ensure_spill(h, pl.spill_bytes))) return rc;
  emdk::Batch b = make_batch(values, offsets, N, order, L.traces > 0 ? L.traces : B, imfs,
                             out_offsets, n_imfs, overflow, debug_n, debug_extno, debug_nzm,
                             stats);
  b.times = times;
  b.times_shared = times_shared ? 1 : 0;
  if ((rc = scratch_acquire(h, st))) return rc;
  rc = enqueue(h, pl, b, B, h->spill, h->counters, h->timing != 0, st);
  const int rc2 = scratch_release(h, st);
  return rc ? rc : rc2;
}

int emd_run_padded(emd_handle* h, const double* signals, int B, int N, double* imfs,
                   int* n_imfs, int* overflow, int* debug_n, int* debug_extno,
                   int* debug_nzm, int* stats, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (N < 0 || N > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "trace length N=%d outside [0, %d]", N, emdk::kMaxN);
  emd_launch L;
  L.threads = h->cfg.threads_per_block;
  L.tile_slots = h->cfg.tile_slots;
  L.wgs_per_cu = 0;
  L.persistent = 0;
  L.traces = 0;
  return run_common(h, signals, nullptr, N, B, N, nullptr, L, imfs, nullptr, n_imfs, overflow,
                    debug_n, debug_extno, debug_nzm, stats, (hipStream_t)stream);
}

int emd_run_ragged(emd_handle* h, const double* values, const int64_t* offsets, int B,
                   int max_n, const int* order, const emd_launch* launch, double* imfs,
                   const int64_t* out_offsets, int* n_imfs, int* overflow, int* debug_n,
                   int* debug_extno, int* debug_nzm, int* stats, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (!offsets) return fail(EMD_ERR_VALUE, "ragged input needs CSR offsets");
  if (max_n < 0 || max_n > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "max trace length %d outside [0, %d]", max_n, emdk::kMaxN);
  emd_launch L;
  if (launch) {
    L = *launch;
  } else {
    L.threads = h->cfg.threads_per_block;
    L.tile_slots = h->cfg.tile_slots;
    L.wgs_per_cu = 0;
    L.persistent = 1;
    L.traces = 0;
  }
  return run_common(h, values, offsets, 0, B, max_n, order, L, imfs, out_offsets, n_imfs,
                    overflow, debug_n, debug_extno, debug_nzm, stats, (hipStream_t)stream);
}

int emd_run_padded_times(emd_handle* h, const double* signals, const double* times,
                         int times_per_signal, int B, int N, double* imfs, int* n_imfs,
                         int* overflow, int* debug_n, int* debug_extno, int* debug_nzm,
                         int* stats, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (N < 0 || N > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "trace length N=%d outside [0, %d]", N, emdk::kMaxN);
  if (B > 0 && N > 0 && !times) return fail(EMD_ERR_VALUE, "null time axis");
  emd_launch L;
  L.threads = h->cfg.threads_per_block;
  L.tile_slots = h->cfg.tile_slots;
  L.wgs_per_cu = 0;
  L.persistent = 0;
  L.traces = 0;
  return run_common(h, signals, nullptr, N, B, N, nullptr, L, imfs, nullptr, n_imfs, overflow,
                    debug_n, debug_extno, debug_nzm, stats, (hipStream_t)stream, times,
                    !times_per_signal);
}

int emd_run_ragged_times(emd_handle* h, const double* values, const double* times,
                         const int64_t* offsets, int B, int max_n, const int* order,
                         const emd_launch* launch, double* imfs, const int64_t* out_offsets,
                         int* n_imfs, int* overflow, int* debug_n, int* debug_extno,
                         int* debug_nzm, int* stats, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (!offsets) return fail(EMD_ERR_VALUE, "ragged input needs CSR offsets");
  if (max_n < 0 || max_n > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "max trace length %d outside [0, %d]", max_n, emdk::kMaxN);
  if (B > 0 && max_n > 0 && !times) return fail(EMD_ERR_VALUE, "null time axis");
  emd_launch L;
  if (launch) {
    L = *launch;
  } else {
    L.threads = h->cfg.threads_per_block;
    L.tile_slots = h->cfg.tile_slots;
    L.wgs_per_cu = 0;
    L.persistent = 1;
    L.traces = 0;
  }
  return run_common(h, values, offsets, 0, B, max_n, order, L, imfs, out_offsets, n_imfs,
                    overflow, debug_n, debug_extno, debug_nzm, stats, (hipStream_t)stream, times,
                    0);
}

int emd_run_ragged_noalloc(emd_handle* h, const double* values, const int64_t* offsets, int B,
                           int max_n, const int* order, const emd_launch* launch, double* imfs,
                           const int64_t* out_offsets, int* n_imfs, int* overflow, int* debug_n,
                           int* debug_extno, int* debug_nzm, int* stats, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (!offsets) return fail(EMD_ERR_VALUE, "ragged input needs CSR offsets");
  if (max_n < 0 || max_n > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "max trace length %d outside [0, %d]", max_n, emdk::kMaxN);
  emd_launch L;
  if (launch) {
    L = *launch;
  } else {
    L.threads = h->cfg.threads_per_block;
    L.tile_slots = h->cfg.tile_slots;
    L.wgs_per_cu = 0;
    L.persistent = 1;
    L.traces = 0;
  }
  int rc = check_io(B, max_n, values, imfs, n_imfs, overflow, debug_n, debug_extno, debug_nzm);
  if (rc) return rc;
  if (L.traces < 0) return fail(EMD_ERR_VALUE, "launch.traces must be >= 0, got %d", L.traces);
  h->timed = 0;
  if (B == 0) return EMD_OK;
  if (!h->static_scratch) return fail(EMD_ERR_RUNTIME, "static scratch address unknown");
  Plan pl;
  if ((rc = make_plan(h, B, max_n, L, 0, &pl))) return rc;
  const size_t room = kStaticScratchBytes - kCounterBytes;
  if (pl.stride > 0) {
    if ((size_t)pl.stride > room)
      return fail(EMD_ERR_VALUE,
                  "this launch needs %lld bytes of global knot tile per "
                  "workgroup; the static scratch holds %zu (lower max_extrema or max_n, or "
                  "use emd_run_ragged)",
                  (long long)pl.stride, room);
    // as many workgroups as the scratch holds regions for (fewer resident workgroups or
    // shorter plain sub-launches; the results do not depend on either)
    pl.grid = (int)std::min<int64_t>(pl.grid, (int64_t)(room / (size_t)pl.stride));
    pl.spill_bytes = (size_t)pl.stride * (size_t)pl.grid;
  }
  const emdk::Batch b = make_batch(values, offsets, 0, order, L.traces > 0 ? L.traces : B,
                                   imfs, out_offsets, n_imfs, overflow, debug_n, debug_extno,
                                   debug_nzm, stats);
  return enqueue(h, pl, b, B, h->static_scratch + kCounterBytes,
                 reinterpret_cast<unsigned*>(h->static_scratch), h->timing != 0,
                 (hipStream_t)stream);
}

}  // extern "C"

// The envelope-only pass, with positions from times (null: sample indices; else [B,N]
// rows, or one shared [N] row when times_shared).
static int envelope_common(emd_handle* h, const double* signals, const double* times,
                           int times_shared, int B, int N, double* max_env, double* min_env,
                           int* n_extrema, int* degenerate, double* max_kpos, double* max_kval,
                           double* min_kpos, double* min_kval, int* n_mirrored, void* stream) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (N < 0 || N > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "trace length N=%d outside [0, %d]", N, emdk::kMaxN);
  if (B < 0) return fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (B == 0) return EMD_OK;
  Shape s;
  int rc = make_shape(h, 256, h->cfg.tile_slots, N, &s, times != nullptr);
  if (rc) return rc;
  const emdk::SiftParams P = params(h, s, 1);
  const int mbuf = h->cfg.max_extrema + 2 * h->cfg.nbsym;
  const int64_t stride = s.spill_slots ? emdk::tile_bytes(s.spill_slots, s.slot) : 0;
  if (stride) {
    rc = ensure_spill(h, (size_t)stride * std::min(B, kMaxPlainBlocks));
    if (rc) return rc;
  }
  if ((rc = scratch_acquire(h, (hipStream_t)stream))) return rc;
  for (int base = 0; base < B; base += kMaxPlainBlocks) {
    const int nb = std::min(kMaxPlainBlocks, B - base);
    emdk::EnvBatch e{};
    e.values = signals + (int64_t)base * N;
    e.times = times ? times + (times_shared ? 0 : (int64_t)base * N) : nullptr;
    e.times_shared = times_shared ? 1 : 0;
    e.N = N;
    e.B = nb;
    e.max_env = max_env + (int64_t)base * N;
    e.min_env = min_env + (int64_t)base * N;
    e.n_extrema = n_extrema + 2 * (int64_t)base;
    e.degenerate = degenerate + base;
    e.max_kpos = max_kpos + (int64_t)base * mbuf;
    e.max_kval = max_kval + (int64_t)base * mbuf;
    e.min_kpos = min_kpos + (int64_t)base * mbuf;
    e.min_kval = min_kval + (int64_t)base * mbuf;
    e.n_mirrored = n_mirrored + 2 * (int64_t)base;
    e.mbuf = mbuf;
    e.spill = stride ? h->spill : nullptr;
    e.spill_stride = stride;
    rc = emdk::launch_envelope(s.wg, e, P, (hipStream_t)stream);
    if (rc) return hip_fail((hipError_t)rc, "envelope launch");
    rc = check_launch("envelope kernel");
    if (rc) return rc;
  }
  return scratch_release(h, (hipStream_t)stream);
}

extern "C" {

int emd_envelope_only(emd_handle* h, const double* signals, int B, int N, double* max_env,
                      double* min_env, int* n_extrema, int* degenerate, double* max_kpos,
                      double* max_kval, double* min_kpos, double* min_kval, int* n_mirrored,
                      void* stream) {
  return envelope_common(h, signals, nullptr, 0, B, N, max_env, min_env, n_extrema, degenerate,
                         max_kpos, max_kval, min_kpos, min_kval, n_mirrored, stream);
}

int emd_envelope_only_times(emd_handle* h, const double* signals, const double* times,
                            int times_per_signal, int B, int N, double* max_env,
                            double* min_env, int* n_extrema, int* degenerate, double* max_kpos,
                            double* max_kval, double* min_kpos, double* min_kval,
                            int* n_mirrored, void* stream) {
  if (B > 0 && N > 0 && !times) return fail(EMD_ERR_VALUE, "null time axis");
  return envelope_common(h, signals, times, !times_per_signal, B, N, max_env, min_env,
                         n_extrema, degenerate, max_kpos, max_kval, min_kpos, min_kval,
                         n_mirrored, stream);
}

}  // extern "C"

struct emd_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  char* spill = nullptr;
  unsigned* counters = nullptr;
};

static void graph_free(emd_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->spill) (void)hipFree(g->spill);
  if (g->counters) (void)hipFree(g->counters);
  delete g;
}

extern "C" {

int emd_chunk_graph_create(emd_handle* h, const emd_chunk* c, emd_graph** out) {
  if (!h || !c || !out) return fail(EMD_ERR_VALUE, "null argument");
  *out = nullptr;
  const int max_n = c->offsets ? c->max_n : c->N;
  if (max_n < 0 || max_n > emdk::kMaxN)
    return fail(EMD_ERR_VALUE, "trace length %d outside [0, %d]", max_n, emdk::kMaxN);
  int rc = check_io(c->B, max_n, c->values, c->imfs, c->n_imfs, c->overflow, c->debug_n,
                    c->debug_extno, c->debug_nzm);
  if (rc) return rc;
  if (c->B < 1) return fail(EMD_ERR_VALUE, "a captured chunk needs B >= 1");
  if (!c->row_offsets != !c->compact)
    return fail(EMD_ERR_VALUE, "compaction needs both row_offsets and compact");
  if (c->launch.traces < 0)
    return fail(EMD_ERR_VALUE, "launch.traces must be >= 0, got %d", c->launch.traces);
  Plan pl;
  if ((rc = make_plan(h, c->B, max_n, c->launch, c->offsets == nullptr, &pl))) return rc;
  emd_graph* g = new emd_graph();
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&g->counters), 256);
  if (e == hipSuccess && pl.spill_bytes)
    e = hipMalloc(reinterpret_cast<void**>(&g->spill), pl.spill_bytes);
  if (e != hipSuccess) {
    graph_free(g);
    return fail(EMD_ERR_MEMORY, "hipMalloc(chunk graph scratch): %s", hipGetErrorString(e));
  }
  hipStream_t cs = nullptr;
  e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    graph_free(g);
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    graph_free(g);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  const emdk::Batch b = make_batch(c->values, c->offsets, c->N, c->order,
                                   c->launch.traces > 0 ? c->launch.traces : c->B, c->imfs,
                                   c->out_offsets, c->n_imfs, c->overflow, c->debug_n,
                                   c->debug_extno, c->debug_nzm, c->stats);
  rc = enqueue(h, pl, b, c->B, g->spill, g->counters, false, cs);
  if (!rc && c->compact) {
    const int r2 = emdk::launch_compact(c->imfs, c->out_offsets, h->cfg.max_imfs, c->n_imfs,
                                        c->offsets, c->N, c->B, c->row_offsets, c->compact, cs);
    if (r2) rc = hip_fail((hipError_t)r2, "compaction launch");
  }
  e = hipStreamEndCapture(cs, &g->graph);
  (void)hipStreamDestroy(cs);
  if (rc) {
    graph_free(g);
    return rc;
  }
  if (e != hipSuccess) {
    graph_free(g);
    return hip_fail(e, "hipStreamEndCapture");
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    graph_free(g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  *out = g;
  return EMD_OK;
}

int emd_chunk_graph_launch(emd_graph* g, void* stream) {
  if (!g || !g->exec) return fail(EMD_ERR_VALUE, "null graph");
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  return e == hipSuccess ? EMD_OK : hip_fail(e, "hipGraphLaunch");
}

int emd_chunk_graph_destroy(emd_graph* g) {
  if (!g) return EMD_OK;
  hipError_t e = hipDeviceSynchronize();
  graph_free(g);
  return e == hipSuccess ? EMD_OK : hip_fail(e, "hipDeviceSynchronize in graph destroy");
}

int emd_set_timing(emd_handle* h, int enable) {
  if (!h) return fail(EMD_ERR_VALUE, "null handle");
  if (enable && !h->ev[0]) {
    for (hipEvent_t& ev : h->ev) {
      hipError_t e = hipEventCreate(&ev);
      if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    }
  }
  h->timing = enable ? 1 : 0;
  h->timed = 0;
  return EMD_OK;
}

int emd_kernel_time_ms(emd_handle* h, float* ms) {
  if (!h || !ms) return fail(EMD_ERR_VALUE, "null argument");
  if (!h->timed) return fail(EMD_ERR_VALUE, "no timed sift launch since timing was enabled");
  hipError_t e = hipEventSynchronize(h->ev[1]);
  if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
  e = hipEventElapsedTime(ms, h->ev[0], h->ev[1]);
  if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
  return EMD_OK;
}

int emd_max_resident(emd_handle* h, const emd_launch* launch, int max_n, int* wgs_per_cu) {
  if (!h || !launch || !wgs_per_cu) return fail(EMD_ERR_VALUE, "null argument");
  Shape s;
  int rc = make_shape(h, launch->threads, launch->tile_slots > 0 ? launch->tile_slots
                                                                 : h->cfg.tile_slots,
                      max_n, &s);
  if (rc) return rc;
  *wgs_per_cu = emdk::max_resident_wgs(launch->persistent ? 1 : 0, s.wg,
                                       emdk::dyn_lds_bytes(s.tile, s.max_n));
  return EMD_OK;
}

}  // extern "C"
