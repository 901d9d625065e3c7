// emd_sift.h -- declarations shared by the gfx950 sift kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emdk {

constexpr int kMaxNbsym = 4;      // mirrored points per side supported by the kernels
constexpr int kMaxN = 16384;      // longest trace (256 blocks of 64 samples)
constexpr int kSlotBytes = 28;    // bytes per knot slot: position int32, value, M, 1/h fp64
constexpr int kSlotBytesTime = 32;  // the same with an fp64 time position
constexpr int kWorkBlockBytes = 568;  // LDS bytes per 64-sample block of the work area

struct SiftParams {
  int max_extrema;    // per-side extrema cap (overflow rule); counts are clamped to it
  int nbsym;          // mirrored points per side (1..kMaxNbsym)
  int max_iteration;  // sift cap per IMF
  int max_imfs;       // output rows per trace (IMF cap = max_imfs - 1, then the residue)
  int tile_slots;     // knot slots of the LDS tile (both envelopes)
  int spill_slots;    // knot slots of each workgroup's global knot tile (0: none), which
                      // holds any trace's knots when the LDS tile does not
  int max_n;          // longest trace of the launch (sizes the LDS work area)
  int zero_fill;      // write zeros into output rows at and beyond n_imfs
};

// Addressing of one sift launch (uniform [B,N] rows or CSR traces).
struct Batch {
  const double* values;
  const int64_t* offsets;      // CSR offsets [B+1] or null (uniform rows of length N)
  int N;                       // uniform length when offsets == null
  int B;                       // traces handled by this launch
  int base;                    // first queue position of this launch
  const int* order;            // processing order or null (FIFO)
  int total;                   // traces the offsets describe: order entries outside
                               // [0, total) are skipped
  double* imfs;
  const int64_t* out_offsets;  // per-signal output base (elements) or null (padded)
  int* n_imfs;
  int* overflow;
  int* dbg_n;
  int* dbg_e;
  int* dbg_z;
  int* stats;
  char* spill;                 // per-workgroup global knot tiles
  int64_t spill_stride;        // bytes per workgroup
  unsigned* queue;             // persistent-kernel queue head
  const double* times;         // time-axis kernels: sample times (normalised), laid out
                               // like values, or one shared row when times_shared
  int times_shared;
};

// Envelope-only pass outputs (reference envelope_batch_kernel ABI, kernels.py:1100-1113).
struct EnvBatch {
  const double* values;
  int N;
  int B;
  double* max_env;
  double* min_env;
  int* n_extrema;   // [B,2]
  int* degenerate;  // [B]
  double* max_kpos; // [B,mbuf]
  double* max_kval;
  double* min_kpos;
  double* min_kval;
  int* n_mirrored;  // [B,2]
  int mbuf;
  char* spill;
  int64_t spill_stride;
  const double* times;  // as in Batch: [B,N] rows, or one shared [N] row when times_shared
  int times_shared;
};


// Bytes of one workgroup's global knot tile of `slots` slots of slot_bytes each.
int64_t tile_bytes(int slots, int slot_bytes = kSlotBytes);
// Static LDS of the sift kernels for a workgroup size.
int static_lds_bytes(int wg);
// Dynamic LDS of a launch: the per-trace work area (IMF + block bookkeeping) for traces of
// up to max_n samples, plus the LDS knot tile.
size_t dyn_lds_bytes(int tile_slots, int max_n, int slot_bytes = kSlotBytes);

// Launch one sift kernel instance.  persistent != 0 selects the queue kernel, whose grid
// is `grid` workgroups; otherwise grid is the number of traces of the launch.  A batch with
// times runs the time-axis instance (knot slots of kSlotBytesTime).  Returns a hipError_t.
int launch_sift(int persistent, int wg, int grid, const Batch& b, const SiftParams& P,
                hipStream_t stream);
int launch_envelope(int wg, const EnvBatch& e, const SiftParams& P, hipStream_t stream);
int reset_queue(unsigned* q, hipStream_t stream);
// Resident workgroups per CU for a kernel instance at the given dynamic LDS (timed: the
// time-axis instance).
int max_resident_wgs(int persistent, int wg, size_t dyn_lds, int timed = 0);
// Raise every kernel's dynamic-LDS limit to the device's per-workgroup LDS (once per handle).
int prepare_kernels(int lds_per_workgroup);

// Enqueue the compaction scan (row_offsets[B+1]) and, when compact is non-null, the row copy.
int launch_compact(const double* imfs, const int64_t* out_offsets, int max_imfs,
                   const int* n_imfs, const int64_t* offsets, int N, int B,
                   int64_t* row_offsets, double* compact, hipStream_t st);

// Sets the thread-local message emd_last_error() returns; returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// fail() for a HIP runtime error: EMD_ERR_RUNTIME with "what: <hip message>".
int hip_fail(hipError_t e, const char* what);

}  // namespace emdk
