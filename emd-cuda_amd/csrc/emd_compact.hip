// emd_compact.hip -- output compaction: keep only each trace's n_imfs rows, packed back to
// back at prefix offsets (exclusive scan of n_imfs x length).  No reference counterpart:
// the reference returns the padded [B, MAX_IMFS, N] array (kernels.py:1353-1372).
#include "../../include/emd_gfx950.h"

#include <hip/hip_runtime.h>

#include <cstdint>

#include "emd_sift.h"

namespace emdk {

namespace {

constexpr int kScanThreads = 1024;

__device__ __forceinline__ int64_t trace_len(const int64_t* offsets, int N, int64_t s) {
  return offsets ? offsets[s + 1] - offsets[s] : (int64_t)N;
}

// One workgroup: each lane sums a contiguous run of traces, the 1024 run totals are
// scanned in LDS, and each lane writes its run's exclusive prefixes.
__global__ __launch_bounds__(kScanThreads) void compact_scan_kernel(
    const int* __restrict__ n_imfs, const int64_t* __restrict__ offsets, int N, int B,
    int64_t* __restrict__ row_offsets) {
  __shared__ int64_t part[kScanThreads];
  const int t = threadIdx.x;
  const int64_t per = ((int64_t)B + kScanThreads - 1) / kScanThreads;
  const int64_t lo = min((int64_t)B, (int64_t)t * per);
  const int64_t hi = min((int64_t)B, lo + per);
  int64_t sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += (int64_t)n_imfs[i] * trace_len(offsets, N, i);
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < kScanThreads; d <<= 1) {
    const int64_t v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  for (int64_t i = lo; i < hi; ++i) {
    row_offsets[i] = run;
    run += (int64_t)n_imfs[i] * trace_len(offsets, N, i);
  }
  if (t == kScanThreads - 1) row_offsets[B] = part[t];
}

// Trace s's rows are contiguous in every output layout (out_offsets[s], or by default the
// trace's input offset x max_imfs), so each trace is one contiguous copy of
// n_imfs[s] * len(s) doubles.
__global__ __launch_bounds__(256) void compact_rows_kernel(
    const double* __restrict__ imfs, const int64_t* __restrict__ out_offsets, int max_imfs,
    const int* __restrict__ n_imfs, const int64_t* __restrict__ offsets, int N, int B,
    const int64_t* __restrict__ row_offsets, double* __restrict__ compact) {
  for (int s = blockIdx.x; s < B; s += gridDim.x) {
    const int64_t len = trace_len(offsets, N, s);
    const int64_t in_off = offsets ? offsets[s] : (int64_t)s * N;
    const double* src = imfs + (out_offsets ? out_offsets[s] : in_off * max_imfs);
    double* dst = compact + row_offsets[s];
    const int64_t cnt = (int64_t)n_imfs[s] * len;
    for (int64_t i = threadIdx.x; i < cnt; i += 256) dst[i] = src[i];
  }
}

}  // namespace

int launch_compact(const double* imfs, const int64_t* out_offsets, int max_imfs,
                   const int* n_imfs, const int64_t* offsets, int N, int B,
                   int64_t* row_offsets, double* compact, hipStream_t st) {
  if (B <= 0) return (int)hipSuccess;
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(kScanThreads), 0, st, n_imfs, offsets,
                     N, B, row_offsets);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !compact) return (int)e;
  hipLaunchKernelGGL(compact_rows_kernel, dim3(B < 8192 ? B : 8192), dim3(256), 0, st, imfs,
                     out_offsets, max_imfs, n_imfs, offsets, N, B, row_offsets, compact);
  return (int)hipGetLastError();
}

}  // namespace emdk

extern "C" {

int emd_compact_offsets(const int* n_imfs, const int64_t* offsets, int N, int B,
                        int64_t* row_offsets, void* stream) {
  if (B < 0) return emdk::fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (B > 0 && (!n_imfs || !row_offsets))
    return emdk::fail(EMD_ERR_VALUE, "null n_imfs or row_offsets");
  const int rc = emdk::launch_compact(nullptr, nullptr, 0, n_imfs, offsets, N, B, row_offsets,
                                      nullptr, (hipStream_t)stream);
  return rc ? emdk::hip_fail((hipError_t)rc, "compaction scan") : EMD_OK;
}

int emd_compact(const double* imfs, const int64_t* out_offsets, int max_imfs, const int* n_imfs,
                const int64_t* offsets, int N, int B, int64_t* row_offsets, double* compact,
                void* stream) {
  if (B < 0) return emdk::fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (B > 0 && (!imfs || !n_imfs || !row_offsets || !compact))
    return emdk::fail(EMD_ERR_VALUE, "null pointer");
  if (!out_offsets && max_imfs < 1)
    return emdk::fail(EMD_ERR_VALUE, "the default row layout needs max_imfs >= 1");
  const int rc = emdk::launch_compact(imfs, out_offsets, max_imfs, n_imfs, offsets, N, B,
                                      row_offsets, compact, (hipStream_t)stream);
  return rc ? emdk::hip_fail((hipError_t)rc, "compaction") : EMD_OK;
}

}  // extern "C"
