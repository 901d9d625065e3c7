// emd_hilbert.hip -- the Hilbert spectral step: instantaneous amplitude, phase and
// frequency of every IMF row, from the row's analytic signal z = x + i H x.
//
// No reference counterpart: the reference stops at the IMFs of traces it describes as
// seismic, sampled at 4 ms (emd.py:28-30); this step is what seismic attribute work runs
// on them next.  H x is taken as in scipy.signal.hilbert: the DFT of the row, times
// -i sgn(k), transformed back.  For the DFT of the row's own length N (any N) the kernel
// computes the equivalent circular convolution with the closed-form discrete Hilbert
// kernel on 2^ceil(log2(2N-1)) points; for a power-of-two length it multiplies by
// -i sgn(k) directly (emd_hilbert.h).
//
// One workgroup per trace: the trace's rows are transformed two at a time (x1 + i x2) with
// the points in registers and the stage exchanges in LDS, then each row's outputs are
// written in natural order, coalesced.  Every launch is one transform class (one length and
// form): a batch whose rows share one class is one launch of a workgroup per trace; a batch
// of mixed lengths takes one launch per class its length range spans, each walking the
// batch with the workgroups resident at once and skipping the traces of other classes.
#include "../../include/emd_gfx950.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "emd_hilbert.h"
#include "emd_sift.h"

namespace emdh {

namespace {

struct Rows {
  const double* imfs;
  const int64_t* row_base;  // first element of each trace's rows, or null (default layout)
  int max_imfs;
  const int* n_imfs;
  const int64_t* offsets;   // CSR offsets [B+1], or null (every trace N samples)
  int N;
  int B;
  int n_fft;
  double inv_2pi_dt;
  double* amp;
  double* phase;
  double* freq;
};

// threadIdx.x behind an empty asm: every stage recomputes its lane-derived positions and
// twiddles instead of the compiler hoisting all of them out of the row loop, which would
// hold them in registers for the whole trace.
__device__ __forceinline__ int fresh_lane() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ void trace_geometry(const Rows& a, int s, int& len, int64_t& base) {
  const int64_t off = a.offsets ? a.offsets[s] : (int64_t)s * a.N;
  len = a.offsets ? (int)(a.offsets[s + 1] - off) : a.N;
  base = a.row_base ? a.row_base[s] : off * a.max_imfs;
}

__constant__ double kAtanPoly[kAtanTerms] = EMDH_ATAN_POLY;
__constant__ double kTrigPoly[2 * kTrigTerms] = EMDH_TRIG_POLY;

// A polynomial table's address behind an empty asm: the table's loads stay where it is
// used (scalar loads into scalar registers) instead of its constants being hoisted out of
// every loop into vector registers held through the whole kernel.
__device__ __forceinline__ const double* table(const double* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// Amplitude and phase of sample n of a row with analytic signal x + i h; returns the phase.
__device__ __forceinline__ double emit_amp_phase(const Rows& a, int64_t out, int n, double xv,
                                                 double h) {
  if (a.amp) a.amp[out + n] = sqrt(xv * xv + h * h);
  const double ph = atan2_poly(h, xv, table(kAtanPoly));
  if (a.phase) a.phase[out + n] = ph;
  return ph;
}

// Frequency of sample n from the row's phases: the increment from m to m+1 (m = n except at
// the last sample) wrapped into [-pi, pi] as numpy.unwrap does (an increment of exactly
// +-pi is kept).
__device__ __forceinline__ void emit_freq(const Rows& a, const double* ph, int64_t out, int n,
                                          int len) {
  double f = 0.0;
  if (len > 1) {
    const int m = n < len - 1 ? n : len - 2;
    double d = ph[m + 1] - ph[m];
    if (d > M_PI) d -= 2.0 * M_PI;
    else if (d < -M_PI) d += 2.0 * M_PI;
    f = d * a.inv_2pi_dt;
  }
  a.freq[out + n] = f;
}

// Points move from their stage-FROM positions to their stage-TO positions through LDS,
// real parts then imaginary parts (one double per point).
template <int LOGF, int FROM, int TO>
__device__ __forceinline__ void exchange(double* xch, cpx* v, bool act) {
  using P = Plan<LOGF>;
  const int t = fresh_lane();
  double* wb = xch + P::lds_base(FROM, t);
  double* rb = xch + P::lds_base(TO, t);
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) wb[P::lds_off(FROM, e)] = v[e].re;
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) v[e].re = rb[P::lds_off(TO, e)];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) wb[P::lds_off(FROM, e)] = v[e].im;
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) v[e].im = rb[P::lds_off(TO, e)];
  }
}

template <int LOGF, int S>
__device__ __forceinline__ void forward(double* xch, cpx* v, bool act) {
  if constexpr (S < Plan<LOGF>::NS) {
    if constexpr (S > 0) exchange<LOGF, S - 1, S>(xch, v, act);
    if (act) stage_fwd<LOGF, S>(v, fresh_lane(), table(kTrigPoly));
    forward<LOGF, S + 1>(xch, v, act);
  }
}

template <int LOGF, int S>
__device__ __forceinline__ void inverse(double* xch, cpx* v, bool act) {
  if constexpr (S >= 0) {
    if (act) stage_inv<LOGF, S>(v, fresh_lane(), table(kTrigPoly));
    if constexpr (S > 0) exchange<LOGF, S, S - 1>(xch, v, act);
    inverse<LOGF, S - 1>(xch, v, act);
  }
}

// Writes the outputs of one row whose H x sits in v[e].re (IM false) or v[e].im (IM true)
// at the stage-0 positions of the lane: H x goes to LDS, each sample's phase replaces it
// there, and the frequencies are the differences of neighbouring phases.
template <int LOGF, bool IM>
__device__ __forceinline__ void emit_row(const Rows& a, double* xch, const double* x,
                                         int64_t out, const cpx* v, int len, bool act) {
  using P = Plan<LOGF>;
  const int t = fresh_lane();
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) {
      const int n = P::pos(0, t, e);
      if (n < len) xch[n] = IM ? v[e].im : v[e].re;
    }
  }
  __syncthreads();
  // samples in natural order, a workgroup-wide stride at a time (coalesced, and a rolled
  // loop: the transcendental outputs are the bulk of the code)
#pragma unroll 1
  for (int n = threadIdx.x; n < len; n += P::WG) xch[n] = emit_amp_phase(a, out, n, x[n], xch[n]);
  if (a.freq) {
    __syncthreads();
#pragma unroll 1
    for (int n = threadIdx.x; n < len; n += P::WG) emit_freq(a, xch, out, n, len);
  }
}

// Both rows of a pair at once: H x1 at xch[0, len), H x2 at xch[len, 2 len) (2 len <= F + 1
// in the convolution form), so no transform register is live through the output loops.
template <int LOGF>
__device__ __forceinline__ void emit_pair(const Rows& a, double* xch, const double* x1,
                                          int64_t out, const cpx* v, int len, bool two,
                                          bool act) {
  using P = Plan<LOGF>;
  const int t = fresh_lane();
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < P::R; ++e) {
      const int n = P::pos(0, t, e);
      if (n < len) {
        xch[n] = v[e].re;
        xch[len + n] = v[e].im;
      }
    }
  }
  __syncthreads();
  // the two rows are adjacent in memory (x2 = x1 + len), as are their outputs
  const int cnt = two ? 2 * len : len;
#pragma unroll 1
  for (int n = threadIdx.x; n < cnt; n += P::WG) xch[n] = emit_amp_phase(a, out, n, x1[n], xch[n]);
  if (a.freq) {
    __syncthreads();
#pragma unroll 1
    for (int n = threadIdx.x; n < cnt; n += P::WG) {
      const int r = n >= len;
      emit_freq(a, xch + r * len, out + r * len, n - r * len, len);
    }
  }
}

template <int LOGF, bool EXACT>
__global__ __launch_bounds__(Plan<LOGF>::WG) void hilbert_kernel(Rows a) {
  using P = Plan<LOGF>;
  extern __shared__ double xch[];
  const bool act = (int)threadIdx.x < P::T;
  for (int s = blockIdx.x; s < a.B; s += gridDim.x) {
    int len;
    int64_t base;
    trace_geometry(a, s, len, base);
    bool exact;
    const int nim = a.n_imfs[s];
    if (nim <= 0 || transform_class(len, a.n_fft, &exact) != LOGF || exact != EXACT) continue;

    // exact form: the Hilbert multiplier -i S[k] / F of the points this lane holds after
    // the forward pass, from the transform of the circular kernel (real and odd, so its
    // transform is -i S)
    double S[EXACT ? P::R : 1];
    if constexpr (EXACT) {
      cpx g[P::R];
      const int t = fresh_lane();
      const double* T = table(kTrigPoly);
#pragma unroll
      for (int e = 0; e < P::R; ++e) g[e] = cpx{circular_kernel(P::pos(0, t, e), len, P::F, T), 0.0};
      forward<LOGF, 0>(xch, g, act);
#pragma unroll
      for (int e = 0; e < P::R; ++e) S[e] = -g[e].im * (1.0 / P::F);
    }

    for (int r = 0; r < nim; r += 2) {
      const bool two = r + 1 < nim;
      const int64_t out = base + (int64_t)r * len;
      const double* x1 = a.imfs + out;
      const double* x2 = x1 + len;
      cpx v[P::R];
      {
        const int t = fresh_lane();
#pragma unroll
        for (int e = 0; e < P::R; ++e) {
          const int n = P::pos(0, t, e);
          const bool in = act && n < len;
          v[e] = cpx{in ? x1[n] : 0.0, in && two ? x2[n] : 0.0};
        }
      }
      forward<LOGF, 0>(xch, v, act);
      {
        // (-i S) (re + i im) = S im - i S re
        const int t = fresh_lane();
#pragma unroll
        for (int e = 0; e < P::R; ++e) {
          const double se = EXACT ? S[EXACT ? e : 0] : P::sgn_over_f(t, e);
          v[e] = cpx{se * v[e].im, -se * v[e].re};
        }
      }
      inverse<LOGF, P::NS - 1>(xch, v, act);
      if constexpr (EXACT) {
        emit_pair<LOGF>(a, xch, x1, out, v, len, two, act);
      } else {
        emit_row<LOGF, false>(a, xch, x1, out, v, len, act);
        if (two) emit_row<LOGF, true>(a, xch, x2, out + len, v, len, act);
      }
    }
  }
}

// Rows of one or two samples (or a fixed transform of one or two points): H x = 0.
__global__ __launch_bounds__(256) void hilbert_trivial_kernel(Rows a) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < a.B; s += gridDim.x * blockDim.x) {
    int len;
    int64_t base;
    trace_geometry(a, s, len, base);
    bool exact;
    const int nim = a.n_imfs[s];
    if (nim <= 0 || len <= 0 || transform_class(len, a.n_fft, &exact) != kTrivial) continue;
    for (int r = 0; r < nim; ++r) {
      const int64_t out = base + (int64_t)r * len;
      const double* x = a.imfs + out;
      double ph[2];
      for (int n = 0; n < len; ++n) ph[n] = emit_amp_phase(a, out, n, x[n], 0.0);
      if (a.freq)
        for (int n = 0; n < len; ++n) emit_freq(a, ph, out, n, len);
    }
  }
}

using KernelFn = void (*)(Rows);

template <int LOGF>
KernelFn kernel_of(bool exact) {
  return exact ? hilbert_kernel<LOGF, true> : hilbert_kernel<LOGF, false>;
}

struct ClassInfo {
  KernelFn fn;
  int wg;
  size_t lds;
};

template <int LOGF>
ClassInfo info_of(bool exact) {
  return ClassInfo{kernel_of<LOGF>(exact), Plan<LOGF>::WG,
                   (size_t)Plan<LOGF>::XCH * sizeof(double)};
}

ClassInfo class_info(int logf, bool exact) {
  switch (logf) {
    case 2: return info_of<2>(exact);
    case 3: return info_of<3>(exact);
    case 4: return info_of<4>(exact);
    case 5: return info_of<5>(exact);
    case 6: return info_of<6>(exact);
    case 7: return info_of<7>(exact);
    case 8: return info_of<8>(exact);
    case 9: return info_of<9>(exact);
    case 10: return info_of<10>(exact);
    case 11: return info_of<11>(exact);
    case 12: return info_of<12>(exact);
    case 13: return info_of<13>(exact);
    case 14: return info_of<14>(exact);
  }
  return ClassInfo{nullptr, 0, 0};
}

// Resident workgroups per CU of each class (queried once per process; the same on every
// MI355X) and the device's CU count.
std::mutex g_mu;
int g_resident[2][kMaxLogF + 1] = {};
int g_cus = 0;

int prepare(int logf, bool exact, int* resident) {
  std::lock_guard<std::mutex> lock(g_mu);
  int& nb = g_resident[exact ? 1 : 0][logf];
  if (nb == 0) {
    const ClassInfo ci = class_info(logf, exact);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(ci.fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)ci.lds);
    if (e != hipSuccess) return emdk::hip_fail(e, "hipFuncSetAttribute(hilbert kernel)");
    int q = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, reinterpret_cast<const void*>(ci.fn),
                                                     ci.wg, ci.lds);
    if (e != hipSuccess) return emdk::hip_fail(e, "occupancy query (hilbert kernel)");
    if (q < 1) return emdk::fail(EMD_ERR_RUNTIME, "hilbert kernel 2^%d does not fit a CU", logf);
    nb = q;
  }
  if (g_cus == 0) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
      e = hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return emdk::hip_fail(e, "CU count");
  }
  *resident = nb * g_cus;
  return EMD_OK;
}

int launch_class(const Rows& a, int logf, bool exact, bool only_class, hipStream_t st) {
  int resident = 0;
  int rc = prepare(logf, exact, &resident);
  if (rc) return rc;
  const ClassInfo ci = class_info(logf, exact);
  // a launch that holds every trace gets one workgroup per trace; one class of several
  // walks the batch with as many workgroups as are resident at once
  const int grid = only_class ? a.B : (a.B < resident ? a.B : resident);
  Rows aa = a;
  void* args[] = {&aa};
  hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(ci.fn), dim3(grid), dim3(ci.wg),
                                 args, ci.lds, st);
  return e == hipSuccess ? EMD_OK : emdk::hip_fail(e, "hilbert kernel launch");
}

int launch_trivial(const Rows& a, hipStream_t st) {
  const int grid = (a.B + 255) / 256 < 1024 ? (a.B + 255) / 256 : 1024;
  hipLaunchKernelGGL(hilbert_trivial_kernel, dim3(grid), dim3(256), 0, st, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? EMD_OK : emdk::hip_fail(e, "hilbert kernel launch");
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

}  // namespace emdh

extern "C" {

int emd_hilbert(const double* imfs, const int64_t* row_base, int max_imfs, const int* n_imfs,
                const int64_t* offsets, int N, int B, int min_n, int max_n, int n_fft, double dt,
                double* amp, double* phase, double* freq, void* stream) {
  using namespace emdh;
  if (B < 0) return emdk::fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (!offsets && N < 0) return emdk::fail(EMD_ERR_VALUE, "N must be >= 0, got %d", N);
  if (offsets && (min_n < 0 || max_n < min_n))
    return emdk::fail(EMD_ERR_VALUE, "need 0 <= min_n <= max_n, got %d and %d", min_n, max_n);
  if (!row_base && max_imfs < 1)
    return emdk::fail(EMD_ERR_VALUE, "the default row layout needs max_imfs >= 1");
  if (!(dt > 0.0) || !std::isfinite(dt))
    return emdk::fail(EMD_ERR_VALUE, "dt must be positive and finite");
  if (n_fft < -1 || (n_fft > 0 && !is_pow2(n_fft)))
    return emdk::fail(EMD_ERR_VALUE,
                      "n_fft must be 0 (the row's own length), -1 (the next power of two) or "
                      "a power of two, got %d", n_fft);
  const int longest = offsets ? max_n : N;
  if (n_fft > 0 && n_fft < longest)
    return emdk::fail(EMD_ERR_VALUE, "n_fft %d is shorter than the longest row (%d)", n_fft,
                      longest);
  if (n_fft > (1 << kMaxLogF) || (n_fft < 0 && longest > (1 << kMaxLogF)))
    return emdk::fail(EMD_ERR_VALUE, "transform length above %d", 1 << kMaxLogF);
  if (n_fft == 0) {
    bool ex;
    const bool ok = offsets ? (longest <= (1 << (kMaxLogF - 1)) ||
                               (min_n == max_n && transform_class(longest, 0, &ex) != 0))
                            : (longest == 0 || transform_class(longest, 0, &ex) != 0);
    if (!ok)
      return emdk::fail(EMD_ERR_VALUE,
                        "rows longer than %d samples need n_fft = -1 or a power of two "
                        "(got rows of up to %d)", 1 << (kMaxLogF - 1), longest);
  }
  if (B == 0 || longest == 0 || (!amp && !phase && !freq)) return EMD_OK;
  if (!imfs || !n_imfs) return emdk::fail(EMD_ERR_VALUE, "null imfs or n_imfs");

  Rows a{imfs, row_base, max_imfs, n_imfs, offsets, N, B, n_fft,
         1.0 / (2.0 * M_PI * dt), amp, phase, freq};
  const hipStream_t st = (hipStream_t)stream;
  bool exact;
  if (!offsets || n_fft > 0) {
    // one transform class for every row
    const int c = transform_class(offsets ? 1 : N, n_fft, &exact);
    if (c == kTrivial) return launch_trivial(a, st);
    return launch_class(a, c, exact, true, st);
  }
  // every class a length in [min_n, max_n] can fall in
  const int lo = min_n < 1 ? 1 : min_n;
  int rc;
  if (lo <= 2 && (rc = launch_trivial(a, st))) return rc;
  if (max_n < 3) return EMD_OK;
  const int lo3 = lo < 3 ? 3 : lo;
  if (n_fft < 0) {
    for (int c = ceil_log2(lo3); c <= ceil_log2(max_n); ++c)
      if ((rc = launch_class(a, c, false, false, st))) return rc;
    return EMD_OK;
  }
  // the DFT of the row's own length: power-of-two lengths directly ...
  for (int c = ceil_log2(lo3 < 4 ? 4 : lo3); (1 << c) <= max_n; ++c)
    if ((rc = launch_class(a, c, false, false, st))) return rc;
  // ... the others as convolutions on 2^ceil(log2(2 len - 1)) points
  if (min_n != max_n || (max_n & (max_n - 1)) != 0)
    for (int c = ceil_log2(2 * lo3 - 1); c <= ceil_log2(2 * max_n - 1); ++c)
      if ((rc = launch_class(a, c, true, false, st))) return rc;
  return EMD_OK;
}

}  // extern "C"
