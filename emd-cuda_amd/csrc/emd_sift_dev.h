// emd_sift_dev.h -- batched Empirical Mode Decomposition (PyEMD EMD() defaults) for gfx950.
//
// One workgroup decomposes one trace at a time.  The current IMF lives in LDS (8 bytes per
// sample) and is processed in blocks of 64 samples, one sample per lane (two adjacent
// samples per lane in 128-sample blocks for the envelope pass): extrema are found with
// wave64 ballots and compacted by popcount over a block-prefix scan, the same masks give
// every sample its envelope interval, and every reduction is a 64-lane butterfly per block
// followed by a perfect binary tree over the block sums.  The knot tile (position, value,
// second derivative and 1/h per knot slot, both envelopes) follows the IMF in LDS; a sift
// whose knots outgrow the launch's LDS tile uses the workgroup's own tile region in global
// memory instead (same code, same arithmetic).
//
// Reference counterparts (SeismicFlow/emd-cuda kernels.py):
//   detect()          <- find_extrema_serial / find_extrema_parallel   kernels.py:115-349
//   mirror()          <- mirror_one_side                                kernels.py:360-558
//   setup/solve/eval  <- notaknot_cubic_spline_block                    kernels.py:585-736
//   stop statistics   <- check_imf_block                                kernels.py:742-821
//   decompose()       <- emd_batch_kernel body                          kernels.py:834-1080
//   envelope kernel   <- envelope_batch_kernel                          kernels.py:1100-1166
//
// Canonical arithmetic: every floating-point sum is the perfect binary tree over
// next_pow2(N) samples in index order (butterfly inside each aligned 64-sample block, then
// a perfect tree over the blocks; zero leaves past N do not change it), the tridiagonal
// solve's operation order depends only on row indices, and the build uses
// -ffp-contract=off with the few fused multiply-adds written out.  Results are therefore
// bit-identical for every workgroup size, LDS tile, tile placement, queue order and kernel
// (plain or persistent).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "emd_sift.h"

// Arrays carved out of a workgroup's LDS or global region: plain pointers, or bounds-checked
// spans in a host emulation build (EMD_BOUNDS_CHECK, emd_bounds.h).
#ifdef EMD_BOUNDS_CHECK
#include "emd_bounds.h"
#else
#define EMD_ARR(T) T*
#define EMD_SPAN(T, ptr, count, name) (ptr)
#endif

namespace emdk {

constexpr double kEnergyRatioThr = 0.2;
constexpr double kStdThr = 0.2;
constexpr double kSvarThr = 0.001;
constexpr double kTotalPowerThr = 0.005;
constexpr double kRangeThr = 0.001;
constexpr double kPowerFloor = 1e-10;
constexpr double kResidueAtol = 1e-8;
constexpr double kSixth = 1.0 / 6.0;
constexpr int kWorkBytesPerBlock = 568;  // 64 fp64 + 4 fp64 + 2 u64 + 2 int per block

enum SplineKind { kConst = 0, kLinear = 1, kThreePoint = 2, kNotAKnot = 3 };

// Knot and sample positions.  IndexAxis: PyEMD's simple mode, where a sample's position
// is its index (int, exact).  TimeAxis: the caller's time axis, already normalised as the
// reference does ((T - T[0]) / min(diff T), emd.py:83-88), read from global memory; the
// reference places knots at T[idx] and mirrored knots at 2 T[sym] - T[idx]
// (kernels.py:471-472, 528-552) and evaluates the envelopes at T[i] (kernels.py:703).
struct IndexAxis {
  using Pos = int;
  static constexpr int kSlot = kSlotBytes;
  __device__ __forceinline__ int at(int i) const { return i; }
  // the position of sample i, where i may be one past the last sample (not used there)
  __device__ __forceinline__ int at_upto(int i, int) const { return i; }
};
struct TimeAxis {
  using Pos = double;
  static constexpr int kSlot = kSlotBytesTime;
  const double* t;  // this trace's N times
  __device__ __forceinline__ double at(int i) const { return t[i]; }
  __device__ __forceinline__ double at_upto(int i, int last) const { return t[min(i, last)]; }
};

template <class P>
struct TileT {
  EMD_ARR(double) Y;  // knot values
  EMD_ARR(double) M;  // chord slope of the interval that starts at each knot slot,
                      // overwritten by the solve with the second derivative at the slot
  EMD_ARR(double) H;  // 1/h of the interval that starts at each knot slot
  EMD_ARR(P) X;       // knot positions (sample index, or time; mirrored knots may lie
                      // before the first sample or after the last)
};
using Tile = TileT<int>;

template <class P>
__device__ __forceinline__ TileT<P> make_tile(double* base, int slots) {
  TileT<P> t;
  t.Y = EMD_SPAN(double, base, slots, "tile.Y");
  t.M = EMD_SPAN(double, base + slots, slots, "tile.M");
  t.H = EMD_SPAN(double, base + 2 * slots, slots, "tile.H");
  P* x = reinterpret_cast<P*>(base + 3 * slots);
  t.X = EMD_SPAN(P, x, slots, "tile.X");
  return t;
}

template <int WG>
struct SiftShared {
  static constexpr int NW = WG / 64;
  double red[NW][6];
  int nz_w[NW];
  int plat_w[NW];
  int bad_w[NW];
  int app_w[NW];
  int tot_max, tot_min;
  // mirror results
  int k0[2], nk[2], kind[2], mtot;
  int xf[2], xl[2];  // first and last true knot position per side
  int bad, bad_m;    // wrong-signed knots: all of them / the mirrored ones
  double p3[2][4];
  double carry[2][3];  // sum of squares, max, min of the IMF left by sift n (slot n & 1)
  int queue_idx;
};

// Per-trace LDS work area (ahead of the knot tile): the IMF and per-64-sample-block
// bookkeeping, laid out for the launch's longest trace.
struct Work {
  EMD_ARR(double) imf;     // [nblk*64] current IMF
  EMD_ARR(double) bsum;    // [4][nblk] block partial sums: the sift's three sums (rows 0-2),
                           // the accept pass's |residual| (row 0) and an IMF start's sum of
                           // squares (row 3)
  EMD_ARR(uint64_t) mmax;  // [nblk] maxima masks of the last detection
  EMD_ARR(uint64_t) mmin;  // [nblk] minima masks
  EMD_ARR(int) pmax;       // [nblk] exclusive prefix counts of maxima over blocks
  EMD_ARR(int) pmin;       // [nblk]
  int nbcap;               // blocks of the layout
};

__device__ __forceinline__ Work make_work(double* base, int nbcap) {
  Work w;
  w.imf = EMD_SPAN(double, base, 64 * nbcap, "work.imf");
  w.bsum = EMD_SPAN(double, base + 64 * nbcap, 4 * nbcap, "work.bsum");
  uint64_t* masks = reinterpret_cast<uint64_t*>(base + 68 * nbcap);
  w.mmax = EMD_SPAN(uint64_t, masks, nbcap, "work.mmax");
  w.mmin = EMD_SPAN(uint64_t, masks + nbcap, nbcap, "work.mmin");
  int* prefix = reinterpret_cast<int*>(masks + 2 * nbcap);
  w.pmax = EMD_SPAN(int, prefix, nbcap, "work.pmax");
  w.pmin = EMD_SPAN(int, prefix + nbcap, nbcap, "work.pmin");
  w.nbcap = nbcap;
  return w;
}

__device__ __forceinline__ int ceil_div(int a, int b) { return (a + b - 1) / b; }
// Values that are equal in every lane by construction (read back from LDS after a
// barrier): make them scalar so branches on them are uniform jumps.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}
// threadIdx.x behind an empty asm: each phase recomputes its lane-derived values and
// addresses (a few VALU) instead of keeping them live, and spilled, across the whole
// decomposition.
__device__ __forceinline__ int fresh_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// 1/q by the hardware reciprocal and one Newton step (the hardware value is good to about
// 2^-50 relative, so one step lands within an ulp or two; only the std-test sum uses it)
__device__ __forceinline__ double recip(double q) {
  const double y = __builtin_amdgcn_rcp(q);
  const double e = __builtin_fma(-q, y, 1.0);
  return __builtin_fma(y, e, y);
}

__device__ __forceinline__ int round_half_even_mid(int s) {
  // np.round(s / 2.0) for integer s >= 0
  int f = s >> 1;
  if ((s & 1) == 0) return f;
  return (f & 1) ? f + 1 : f;
}

// Cross-lane moves for the butterflies below (all 64 lanes active).  DPP covers the
// levels inside 16-lane rows: quad permutations for lane^1 and lane^2, then row mirrors
// (lane i <- 7-i, lane i <- 15-i), which pair each lane with the other half of its group;
// once a group holds one value in every lane that is the same pairing as lane^4 / lane^8.
// The last two levels either read the four rows' values into scalars (max/min) or use
// row broadcasts (sums).
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror
constexpr int kDppBcast15 = 0x142;    // row_bcast:15 (lane 15 of each row -> the next row)
constexpr int kDppBcast31 = 0x143;    // row_bcast:31 (lane 31 -> rows 2 and 3)
constexpr int kDppRowShr = 0x110;     // + n: row_shr:n (lane i <- lane i-n of its row)
constexpr int kDppWaveShl1 = 0x130;   // wave_shl:1 (lane i <- lane i+1)
constexpr int kDppWaveShr1 = 0x138;   // wave_shr:1 (lane i <- lane i-1)

// A lane whose source lane does not exist (wave shifts past the ends) reads 0; lanes of
// rows outside ROWS are left undefined, which only the row broadcasts of
// wave_tree_sum_last use (their result is read from lane 63 alone).  With no `old`
// operand the compiler emits no register initialisation per move.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, true);
  return __longlong_as_double(
      (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}
__device__ __forceinline__ double lane_d(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders the LDS accesses of one wave: its lanes' earlier writes are visible to its
// later reads (a wave's LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same for a knot tile that may live in global memory (the workgroup's global tile):
// workgroup-scope fences also order the wave's global stores before its later loads.
__device__ __forceinline__ void tile_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Perfect binary tree over the 64 lanes (lane order): level d adds the partner at lane^d,
// i.e. ((x0+x1)+(x2+x3))+... (fp addition commutes, so both partners form bit-identical
// sums).  The first four levels run on every lane; the last two are row broadcasts, so
// only lane 63 holds the sum (row 3 adds row 2's value, then rows 0+1 from lane 31).
__device__ __forceinline__ double wave_tree_sum_last(double v) {
  v += dpp_d<kDppXor1>(v);
  v += dpp_d<kDppXor2>(v);
  v += dpp_d<kDppHalfMirror>(v);
  v += dpp_d<kDppMirror>(v);
  v += dpp_d<kDppBcast15, 0xA>(v);
  v += dpp_d<kDppBcast31, 0x8>(v);
  return v;
}
// Three wave_tree_sum_last trees, level by level (independent moves fill each other's
// DPP read hazards).
__device__ __forceinline__ void wave_tree_sum3_last(double& a, double& b, double& c) {
#define EMD_LEVEL(CTRL, ROWS)        \
  {                                  \
    const double ta = dpp_d<CTRL, ROWS>(a); \
    const double tb = dpp_d<CTRL, ROWS>(b); \
    const double tc = dpp_d<CTRL, ROWS>(c); \
    a += ta;                         \
    b += tb;                         \
    c += tc;                         \
  }
  EMD_LEVEL(kDppXor1, 0xF)
  EMD_LEVEL(kDppXor2, 0xF)
  EMD_LEVEL(kDppHalfMirror, 0xF)
  EMD_LEVEL(kDppMirror, 0xF)
  EMD_LEVEL(kDppBcast15, 0xA)
  EMD_LEVEL(kDppBcast31, 0x8)
#undef EMD_LEVEL
}
// Inclusive prefix sum over the 64 lanes: row-local shifts by 1, 2, 4 and 8 (a lane whose
// source lies outside its 16-lane row adds 0), then the row totals by broadcast (rows
// outside ROWS add 0: the `old` operand).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned dpp_u(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v) {
  v += dpp_u<kDppRowShr + 1>(v);
  v += dpp_u<kDppRowShr + 2>(v);
  v += dpp_u<kDppRowShr + 4>(v);
  v += dpp_u<kDppRowShr + 8>(v);
  v += dpp_u<kDppBcast15, 0xA>(v);
  v += dpp_u<kDppBcast31, 0xC>(v);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_d<kDppXor1>(v));
  v = fmax(v, dpp_d<kDppXor2>(v));
  v = fmax(v, dpp_d<kDppHalfMirror>(v));
  v = fmax(v, dpp_d<kDppMirror>(v));
  return fmax(fmax(lane_d(v, 0), lane_d(v, 16)), fmax(lane_d(v, 32), lane_d(v, 48)));
}
__device__ __forceinline__ double wave_min(double v) {
  v = fmin(v, dpp_d<kDppXor1>(v));
  v = fmin(v, dpp_d<kDppXor2>(v));
  v = fmin(v, dpp_d<kDppHalfMirror>(v));
  v = fmin(v, dpp_d<kDppMirror>(v));
  return fmin(fmin(lane_d(v, 0), lane_d(v, 16)), fmin(lane_d(v, 32), lane_d(v, 48)));
}
// =======================================================================================
// Extrema detection.  Semantics of PyEMD _find_extrema_simple: strict extrema where
// d1*d2 < 0; plateau extrema at the half-to-even rounded midpoint of d==0 runs (first run
// dropped when it starts at d-index 1, a run reaching the end dropped, d[-1] wraparound);
// zero-crossing count = #(S[i]*S[i+1] < 0) + #maximal runs of exact zeros.
// Knot counts are clamped to max_extrema per side with overflow when a count reaches it
// (reference rule, kernels.py:196-203, 318-344).
// =======================================================================================
struct DetectResult {
  int tm, tn;      // true counts (unclamped)
  int nm, nn;      // counts used for the envelopes
  int nzer;
  int in_lds;      // the knots of this detection are in the LDS tile (else the global one)
  int ovf;
};

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (uint64_t(1) << lane) - 1; }

// Lane masks straight from the compare (one v_cmp into a scalar register pair; the masks
// are then combined on the scalar unit).  Predicates: LLVM's fcmp / icmp codes.
constexpr int kFOeq = 1, kFOgt = 2, kFOlt = 4, kFUne = 14, kISlt = 40;
template <int PRED>
__device__ __forceinline__ uint64_t fmask(double a) {
  return __builtin_amdgcn_fcmp(a, 0.0, PRED);
}
__device__ __forceinline__ uint64_t lt_mask(int a, int b) {
  return __builtin_amdgcn_sicmp(a, b, kISlt);
}
// Lanes 0..n-1 of a 64-lane block (n <= 0: none, n >= 64: all), on the scalar unit.
__device__ __forceinline__ uint64_t lanes_below(int n) {
  return n <= 0 ? uint64_t(0) : n >= 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1;
}
// A 64-bit value that is equal in every lane, as a scalar.
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (uint64_t(hi) << 32) | lo;
}
// Set bits of a scalar lane mask below this lane.
__device__ __forceinline__ int mbcnt(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Serial scan on one lane over a staged copy of the IMF: emit(idx) for every maximum
// (want_max) or every minimum, in index order.
template <class Emit>
__device__ __forceinline__ void serial_extrema(const double* S, int N, bool want_max,
                                               Emit&& emit) {
  if (N < 3) return;
  const double d_last = S[N - 1] - S[N - 2];  // d[-1]
  double dprev = 0.0;
  int run_a = -1;
  double run_dbefore = 0.0;
  bool first_run_seen = false;
  for (int k = 0; k <= N - 2; ++k) {
    const double dk = S[k + 1] - S[k];
    if (k >= 1 && dprev * dk < 0.0 && (dprev > 0.0) == want_max) emit(k);
    if (dk == 0.0) {
      if (run_a < 0) {
        run_a = k;
        run_dbefore = (k >= 1) ? dprev : d_last;
      }
    } else if (run_a >= 0) {
      const int a = run_a, b = k;  // d[a..b-1] == 0, S[a..b] equal
      run_a = -1;
      const bool first = !first_run_seen;
      first_run_seen = true;
      if (!(first && a == 1)) {
        const int mid = round_half_even_mid(a + b);
        if (run_dbefore > 0.0) {
          if (dk < 0.0 && want_max) emit(mid);
        } else {
          if (dk > 0.0 && !want_max) emit(mid);
        }
      }
    }
    dprev = dk;
  }
  // a run still open here reaches the end of the signal and is dropped
}

// Blocks first, first + step, ... of the calling wave.
template <class AX>
__device__ __forceinline__ void compact_knots(const Work& wk, int N,
                                              const TileT<typename AX::Pos>& t, const AX& ax,
                                              int nm, int nn, int pad, uint64_t& bad,
                                              int first, int step) {
  const int lane = fresh_tid() & 63;
  const int nblk = ceil_div(N, 64);
  const int tb0 = pad, tb1 = 3 * pad + nm;
  for (int blk = first; blk < nblk; blk += step) {
    const int i = blk * 64 + lane;
    // every load of the block is issued before the first use (one LDS round trip); the
    // masks stay in vector registers, so no scalar read waits for them
    const uint64_t bm = wk.mmax[blk], bn = wk.mmin[blk];
    const int pm = wk.pmax[blk], pn = wk.pmin[blk];
    const double y = wk.imf[i];
    const int km = pm + mbcnt(bm);
    const int kn = pn + mbcnt(bn);
    const bool sm = ((bm >> lane) & 1ull) && km < nm;  // this lane stores a maximum
    const bool sn = ((bn >> lane) & 1ull) && kn < nn;  // or a minimum (max_extrema clamp)
    if (sm) {
      t.X[tb0 + km] = ax.at(i);
      t.Y[tb0 + km] = y;
    }
    if (sn) {
      t.X[tb1 + kn] = ax.at(i);
      t.Y[tb1 + kn] = y;
    }
    bad |= (__ballot(sm) & fmask<kFOlt>(y)) | (__ballot(sn) & fmask<kFOgt>(y));
  }
}

template <int WG>
__device__ __forceinline__ DetectResult detect(const Work& wk, int N, SiftShared<WG>& sh,
                                               const SiftParams& P) {
  constexpr int NW = WG / 64;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  const int pad = P.nbsym;
  const int nblk = ceil_div(N, 64);
  const EMD_ARR(double) imf = wk.imf;

  // ---- pass A: per-block lane masks of maxima/minima, zero crossings, plateaus --------
  int nz = 0, plat = 0;
  // lanes whose sample is in the trace (i < N) and has a successor (i <= N - 2): every
  // lane of every block but the last (N > 64 (nblk - 1)), so the last block's masks are
  // made once, off the loop's scalar path
  const int last = nblk - 1;
  const uint64_t in_last = lanes_below(N - last * 64), inner_last = lanes_below(N - 1 - last * 64);
  for (int blk = w; blk < nblk; blk += NW) {
    const int i = blk * 64 + lane;
    // all three loads stay inside the IMF array (the neighbours are clamped at its ends);
    // values outside the trace are masked off below
    const double x = imf[i];
    const double prev = imf[i >= 1 ? i - 1 : 0];
    const double next = imf[min(i + 1, 64 * wk.nbcap - 1)];
    const double d2 = next - x;
    const double d1 = x - prev;
    // extrema also need a predecessor (i >= 1)
    const uint64_t in = blk == last ? in_last : ~uint64_t(0);
    const uint64_t inner = blk == last ? inner_last : ~uint64_t(0);
    const uint64_t inner1 = blk == 0 ? inner & ~uint64_t(1) : inner;
    const uint64_t ext = fmask<kFOlt>(d1 * d2) & inner1;
    const uint64_t bm = ext & fmask<kFOgt>(d1), bn = ext & fmask<kFOlt>(d1);
    const uint64_t zc = fmask<kFOlt>(x * next) & inner;
    const uint64_t zr = fmask<kFOeq>(x) & in & (fmask<kFUne>(prev) | (blk == 0 ? 1ull : 0ull));
    nz += __popcll(zc) + __popcll(zr);
    plat |= (fmask<kFOeq>(d2) & inner) != 0ull;
    if (lane == 0) {
      wk.mmax[blk] = bm;
      wk.mmin[blk] = bn;
      // the block's counts (maxima | minima << 16) for the scan below
      wk.pmax[blk] = __popcll(bm) | (__popcll(bn) << 16);
    }
  }
  if (lane == 0) {
    sh.nz_w[w] = nz;
    sh.plat_w[w] = plat;
  }
  __syncthreads();
  int nzt = 0, plt = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    nzt += sh.nz_w[q];
    plt |= sh.plat_w[q];
  }
  if (uni(plt)) {
    // plateau present: one lane applies the full run rules and rewrites the maxima
    // masks, and one lane of another wave (the same lane when there is one wave) the
    // minima masks; the two scans are independent
    constexpr int kMinLane = NW > 1 ? 64 : 0;
    if (tid == 0) {
      for (int b = 0; b < nblk; ++b) wk.mmax[b] = 0ull;
      serial_extrema(imf, N, true, [&](int idx) { wk.mmax[idx >> 6] |= uint64_t(1) << (idx & 63); });
    }
    if (tid == kMinLane) {
      for (int b = 0; b < nblk; ++b) wk.mmin[b] = 0ull;
      serial_extrema(imf, N, false, [&](int idx) { wk.mmin[idx >> 6] |= uint64_t(1) << (idx & 63); });
    }
    __syncthreads();
  }
  // ---- exclusive prefix of per-block counts (wave 0; up to 4 blocks per lane) ---------
  if (w == 0) {
    // One wave works while the workgroup waits at the barrier below; raised issue
    // priority lets it go ahead of the other workgroups' waves sharing its SIMD.
    __builtin_amdgcn_s_setprio(2);
    const int per = ceil_div(nblk, 64);
    // block counts: written by pass A, or recounted from the masks the plateau rules
    // rewrote
    const bool recount = uni(plt) != 0;
    auto count = [&](int b) -> unsigned {
      return recount ? (__popcll(wk.mmax[b]) | (unsigned(__popcll(wk.mmin[b])) << 16))
                     : unsigned(wk.pmax[b]);
    };
    unsigned loc = 0;
    for (int k = 0; k < per; ++k) {
      const int b = lane * per + k;
      if (b < nblk) loc += count(b);
    }
    const unsigned incl = wave_incl_scan(loc);
    unsigned run = incl - loc;
    for (int k = 0; k < per; ++k) {
      const int b = lane * per + k;
      if (b < nblk) {
        const unsigned c = count(b);
        wk.pmax[b] = int(run & 0xffffu);
        wk.pmin[b] = int(run >> 16);
        run += c;
      }
    }
    if (lane == 63) {
      sh.tot_max = int(incl & 0xffffu);
      sh.tot_min = int(incl >> 16);
    }
    __builtin_amdgcn_s_setprio(0);
  }
  __syncthreads();
  DetectResult R;
  R.nzer = uni(nzt);
  R.tm = uni(sh.tot_max);
  R.tn = uni(sh.tot_min);
  R.nm = min(R.tm, P.max_extrema);
  R.nn = min(R.tn, P.max_extrema);
  R.ovf = (R.tm >= P.max_extrema || R.tn >= P.max_extrema);
  R.in_lds = (4 * pad + R.nm + R.nn <= P.tile_slots);
  // the knots are compacted into the tile by the sift that uses them (build_envelopes)
  return R;
}

// =======================================================================================
// Boundary mirroring (one lane).  PyEMD _prepare_points_simple with T = arange(N).
// =======================================================================================
// A mirror list of PyEMD _prepare_points_simple: a slice of one side's true extrema (in
// ascending or reversed order) plus at most one edge sample (index 0 or N-1) placed before
// or after the slice.  Kept as a descriptor so the single mirroring lane needs no arrays.
struct KDesc {
  int from, to;  // slice [from, to) of the side's true extrema
  int rev;       // 1: elements ind[to-1], ..., ind[from]
  int sp;        // 0: no edge sample; 1: sample N-1 first (right lists); 2: sample 0 last
                 // (left lists)
};

__device__ __forceinline__ KDesc kd(int from, int to, int rev) {
  KDesc d;
  d.from = from;
  d.to = to;
  d.rev = rev;
  d.sp = 0;
  return d;
}
__device__ __forceinline__ int kd_len(const KDesc& d) {
  return max(0, d.to - d.from) + (d.sp ? 1 : 0);
}
// slot (within the side's true-extrema array) of element q, or -1 for the edge sample
__device__ __forceinline__ int kd_slot(const KDesc& d, int q) {
  const int nr = max(0, d.to - d.from);
  if (d.sp == 1) {
    if (q == 0) return -1;
    --q;
  } else if (d.sp == 2 && q == nr) {
    return -1;
  }
  return d.rev ? d.to - 1 - q : d.from + q;
}
// The first and last nbsym+1 true knots of each side, which are all the mirror reads:
// lanes 0..4F-1 of the mirroring wave each hold one (position, value), and the
// wave-uniform mirror logic fetches them with readlane.
struct KnotEnds {
  int xi;      // this lane's cached knot position
  double yv;   // this lane's cached knot value
  int F;       // knots cached at each end of a side
  int n[2];    // true knots per side
  int last;    // N - 1: the right edge sample
  double S0, SN;  // edge sample values
};

// Position of set bit j (counting from 0 at the lowest) of m; j < popcount(m).
__device__ __forceinline__ int nth_bit64(uint64_t m, int j) {
  int pos = 0;
#pragma unroll
  for (int half = 32; half >= 1; half >>= 1) {
    const uint64_t low = m & ((uint64_t(1) << half) - 1);
    const int c = __popcll(low);
    if (j >= c) {
      j -= c;
      m >>= half;
      pos += half;
    } else {
      m = low;
    }
  }
  return pos;
}

// The end knots straight from the extrema masks and the block prefix counts of the last
// detection (the compaction that writes the tile runs at the same time): lane
// L = side * 2F + half * F + k caches knot slot clamp(half ? n - F + k : k, 0, n - 1) of
// its side.  Knot g of a side lies in the last block whose exclusive prefix is <= g (a
// ballot over the blocks, whose prefixes do not decrease), at the set bit g - prefix of
// the block's mask.  Three LDS round trips: prefixes, the lane's block, its sample.
__device__ __forceinline__ KnotEnds ends_from_masks(const Work& wk, int nblk, int nm, int nn,
                                                    int F, int N, double S0, double SN) {
  KnotEnds c;
  c.last = N - 1;
  c.S0 = S0;
  c.SN = SN;
  c.F = F;
  c.n[0] = nm;
  c.n[1] = nn;
  c.xi = 0;
  c.yv = 0.0;
  const int lane = fresh_tid() & 63;
  int myblk = 0, myg = 0, myside = 0;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    const int pm = b < nblk ? wk.pmax[b] : INT_MAX, pn = b < nblk ? wk.pmin[b] : INT_MAX;
    // target L: side, first or last F, k-th of those
    auto target = [&](int L, int FF) {
      const int side = L >= 2 * FF ? 1 : 0;
      const int r = L - side * 2 * FF, half = r >= FF ? 1 : 0, k = r - half * FF;
      const int n = side ? nn : nm;
      const int g = max(0, min(half ? n - FF + k : k, n - 1));
      const uint64_t le = __ballot((side ? pn : pm) <= g);
      if (le != 0ull && lane == L) {
        myblk = b0 + 63 - __builtin_clzll(le);
        myg = g;
        myside = side;
      }
    };
    if (F == 3) {
      // nbsym = 2, the default: the targets' sides and slots are constants
#pragma unroll
      for (int L = 0; L < 12; ++L) target(L, 3);
    } else {
      for (int L = 0; L < 4 * F; ++L) target(L, F);
    }
  }
  if (lane < 4 * F && (myside ? nn : nm) > 0) {
    const uint64_t m = myside ? wk.mmin[myblk] : wk.mmax[myblk];
    const int pre = myside ? wk.pmin[myblk] : wk.pmax[myblk];
    c.xi = myblk * 64 + nth_bit64(m, myg - pre);
    c.yv = wk.imf[c.xi];
  }
  return c;
}
// lane holding true knot `slot` of `side` (slot within the first or last F)
__device__ __forceinline__ int ends_lane(const KnotEnds& c, int side, int slot) {
  const int n = side ? c.n[1] : c.n[0];
  const int k = slot < c.F ? slot : c.F + slot - (n - c.F);
  return side * 2 * c.F + k;
}
__device__ __forceinline__ int ends_x(const KnotEnds& c, int side, int slot) {
  return __builtin_amdgcn_readlane(c.xi, ends_lane(c, side, slot));
}
__device__ __forceinline__ double ends_y(const KnotEnds& c, int side, int slot) {
  return lane_d(c.yv, ends_lane(c, side, slot));
}

__device__ __forceinline__ int kd_idx(const KDesc& d, const KnotEnds& c, int side, int q) {
  const int s = kd_slot(d, q);
  return s < 0 ? (d.sp == 1 ? c.last : 0) : ends_x(c, side, s);
}
__device__ __forceinline__ double kd_val(const KDesc& d, const KnotEnds& c, int side, int q) {
  const int s = kd_slot(d, q);
  return s < 0 ? (d.sp == 1 ? c.SN : c.S0) : ends_y(c, side, s);
}

// Knot counts of one side after PyEMD's removal of duplicate adjacent positions.
template <class AX>
__device__ __forceinline__ void side_counts(const KnotEnds& c, const AX& ax, int side, int n,
                                            int tb, const KDesc& L, int lsym, const KDesc& R,
                                            int rsym, int& nl, int& rs, int& k0, int& nk,
                                            int& kind) {
  const int ln = kd_len(L), rn = kd_len(R);
  nl = ln;
  if (ln > 0 && 2 * ax.at(lsym) - ax.at(kd_idx(L, c, side, ln - 1)) == ax.at(ends_x(c, side, 0)))
    --nl;
  rs = (rn > 0 &&
        2 * ax.at(rsym) - ax.at(kd_idx(R, c, side, 0)) == ax.at(ends_x(c, side, n - 1)))
           ? 1
           : 0;
  k0 = tb - nl;
  nk = nl + n + (rn - rs);
  kind = nk >= 4 ? kNotAKnot : nk == 3 ? kThreePoint : nk == 2 ? kLinear : kConst;
}

// The four mirror lists (left/right x maxima/minima) are written in one parallel step:
// lane q takes element q % R of list q / R (R = nbsym + 1 entries at most), so every
// mirrored knot is gathered, checked and written at once.  Left knots sit at positions
// 2*lsym - idx; a left knot equal in position to the first true knot is the same sample
// (same value) and is dropped, which is what PyEMD's delete of equal adjacent positions
// does; likewise on the right.  Right knots go after the true ones.
template <class AX>
__device__ __forceinline__ void finish_lists(const TileT<typename AX::Pos>& t, const AX& ax,
                                             const KnotEnds& c, int tb0, int tb1, int nm, int nn,
                                             const KDesc& lmax, const KDesc& lmin,
                                             const KDesc& rmax, const KDesc& rmin, int lsym,
                                             int rsym, int* k0, int* nk, int* kind, int& bad) {
  constexpr int kR = kMaxNbsym + 1;
  const int lane = fresh_tid() & 63;
  int nl[2], rs[2];
  side_counts(c, ax, 0, nm, tb0, lmax, lsym, rmax, rsym, nl[0], rs[0], k0[0], nk[0], kind[0]);
  side_counts(c, ax, 1, nn, tb1, lmin, lsym, rmin, rsym, nl[1], rs[1], k0[1], nk[1], kind[1]);
  // list of this lane: 0 left maxima, 1 right maxima, 2 left minima, 3 right minima
  const int list = lane / kR, e = lane - list * kR;
  const int side = list >> 1;
  const bool right = (list & 1) != 0;
  KDesc d;  // field-wise select (a per-lane choice of a whole struct would go to scratch)
#define EMD_PICK(f) (list == 0 ? +lmax.f : list == 1 ? +rmax.f : list == 2 ? +lmin.f : +rmin.f)
  d.from = EMD_PICK(from);
  d.to = EMD_PICK(to);
  d.rev = EMD_PICK(rev);
  d.sp = EMD_PICK(sp);
#undef EMD_PICK
  const bool act = list < 4 && e < kd_len(d);
  const int slot = act ? kd_slot(d, e) : 0;
  const int src = ends_lane(c, side, max(slot, 0));
  const int gx = __shfl(c.xi, src);
  const double gy = __shfl(c.yv, src);
  const int idx = slot < 0 ? (d.sp == 1 ? c.last : 0) : gx;
  const double val = slot < 0 ? (d.sp == 1 ? c.SN : c.S0) : gy;
  const bool viol = act && (side == 0 ? (val < 0.0) : (val > 0.0));
  bad |= (__ballot(viol) != 0ull) ? 1 : 0;
  if (act) {
    const int n = side ? nn : nm, tb = side ? tb1 : tb0;
    const int nls = side ? nl[1] : nl[0], rss = side ? rs[1] : rs[0];
    if (!right && e < nls) {
      t.X[tb - nls + e] = 2 * ax.at(lsym) - ax.at(idx);
      t.Y[tb - nls + e] = val;
    } else if (right && e >= rss) {
      t.X[tb + n + e - rss] = 2 * ax.at(rsym) - ax.at(idx);
      t.Y[tb + n + e - rss] = val;
    }
  }
}

// PyEMD cubic_spline_3pts: the natural cubic spline through the three knots from slot q0,
// in slope form, solved in closed form (coefficients a1, b1, a2, b2 of its two intervals).
template <class P>
__device__ __forceinline__ void three_point_coeffs(const TileT<P>& t, int q0, double* p3) {
  const double x0 = t.X[q0], x1 = t.X[q0 + 1], x2 = t.X[q0 + 2];
  const double y0 = t.Y[q0], y1 = t.Y[q0 + 1], y2 = t.Y[q0 + 2];
  const double h1 = x1 - x0, h2 = x2 - x1;
  const double dy1 = y1 - y0, dy2 = y2 - y1;
  const double i1 = 1.0 / h1, i2 = 1.0 / h2;
  const double v1 = 3.0 * dy1 * i1 * i1;
  const double v3 = 3.0 * dy2 * i2 * i2;
  const double k1 = (v1 + v3) / (3.0 * (i1 + i2));
  const double kk0 = (v1 - i1 * k1) / (2.0 * i1);
  const double kk2 = (v3 - i2 * k1) / (2.0 * i2);
  p3[0] = kk0 * h1 - dy1;   // a1
  p3[1] = -k1 * h1 + dy1;   // b1
  p3[2] = k1 * h2 - dy2;    // a2
  p3[3] = -kk2 * h2 + dy2;  // b2
}

// Boundary mirroring, run wave-uniformly by one wave (its results and knot writes are the
// same in every lane; lane 0 writes).  PyEMD _prepare_points_simple (positions from the
// axis: T = arange(N) in PyEMD's simple mode); the reference's mirror_one_side
// (kernels.py:360-558) follows the same branch table.  xf, xl: sample indices of the first
// and last true knot of each side.
template <class AX>
__device__ __forceinline__ void mirror(const TileT<typename AX::Pos>& t, const AX& ax,
                                       const Work& wk, int N, int nm, int nn, int nb, double S0,
                                       double SN, int* k0, int* nk, int* kind, int& bad, int* xf,
                                       int* xl) {
  const int tb0 = nb, tb1 = 3 * nb + nm;
  const KnotEnds c = ends_from_masks(wk, ceil_div(N, 64), nm, nn, nb + 1, N, S0, SN);
  const int Xm0 = ends_x(c, 0, 0), XmL = ends_x(c, 0, nm - 1);
  const int Xn0 = ends_x(c, 1, 0), XnL = ends_x(c, 1, nn - 1);
  xf[0] = Xm0;
  xl[0] = XmL;
  xf[1] = Xn0;
  xl[1] = XnL;
  const double Ym0 = ends_y(c, 0, 0), YmL = ends_y(c, 0, nm - 1);
  const double Yn0 = ends_y(c, 1, 0), YnL = ends_y(c, 1, nn - 1);
  KDesc lmax, lmin, rmax, rmin;
  int lsym, rsym;
  if (Xm0 < Xn0) {
    if (uni(S0 > Yn0 ? 1 : 0)) {
      lmax = kd(1, min(nm, nb + 1), 1);
      lmin = kd(0, min(nn, nb), 1);
      lsym = Xm0;
    } else {
      lmax = kd(0, min(nm, nb), 1);
      lmin = kd(0, min(nn, nb - 1), 1);
      lmin.sp = 2;
      lsym = 0;
    }
  } else {
    if (uni(S0 < Ym0 ? 1 : 0)) {
      lmax = kd(0, min(nm, nb), 1);
      lmin = kd(1, min(nn, nb + 1), 1);
      lsym = Xn0;
    } else {
      lmax = kd(0, min(nm, nb - 1), 1);
      lmax.sp = 2;
      lmin = kd(0, min(nn, nb), 1);
      lsym = 0;
    }
  }
  if (XmL < XnL) {
    if (uni(SN < YmL ? 1 : 0)) {
      rmax = kd(max(nm - nb, 0), nm, 1);
      rmin = kd(max(nn - nb - 1, 0), nn - 1, 1);
      rsym = XnL;
    } else {
      rmax = kd(max(nm - nb + 1, 0), nm, 1);
      rmax.sp = 1;
      rmin = kd(max(nn - nb, 0), nn, 1);
      rsym = N - 1;
    }
  } else {
    if (uni(SN > YnL ? 1 : 0)) {
      rmax = kd(max(nm - nb - 1, 0), nm - 1, 1);
      rmin = kd(max(nn - nb, 0), nn, 1);
      rsym = XmL;
    } else {
      rmax = kd(max(nm - nb, 0), nm, 1);
      rmin = kd(max(nn - nb + 1, 0), nn, 1);
      rmin.sp = 1;
      rsym = N - 1;
    }
  }
  // PyEMD: an empty list falls back to the whole ascending index array (for nbsym >= 1
  // this happens only when that side has a single extremum)
  if (kd_len(lmin) == 0) lmin = kd(0, nn, 0);
  if (kd_len(rmin) == 0) rmin = kd(0, nn, 0);
  if (kd_len(lmax) == 0) lmax = kd(0, nm, 0);
  if (kd_len(rmax) == 0) rmax = kd(0, nm, 0);

  // mirrored points must reach the signal edges; otherwise mirror about the edge
  if (2 * ax.at(lsym) - ax.at(kd_idx(lmin, c, 1, 0)) > ax.at(0) ||
      2 * ax.at(lsym) - ax.at(kd_idx(lmax, c, 0, 0)) > ax.at(0)) {
    if (lsym == Xm0) lmax = kd(0, min(nm, nb), 1);
    else lmin = kd(0, min(nn, nb), 1);
    lsym = 0;
  }
  if (2 * ax.at(rsym) - ax.at(kd_idx(rmin, c, 1, kd_len(rmin) - 1)) < ax.at(N - 1) ||
      2 * ax.at(rsym) - ax.at(kd_idx(rmax, c, 0, kd_len(rmax) - 1)) < ax.at(N - 1)) {
    if (rsym == XmL) rmax = kd(max(nm - nb, 0), nm, 1);
    else rmin = kd(max(nn - nb, 0), nn, 1);
    rsym = N - 1;
  }
  finish_lists(t, ax, c, tb0, tb1, nm, nn, lmax, lmin, rmax, rmin, lsym, rsym, k0, nk, kind,
               bad);
}

// =======================================================================================
// Not-a-knot spline solves, one wave per envelope.  Each side's system has m = knots - 2
// rows (M_0 and M_{n-1} eliminated) and is solved by a partition method: lane l owns a
// contiguous chunk [s, e] of rows; a backward sweep over the chunk's interior expresses
// x_{s+1} and x_{e-1} in terms of the chunk ends x_s and x_e; the chunk end rows then form
// a tridiagonal system in (x_s, x_e) over all chunks, which an odd-even step reduces to
// one unknown per lane and parallel cyclic reduction across lanes solves with shuffles;
// a forward substitution then fills the chunk's interior.  The sweep parks its factors in
// the chunk's own knot slots of H and M (no registers or global memory per row, so any
// chunk length takes the same path), and the substitution overwrites them with the
// solution and 1/h.  The systems are diagonally dominant, so no pivoting is needed.  The
// solution goes to M by knot slot; the end second derivatives are then recovered.
// =======================================================================================
// Per interval k (knots k, k+1) of a not-a-knot side: H[k] = 1/h and, until the solve
// overwrites it, M[k] = the chord slope (y_{k+1} - y_k) / h as (y_{k+1} - y_k) * (1/h).
template <class P>
__device__ __forceinline__ void interval_setup(const TileT<P>& t, int k) {
  const double ih = recip(double(t.X[k + 1] - t.X[k]));
  t.H[k] = ih;
  t.M[k] = (t.Y[k + 1] - t.Y[k]) * ih;
}

// Row j of the system of a side whose knots start at slot k0: a x_{j-1} + b x_j + c x_{j+1} = d,
// from the chord slopes sl, sr of the intervals left and right of knot k0+j+1.
template <class P>
__device__ __forceinline__ void nak_row_sl(const TileT<P>& t, int k0, int m, int j, double sl,
                                           double sr, double& a, double& b, double& c,
                                           double& d) {
  const int i = k0 + j + 1;
  const double xa = t.X[i - 1], xb = t.X[i], xc = t.X[i + 1];
  const double hl = xb - xa, hr = xc - xb;
  const double rhs = 6.0 * (sr - sl);
  if (j == 0) {
    a = 0.0;
    b = hl + 2.0 * hr;
    c = hr - hl;
    d = hr * rhs / (hl + hr);
  } else if (j == m - 1) {
    a = hl - hr;
    b = 2.0 * hl + hr;
    c = 0.0;
    d = hl * rhs / (hl + hr);
  } else {
    a = hl;
    b = 2.0 * (hl + hr);
    c = hr;
    d = rhs;
  }
}

// The same row with the chord slopes the interval setup left in M.
template <class P>
__device__ __forceinline__ void nak_row(const TileT<P>& t, int k0, int m, int j, double& a,
                                        double& b, double& c, double& d) {
  const int i = k0 + j + 1;
  nak_row_sl(t, k0, m, j, t.M[i - 1], t.M[i], a, b, c, d);
}

// The end rows of chunk [s, e] after eliminating its interior rows.  The chunk sweeps
// backward from its end, x_j = p_j - q_j x_{j-1} - r_j x_e, carrying a rolling window of
// knots and slopes; q_j goes to H[k0+j] and p_j to M[k0+j+1] (slots of this chunk that no
// later row reads; the slope left of the end row, which the sweep overwrites, is kept in a
// register), and the unrolled recurrence for x_{e-1} is summed on the way.  r_{s+1}, where
// the substitution starts, is parked in M[k0+s+1] once the start row has read that slope.
struct EndRows {
  double sa, sb, sc, sd;  // start row: sa x_{s-1} + sb x_s + sc x_e = sd
  double ea, eb, ec, ed;  // end row:   ea x_s + eb x_e + ec x_{e+1} = ed
};

template <class P>
__device__ __forceinline__ EndRows chunk_ends(const TileT<P>& t, int k0, int m, int s, int e) {
  const int len = e - s - 1;
  const double sle = t.M[k0 + e];
  P xm = t.X[k0 + e], xr = t.X[k0 + e + 1];
  double sr = sle;
  double p = 0.0, pq = 0.0, pr = 0.0, Q = 1.0, su = 0.0, sw = 0.0;
  // each row's left knot and slope are loaded one row ahead (the last such load is of the
  // start row's, always a slot of this chunk)
  P xl_n = t.X[k0 + e - 1];
  double sl_n = t.M[k0 + e - 1];
  for (int q = 0; q < len; ++q) {
    const int j = e - 1 - q;  // knots k0+j .. k0+j+2
    const P xl = xl_n;
    const double sl = sl_n;
    xl_n = t.X[k0 + j - 1];
    sl_n = t.M[k0 + j - 1];
    const double hl = double(xm - xl), hr = double(xr - xm);
    const double b = 2.0 * (hl + hr), d = 6.0 * (sr - sl);
    const double rb = recip(q == 0 ? b : b - hr * pq);
    p = (q == 0 ? d : d - hr * p) * rb;
    pr = (q == 0 ? hr : -(hr * pr)) * rb;
    pq = hl * rb;
    t.H[k0 + j] = pq;
    t.M[k0 + j + 1] = p;
    su += Q * p;
    sw += Q * pr;
    Q = -(Q * pq);
    xr = xm;
    xm = xl;
    sr = sl;
  }
  // x_{s+1} = p - pq x_s - pr x_e and x_{e-1} = su + Q x_s - sw x_e (with no interior
  // rows the start row's x_{s+1} is x_e itself)
  double a, b, c, d;
  EndRows r;
  nak_row(t, k0, m, s, a, b, c, d);
  r.sa = a;
  r.sb = b - c * pq;
  r.sc = len > 0 ? -(c * pr) : c;
  r.sd = d - c * p;
  if (len > 0) t.M[k0 + s + 1] = pr;
  nak_row_sl(t, k0, m, e, sle, t.M[k0 + e + 1], a, b, c, d);
  r.ea = a * Q;
  r.eb = b - a * sw;
  r.ec = c;
  r.ed = d - a * su;
  return r;
}

// The interior of chunk [s, e] once x_s and x_e are known: forward substitution over the
// parked (q_j, p_j), x_j = p_j - q_j x_{j-1} - r_j x_e, with r_j rebuilt from r_{s+1} by
// r_{j+1} = -r_j (b_j - c_j q_{j+1}) / c_j; x_j replaces p_j in M[k0+j+1], 1/h of interval
// k0+j goes back into H[k0+j], and x_s, x_e are stored last.
template <class P>
__device__ __forceinline__ void chunk_interior(const TileT<P>& t, int k0, int s, int e,
                                               double xs, double xe) {
  const int len = e - s - 1;
  if (len > 0) {
    double x = xs, r = t.M[k0 + s + 1], qj = t.H[k0 + s + 1];
    P xl = t.X[k0 + s + 1], xm = t.X[k0 + s + 2];
    double ihl = recip(double(xm - xl));
    // each row's right knot, parked p and next q are loaded one row ahead (the last such
    // loads reach knot k0+e+2 and slot k0+e+1, still this side's)
    P xr_n = t.X[k0 + s + 3];
    double p_n = t.M[k0 + s + 2], q_n = t.H[k0 + s + 2];
    for (int q = 0; q < len; ++q) {
      const int j = s + 1 + q;  // knots k0+j .. k0+j+2
      const P xr = xr_n;
      const double pj = p_n, qn = q_n;
      xr_n = t.X[k0 + j + 3];
      p_n = t.M[k0 + j + 2];
      q_n = t.H[k0 + j + 2];
      const double hl = double(xm - xl), hr = double(xr - xm);
      const double ihr = recip(hr);
      x = pj - qj * x - r * xe;
      t.M[k0 + j + 1] = x;
      if (q + 1 < len) {
        r = -(r * ((2.0 * (hl + hr) - hr * qn) * ihr));
        qj = qn;
      }
      t.H[k0 + j] = ihl;
      ihl = ihr;
      xl = xm;
      xm = xr;
    }
  }
  t.M[k0 + 1 + s] = xs;
  t.M[k0 + 1 + e] = xe;
}

// Solve one side's system (knots from slot k0, m rows) with the calling wave; every lane
// of the wave must call it.  Writes M[k0+1 .. k0+m].
template <class P>
__device__ __forceinline__ void spline_system_wave(const TileT<P>& t, int side, int k0, int m,
                                                   int lane) {
  const int K = min(64, m / 2);  // chunks of >= 2 rows
  // m / K and m % K without a division: K = 64 for m >= 128; otherwise K = m / 2, which
  // leaves chunks of 2 (plus one of 3 when m is odd), except m = 3 (one chunk of 3)
  const int base = m >= 128 ? (m >> 6) : (m == 3 ? 3 : 2);
  const int rem = m >= 128 ? (m & 63) : (m == 3 ? 0 : (m & 1));
  const bool act = lane < K;
  const int s = act ? lane * base + min(lane, rem) : 0;
  const int e = act ? s + base + (lane < rem ? 1 : 0) - 1 : 0;
  // chunk end rows after eliminating the interior (identity rows on idle lanes)
  double rsa = 0.0, rsb = 1.0, rsc = 0.0, rsd = 0.0;
  double rea = 0.0, reb = 1.0, rec = 0.0, red = 0.0;
  if (act) {
    const EndRows r = chunk_ends(t, k0, m, s, e);
    rsa = r.sa;
    rsb = r.sb;
    rsc = r.sc;
    rsd = r.sd;
    rea = r.ea;
    reb = r.eb;
    rec = r.ec;
    red = r.ed;
  }
  // odd-even step: eliminate x_s of this chunk and of the next from the e-row
  const double nsa = dpp_d<kDppWaveShl1>(rsa), nsb = dpp_d<kDppWaveShl1>(rsb);
  const double nsc = dpp_d<kDppWaveShl1>(rsc), nsd = dpp_d<kDppWaveShl1>(rsd);
  const bool has_next = lane + 1 < K;
  const double k1 = rea * recip(rsb);
  const double k2 = has_next ? rec * recip(nsb) : 0.0;
  const double A = -(k1 * rsa);
  const double B = reb - k1 * rsc - (has_next ? k2 * nsa : 0.0);
  const double C = has_next ? -(k2 * nsc) : 0.0;
  const double D = red - k1 * rsd - (has_next ? k2 * nsd : 0.0);
  // parallel cyclic reduction over the chunk e-rows, each row scaled to a unit diagonal
  // (alpha x_{l-dd} + x_l + gamma x_{l+dd} = delta), so a level exchanges three values
  // each way and divides once; distance 1 by whole-wave DPP shifts
  double ib = recip(B);
  double al = A * ib, ga = C * ib, de = D * ib;
  for (int dd = 1; dd < K; dd <<= 1) {
    double au, gu, du, ad, gd, ddn;
    if (dd == 1) {
      au = dpp_d<kDppWaveShr1>(al);
      gu = dpp_d<kDppWaveShr1>(ga);
      du = dpp_d<kDppWaveShr1>(de);
      ad = dpp_d<kDppWaveShl1>(al);
      gd = dpp_d<kDppWaveShl1>(ga);
      ddn = dpp_d<kDppWaveShl1>(de);
    } else {
      au = __shfl_up(al, dd);
      gu = __shfl_up(ga, dd);
      du = __shfl_up(de, dd);
      ad = __shfl_down(al, dd);
      gd = __shfl_down(ga, dd);
      ddn = __shfl_down(de, dd);
    }
    const bool hu = lane >= dd, hd = lane + dd < 64;
    // neighbours past the ends contribute nothing: their coefficients are taken as 0
    const double au0 = hu ? au : 0.0, gu0 = hu ? gu : 0.0, du0 = hu ? du : 0.0;
    const double ad0 = hd ? ad : 0.0, gd0 = hd ? gd : 0.0, dd0 = hd ? ddn : 0.0;
    const double nA = -(al * au0);
    const double nC = -(ga * gd0);
    const double nB = __builtin_fma(-ga, ad0, __builtin_fma(-al, gu0, 1.0));
    const double nD = __builtin_fma(-ga, dd0, __builtin_fma(-al, du0, de));
    ib = recip(nB);
    al = nA * ib;
    ga = nC * ib;
    de = nD * ib;
  }
  const double xe = de;
  const double xprev = dpp_d<kDppWaveShr1>(xe);
  if (act) {
    const double xs = (rsd - (lane > 0 ? rsa * xprev : 0.0) - rsc * xe) * recip(rsb);
    chunk_interior(t, k0, s, e, xs, xe);
  }
}

// Not-a-knot solve of one side by the calling wave (every lane calls): interval setup of
// the intervals that touch mirrored knots (the true-knot intervals were set up while the
// mirror ran), the solve, then M_0 and M_{n-1} from the not-a-knot relations.
template <class P>
__device__ __forceinline__ void solve_side(const TileT<P>& t, int side, int k0, int nk, int tb,
                                           int n, int lane) {
  const int nl = tb - k0;  // left mirrored knots
  const int nmir = nk - n;  // intervals with a mirrored end: nl on the left, the rest right
  if (lane < nmir) interval_setup(t, lane < nl ? k0 + lane : tb + n - 1 + (lane - nl));
  tile_sync();
  spline_system_wave(t, side, k0, nk - 2, lane);
  tile_sync();
  if (lane == 0) {
    // all loads first, so they are in flight together
    const double x0 = t.X[k0], x1 = t.X[k0 + 1], x2 = t.X[k0 + 2];
    const double M1 = t.M[k0 + 1], M2 = t.M[k0 + 2];
    const double xa = t.X[k0 + nk - 3], xb = t.X[k0 + nk - 2], xc = t.X[k0 + nk - 1];
    const double Mb = t.M[k0 + nk - 2], Ma = t.M[k0 + nk - 3];
    const double h0 = x1 - x0, h1 = x2 - x1;
    const double hnm3 = xb - xa, hnm2 = xc - xb;
    t.M[k0] = ((h0 + h1) * M1 - h0 * M2) / h1;
    t.M[k0 + nk - 1] = ((hnm3 + hnm2) * Mb - hnm2 * Ma) / hnm3;
  }
}

template <int WG, class P>
__device__ __forceinline__ void spline_solve(const TileT<P>& t, const SiftShared<WG>& sh,
                                             const DetectResult& D, int nb) {
  const int tid = fresh_tid(), lane = tid & 63, wv = uni(tid >> 6);
  const int mtot = uni(sh.mtot);
  if (mtot > 0) {
    // wave 0 solves the maxima envelope and wave 1 the minima (wave 0 both when WG = 64);
    // their knot slots are disjoint
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int owner = WG >= 128 ? side : 0;
      if (wv == owner && uni(sh.kind[side]) == kNotAKnot) {
        __builtin_amdgcn_s_setprio(2);  // one wave per side; the others wait at the barrier
        solve_side(t, side, uni(sh.k0[side]), uni(sh.nk[side]), side ? 3 * nb + D.nm : nb,
                   side ? D.nn : D.nm, lane);
        __builtin_amdgcn_s_setprio(0);
      }
    }
  }
  __syncthreads();
}

// Envelope value at a sample for the closed-form spline kinds (3 knots: PyEMD's natural
// cubic; 2 knots: straight line; 1 knot: constant).
template <class P>
__device__ __forceinline__ double eval_small(const TileT<P>& t, int k0, int kind,
                                             const double* p3, double tt) {
  const double X0 = t.X[k0], Y0 = t.Y[k0];
  if (kind == kConst) return Y0;
  const double X1 = t.X[k0 + 1], Y1 = t.Y[k0 + 1];
  if (kind == kLinear) return Y0 + (Y1 - Y0) / (X1 - X0) * (tt - X0);
  const double X2 = t.X[k0 + 2], Y2 = t.Y[k0 + 2];
  if (tt < X1) {
    const double u = (tt - X0) / (X1 - X0);
    const double w = 1.0 - u;
    return w * Y0 + u * Y1 + u * w * (p3[0] * w + p3[1] * u);
  }
  const double u = (tt - X1) / (X2 - X1);
  const double w = 1.0 - u;
  return w * Y1 + u * Y2 + u * w * (p3[2] * w + p3[3] * u);
}

// Interval of not-a-knot side `side` holding sample i (block blk, lane) from the masks of
// the detection that produced the knots: true extrema at or before i plus the left knots,
// corrected for left/right mirrored knots that straddle i (at most nbsym steps).
template <int WG, class P>
__device__ __forceinline__ int interval_of(const TileT<P>& t, const SiftShared<WG>& sh, int side,
                                           const Work& wk, int blk, int lane, P pi, int tb) {
  const int k0 = sh.k0[side], nk = sh.nk[side];
  const uint64_t m = side == 0 ? wk.mmax[blk] : wk.mmin[blk];
  const int pre = side == 0 ? wk.pmax[blk] : wk.pmin[blk];
  const int le = pre + __popcll(m & (lanemask_lt(lane) | (uint64_t(1) << lane)));
  int j = (tb - k0) + le - 1;
  j = max(0, min(j, nk - 2));
  while (j < nk - 2 && t.X[k0 + j + 1] <= pi) ++j;
  while (j > 0 && t.X[k0 + j] > pi) --j;
  return j;
}

// Envelope of side `side` at sample i, whose position is pi.
template <int WG, class P>
__device__ __forceinline__ double envelope_at(const TileT<P>& t, const SiftShared<WG>& sh,
                                              int side, const Work& wk, int blk, int lane, P pi,
                                              int tb) {
  const int kind = sh.kind[side];
  const int k0 = sh.k0[side];
  if (kind != kNotAKnot) return eval_small(t, k0, kind, sh.p3[side], double(pi));
  const int j = interval_of(t, sh, side, wk, blk, lane, pi, tb) + k0;
  const P x0 = t.X[j], x1 = t.X[j + 1];
  const double h = double(x1 - x0);
  const double A = double(x1 - pi) * t.H[j];
  const double B = double(pi - x0) * t.H[j];
  return A * t.Y[j] + B * t.Y[j + 1] +
         ((A * A * A - A) * t.M[j] + (B * B * B - B) * t.M[j + 1]) * (h * h * kSixth);
}

// Perfect binary tree over nblk block partials (zero-padded to a power of two), computed
// identically by every lane of every wave.
__device__ __forceinline__ double blocks_tree(const double* v, int nblk) {
  const int lane = fresh_tid() & 63;
  int P = 1;
  while (P < nblk) P <<= 1;
  const int per = P > 64 ? P / 64 : 1;  // leaves per lane (1, 2 or 4)
  const int last = nblk - 1;            // loads clamp to the last leaf; selects zero-pad
  double a = 0.0;
  if (per == 1) {
    const double x0 = v[min(lane, last)];
    a = lane < nblk ? x0 : 0.0;
  } else if (per == 2) {
    const int b = 2 * lane;
    const double x0 = v[min(b, last)], x1 = v[min(b + 1, last)];
    a = (b < nblk ? x0 : 0.0) + (b + 1 < nblk ? x1 : 0.0);
  } else {
    const int b = 4 * lane;
    const double x0 = v[min(b, last)], x1 = v[min(b + 1, last)];
    const double x2 = v[min(b + 2, last)], x3 = v[min(b + 3, last)];
    a = ((b < nblk ? x0 : 0.0) + (b + 1 < nblk ? x1 : 0.0)) +
        ((b + 2 < nblk ? x2 : 0.0) + (b + 3 < nblk ? x3 : 0.0));
  }
  // lanes past the leaves hold zeros: the 64-leaf tree equals the tree over P leaves
  return lane_d(wave_tree_sum_last(a), 63);
}

// blocks_tree of three rows of block partials (row stride `stride`) at once: the same
// leaves, pairs and order as three blocks_tree calls, with the three DPP trees interleaved.
__device__ __forceinline__ void blocks_tree3(const double* v, int stride, int nblk, double& s0,
                                             double& s1, double& s2) {
  const int lane = fresh_tid() & 63;
  int P = 1;
  while (P < nblk) P <<= 1;
  const int per = P > 64 ? P / 64 : 1;  // leaves per lane (1, 2 or 4)
  const int last = nblk - 1;            // loads clamp to the last leaf; selects zero-pad
  double a[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double* w = v + r * stride;
    if (per == 1) {
      const double x0 = w[min(lane, last)];
      a[r] = lane < nblk ? x0 : 0.0;
    } else if (per == 2) {
      const int b = 2 * lane;
      const double x0 = w[min(b, last)], x1 = w[min(b + 1, last)];
      a[r] = (b < nblk ? x0 : 0.0) + (b + 1 < nblk ? x1 : 0.0);
    } else {
      const int b = 4 * lane;
      const double x0 = w[min(b, last)], x1 = w[min(b + 1, last)];
      const double x2 = w[min(b + 2, last)], x3 = w[min(b + 3, last)];
      a[r] = ((b < nblk ? x0 : 0.0) + (b + 1 < nblk ? x1 : 0.0)) +
             ((b + 2 < nblk ? x2 : 0.0) + (b + 3 < nblk ? x3 : 0.0));
    }
  }
  wave_tree_sum3_last(a[0], a[1], a[2]);
  s0 = lane_d(a[0], 63);
  s1 = lane_d(a[1], 63);
  s2 = lane_d(a[2], 63);
}

struct SiftStats {
  double sq_new, sq_diff, sq_old, std_sum, old_max, old_min;
};

// Wave-uniform description of one not-a-knot side for the envelope pass.
struct SideEval {
  int k0;   // first knot slot
  int nl;   // left (mirrored) knots before the first true knot
  int nk2;  // knots - 2: the last interval index
  int xf;   // positions of the first and last true knot: a sample in [xf, xl) lies
  int xl;   // between two true knots, in the interval the extrema masks give
};

template <int WG>
__device__ __forceinline__ SideEval side_eval(const SiftShared<WG>& sh, int side, int tb) {
  SideEval e;
  e.k0 = uni(sh.k0[side]);
  e.nl = tb - e.k0;
  e.nk2 = uni(sh.nk[side]) - 2;
  e.xf = uni(sh.xf[side]);
  e.xl = uni(sh.xl[side]);
  return e;
}

// Not-a-knot envelope of one side at samples i and i+1 (bits p and p+1 of 64-sample block
// b).  The interval comes from the extrema masks of the detection that produced the knots:
// true extrema at or before the sample plus the left knots.  Only samples before the
// first or after the last true knot (where mirrored knots may straddle them) walk to the
// enclosing interval; for every other sample the walk would not move.  pi0, pi1: the
// positions of samples i and i+1.
template <int SIDE, class P>
__device__ __forceinline__ void envelope_pair(const TileT<P>& t, const SideEval& s,
                                              const Work& wk, int b, int p, int i, P pi0, P pi1,
                                              double& e0, double& e1) {
  const uint64_t m = SIDE == 0 ? wk.mmax[b] : wk.mmin[b];
  const int pre = SIDE == 0 ? wk.pmax[b] : wk.pmin[b];
  const uint64_t le_mask = (uint64_t(2) << p) - 1;  // bits 0..p (p is even, so <= 62)
  const int le0 = pre + __popcll(m & le_mask);
  const int le1 = le0 + (int)((m >> (p + 1)) & 1);  // p is even, so p + 1 <= 63
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ii = i + q;
    const P pi = q ? pi1 : pi0;
    int j = s.nl + (q ? le1 : le0) - 1;
    if (!(ii >= s.xf && ii < s.xl)) {
      j = max(0, min(j, s.nk2));
      while (j < s.nk2 && t.X[s.k0 + j + 1] <= pi) ++j;
      while (j > 0 && t.X[s.k0 + j] > pi) --j;
    }
    const int jj = s.k0 + j;
    const P xa = t.X[jj], xb = t.X[jj + 1];
    const double B = double(pi - xa) * t.H[jj];
    const double A = 1.0 - B;
    const double hh = double(xb - xa);
    // A y0 + B y1 + h^2/6 ((A^3 - A) M0 + (B^3 - B) M1), with fused multiply-adds
    const double ca = __builtin_fma(A * A, A, -A), cb = __builtin_fma(B * B, B, -B);
    const double curv = __builtin_fma(ca, t.M[jj], cb * t.M[jj + 1]);
    const double lin = __builtin_fma(A, t.Y[jj], B * t.Y[jj + 1]);
    const double e = __builtin_fma(curv, hh * hh * kSixth, lin);
    if (q == 0) e0 = e; else e1 = e;
  }
}

// One 128-sample block c of the envelope pass: both envelopes at samples i and i+1 of this
// lane, the IMF update, and the lane's pair sums of the three stop statistics (in sq).
// NAK: both sides are known to be not-a-knot.
template <int WG, bool NAK, class AX>
__device__ __forceinline__ void envelope_block(const Work& wk,
                                               const TileT<typename AX::Pos>& t, const AX& ax,
                                               const SiftShared<WG>& sh, int N, int c, int lane,
                                               const SideEval& s0, const SideEval& s1,
                                               int kind0, int kind1, double* sq, double& nmax,
                                               double& nmin) {
  using P = typename AX::Pos;
  const int i = c * 128 + 2 * lane;
  sq[0] = sq[1] = sq[2] = 0.0;
  if (i < N) {
    const int b = 2 * c + (lane >> 5), p = (2 * lane) & 63;
    const P pi0 = ax.at(i), pi1 = ax.at_upto(i + 1, N - 1);
    double emax0, emax1, emin0, emin1;
    if (NAK || kind0 == kNotAKnot) {
      envelope_pair<0>(t, s0, wk, b, p, i, pi0, pi1, emax0, emax1);
    } else {
      emax0 = eval_small(t, sh.k0[0], kind0, sh.p3[0], double(pi0));
      emax1 = eval_small(t, sh.k0[0], kind0, sh.p3[0], double(pi1));
    }
    if (NAK || kind1 == kNotAKnot) {
      envelope_pair<1>(t, s1, wk, b, p, i, pi0, pi1, emin0, emin1);
    } else {
      emin0 = eval_small(t, sh.k0[1], kind1, sh.p3[1], double(pi0));
      emin1 = eval_small(t, sh.k0[1], kind1, sh.p3[1], double(pi1));
    }
    // i + 1 < 64 * nbcap always (i is even and < N), so the pair read stays in the work area
    const double old0 = wk.imf[i], old1 = wk.imf[i + 1];
    const double nw0 = __builtin_fma(-0.5, emax0 + emin0, old0);
    const double d0 = nw0 - old0;
    const double r0 = d0 * recip(nw0);
    sq[0] = nw0 * nw0;
    sq[1] = d0 * d0;
    sq[2] = r0 * r0;
    double hi = nw0, lo = nw0;
    if (i + 1 < N) {
      const double nw1 = __builtin_fma(-0.5, emax1 + emin1, old1);
      const double d1 = nw1 - old1;
      const double r1 = d1 * recip(nw1);
      sq[0] = __builtin_fma(nw1, nw1, sq[0]);
      sq[1] = __builtin_fma(d1, d1, sq[1]);
      sq[2] = __builtin_fma(r1, r1, sq[2]);
      hi = fmax(nw0, nw1);
      lo = fmin(nw0, nw1);
      wk.imf[i + 1] = nw1;
    }
    wk.imf[i] = nw0;
    nmax = fmax(nmax, hi);
    nmin = fmin(nmin, lo);
  }
}

// Sub-blocks of 128 samples per envelope-pass superblock: a function of N alone, so the
// association of the stop-statistic sums is the same for every launch shape.
__device__ __forceinline__ int sub_blocks(int nblk2) { return min(4, max(1, nblk2 >> 2)); }

// One sift: evaluate both envelopes at every sample, subtract the mean, accumulate the
// PyEMD stop statistics.  The IMF in LDS is updated in place.  Each lane takes two
// adjacent samples of a 128-sample block and adds them; a superblock of K = sub_blocks()
// consecutive 128-sample blocks is one wave's unit, whose lane pair sums are added as a
// perfect tree over the K blocks, then over the 64 lanes; the superblock sums are then
// added as a perfect tree over superblocks (zero-padded).  Every sum therefore has one
// association for a given N, whatever the workgroup size, tile or queue.
// Sift n's imf_old is the IMF that sift n-1 left, bit for bit (the detection in between
// does not touch it), so its sum of squares and its range come from sh.carry[(n-1)&1]
// instead of being recomputed; the IMF's start pass fills carry[0] for n == 1.
template <int WG, class AX>
__device__ __forceinline__ SiftStats envelope_update(const Work& wk,
                                                     const TileT<typename AX::Pos>& t,
                                                     const AX& ax, SiftShared<WG>& sh, int N,
                                                     const DetectResult& D, int nb, int n) {
  constexpr int NW = WG / 64;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  const int nblk2 = ceil_div(N, 128);
  const int K = sub_blocks(nblk2);
  const int nsb = ceil_div(nblk2, K);
  const int kind0 = uni(sh.kind[0]), kind1 = uni(sh.kind[1]);
  const SideEval s0 = side_eval(sh, 0, nb), s1 = side_eval(sh, 1, 3 * nb + D.nm);
  double nmax = -INFINITY, nmin = INFINITY;
  const bool nak = kind0 == kNotAKnot && kind1 == kNotAKnot;
  for (int sb = w; sb < nsb; sb += NW) {
    // a: the running pair sum, first of blocks 0 and 1, then of blocks 2 and 3; z: blocks
    // 0+1 once the second pair starts.  The sums are of squares (never -0), so 0 + v == v
    // and the association stays (v0 + v1) + (v2 + v3) for every K
    double a[3] = {0.0, 0.0, 0.0}, z[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < K; ++q) {
      const int c = sb * K + q;
      if (c >= nblk2) break;
      if (q == 2) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          z[k] = a[k];
          a[k] = 0.0;
        }
      }
      double v[3];
      // the common case (both envelopes not-a-knot) gets a loop without the small kinds
      if (nak)
        envelope_block<WG, true>(wk, t, ax, sh, N, c, lane, s0, s1, kind0, kind1, v, nmax,
                                 nmin);
      else
        envelope_block<WG, false>(wk, t, ax, sh, N, c, lane, s0, s1, kind0, kind1, v, nmax,
                                  nmin);
#pragma unroll
      for (int k = 0; k < 3; ++k) a[k] += v[k];
    }
    double s_new = a[0] + z[0], s_diff = a[1] + z[1], s_std = a[2] + z[2];
    wave_tree_sum3_last(s_new, s_diff, s_std);
    if (lane == 63) {
      wk.bsum[sb] = s_new;
      wk.bsum[wk.nbcap + sb] = s_diff;
      wk.bsum[2 * wk.nbcap + sb] = s_std;
    }
  }
  nmax = wave_max(nmax);
  nmin = wave_min(nmin);
  if (lane == 0) {
    sh.red[w][2] = nmax;
    sh.red[w][3] = nmin;
  }
  __syncthreads();
  SiftStats st;
  blocks_tree3(wk.bsum, wk.nbcap, nsb, st.sq_new, st.sq_diff, st.std_sum);
  double gnx = sh.red[0][2], gnn = sh.red[0][3];
#pragma unroll
  for (int q = 1; q < NW; ++q) {
    gnx = fmax(gnx, sh.red[q][2]);
    gnn = fmin(gnn, sh.red[q][3]);
  }
  {
    const double* c = sh.carry[(n - 1) & 1];
    st.sq_old = c[0];
    st.old_max = c[1];
    st.old_min = c[2];
  }
  // slot n&1 was last read by sift n-1, which every thread has finished (barriers since)
  if (tid == 0) {
    double* c = sh.carry[n & 1];
    c[0] = st.sq_new;
    c[1] = gnx;
    c[2] = gnn;
  }
  return st;
}

__device__ __forceinline__ bool check_imf(const SiftStats& s, int bad) {
  if (bad) return false;
  if (s.sq_new < kPowerFloor) return false;
  const double svar = s.sq_diff / (s.old_max - s.old_min);
  if (svar < kSvarThr) return true;
  if (s.std_sum < kStdThr) return true;
  const double er = s.sq_diff / s.sq_old;
  if (er < kEnergyRatioThr) return true;
  return false;
}

// Knots of detection D into the tile, mirror, interval setup and spline solve (shared by the
// sift and the envelope-only kernel).  Wave 0 mirrors (its end knots come from the extrema
// masks) while the other waves compact the true extrema into the tile; after a barrier every
// wave sets up the intervals between true knots, and waves 0 and 1 solve.
template <int WG, class AX>
__device__ __forceinline__ void build_envelopes(const TileT<typename AX::Pos>& t, const AX& ax,
                                                SiftShared<WG>& sh, const DetectResult& D,
                                                int N, const SiftParams& P, const Work& wk) {
  constexpr int NW = WG / 64;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  if (w == 0) {
    __builtin_amdgcn_s_setprio(2);  // the mirror is the longest single-wave step
    int bad = 0;
    int k0[2], nk[2], kind[2], xf[2], xl[2];
    mirror(t, ax, wk, N, D.nm, D.nn, P.nbsym, wk.imf[0], wk.imf[N - 1], k0, nk, kind, bad, xf,
           xl);
    if (tid == 0) {
      sh.bad_m = bad;
      for (int q = 0; q < 2; ++q) {
        sh.k0[q] = k0[q];
        sh.nk[q] = nk[q];
        sh.kind[q] = kind[q];
        sh.xf[q] = xf[q];
        sh.xl[q] = xl[q];
      }
      const int m0 = kind[0] == kNotAKnot ? nk[0] - 2 : 0;
      const int m1 = kind[1] == kNotAKnot ? nk[1] - 2 : 0;
      sh.mtot = m0 + m1;
    }
    __builtin_amdgcn_s_setprio(0);
  }
  // the true extrema into the knot tile: waves 1.. while wave 0 mirrors (the slots are
  // disjoint); wave 0 itself when it is the only one
  if (NW == 1 || w > 0) {
    // the compaction waves take longer than the mirror: raised issue priority (below the
    // mirror's) lets them go ahead of other workgroups' waves on their SIMDs
    if (NW > 1) __builtin_amdgcn_s_setprio(1);
    uint64_t bad = 0;
    compact_knots(wk, N, t, ax, D.nm, D.nn, P.nbsym, bad, NW == 1 ? 0 : w - 1,
                  NW == 1 ? 1 : NW - 1);
    if (lane == 0) sh.bad_w[w] = bad != 0ull ? 1 : 0;
    if (NW > 1) __builtin_amdgcn_s_setprio(0);
  }
  __syncthreads();
  if (tid == 0) {
    int b = sh.bad_m;
    for (int q = NW == 1 ? 0 : 1; q < NW; ++q) b |= sh.bad_w[q];
    sh.bad = b;
#pragma unroll
    for (int side = 0; side < 2; ++side)
      if (sh.kind[side] == kThreePoint) three_point_coeffs(t, sh.k0[side], sh.p3[side]);
  }
  // interval setup between consecutive true knots of both sides (the mirror does not move
  // them)
  {
    const int nt0 = max(D.nm - 1, 0), nt = nt0 + max(D.nn - 1, 0);
    const int tb0 = P.nbsym, tb1 = 3 * P.nbsym + D.nm;
    for (int q = tid; q < nt; q += WG) interval_setup(t, q < nt0 ? tb0 + q : tb1 + (q - nt0));
  }
  __syncthreads();
  spline_solve<WG>(t, sh, D, P.nbsym);
}

// =======================================================================================
// Whole decomposition of one trace (PyEMD EMD.emd default loop).
// =======================================================================================
struct SignalIO {
  const double* S;
  const double* T;  // the trace's times (time-axis kernels) or null
  int N;
  double* out;      // row k at out + k*N
  int* n_imfs;
  int* overflow;
  int* dbg_n;
  int* dbg_e;
  int* dbg_z;
  int* stats;       // 4 ints or null
};

constexpr int kPrefetch = 4;  // 64-sample blocks per wave whose global loads are batched

// Start of an IMF: imf = S - acc (acc = the running sum of accepted rows, first updated
// with the IMF in LDS when `add`), written to LDS with its sum of squares and range for
// the first sift's imf_old (sh.carry[0]).
template <int WG>
__device__ __forceinline__ void start_imf(const SignalIO& io, SiftShared<WG>& sh, const Work& wk,
                                          double* acc, bool have_acc, bool add) {
  constexpr int NW = WG / 64;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  const int N = io.N, nblk = ceil_div(N, 64);
  // block sums in the fourth row of the work area's sums: the end-condition reduction
  // that precedes this pass may still be reading the first
  EMD_ARR(double) bs = wk.bsum + 3 * wk.nbcap;
  double omax = -INFINITY, omin = INFINITY;
  for (int blk0 = w; blk0 < nblk; blk0 += kPrefetch * NW) {
    double sv[kPrefetch], av[kPrefetch];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int i = (blk0 + u * NW) * 64 + lane;
      sv[u] = i < N ? io.S[i] : 0.0;
      av[u] = (i < N && have_acc) ? acc[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int blk = blk0 + u * NW;
      if (blk >= nblk) break;
      const int i = blk * 64 + lane;
      double lo = 0.0;
      if (i < N) {
        double a = av[u];
        if (add) {
          a = a + wk.imf[i];
          acc[i] = a;
        }
        const double v = sv[u] - a;
        wk.imf[i] = v;
        lo = v * v;
        omax = fmax(omax, v);
        omin = fmin(omin, v);
      }
      lo = wave_tree_sum_last(lo);
      if (lane == 63) bs[blk] = lo;
    }
  }
  omax = wave_max(omax);
  omin = wave_min(omin);
  if (lane == 0) {
    sh.red[w][2] = omax;
    sh.red[w][3] = omin;
  }
  __syncthreads();
  const double sq = blocks_tree(bs, nblk);
  double gmx = sh.red[0][2], gmn = sh.red[0][3];
#pragma unroll
  for (int q = 1; q < NW; ++q) {
    gmx = fmax(gmx, sh.red[q][2]);
    gmn = fmin(gmn, sh.red[q][3]);
  }
  if (tid == 0) {
    sh.carry[0][0] = sq;
    sh.carry[0][1] = gmx;
    sh.carry[0][2] = gmn;
  }
}

// tl: the launch's LDS knot tile; tg: this workgroup's global tile (holds any trace); ax:
// the positions of this trace's samples.
template <int WG, class AX>
__device__ __forceinline__ void decompose(const SignalIO& io, const AX& ax, SiftShared<WG>& sh,
                                          const Work& wk, const TileT<typename AX::Pos>& tl,
                                          const TileT<typename AX::Pos>& tg,
                                          const SiftParams& P) {
  constexpr int NW = WG / 64;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  const int N = io.N;
  const int nblk = ceil_div(N, 64);
  const int nb = P.nbsym;

  if (tid == 0) {
    for (int k = 0; k < P.max_imfs; ++k) {
      io.dbg_n[k] = -1;
      io.dbg_e[k] = -1;
      io.dbg_z[k] = -1;
    }
  }
  if (N > P.max_n || N < 0) {
    // the launch was sized for shorter traces (or the offsets decrease): the LDS work
    // area cannot hold this one; it is reported with overflow = -1 and no rows
    if (tid == 0) {
      *io.n_imfs = 0;
      *io.overflow = -1;
      if (io.stats) io.stats[0] = io.stats[1] = io.stats[2] = io.stats[3] = 0;
    }
    return;
  }
  // The running sum of accepted IMFs (PyEMD's np.sum(IMF, axis=0), rows added in order)
  // is kept in the last output row, which no IMF occupies before the residue is written.
  double* acc = io.out + (int64_t)(P.max_imfs - 1) * N;
  bool have_acc = false;

  int imf_count = 0;
  bool finished = false;
  bool res_in_lds = false;  // the residue is already in the work area (no HBM re-read)
  int extNo_last = -1;
  int overflow = 0;
  int sifts = 0, peak = 0, knot_sum = 0, flags = 0;

  while (!finished && imf_count < P.max_imfs - 1) {
    if (imf_count == 0) start_imf<WG>(io, sh, wk, acc, false, false);
    DetectResult D = detect<WG>(wk, N, sh, P);
    overflow |= D.ovf;
    peak = max(peak, max(D.tm, D.tn));
    int n = 0;
    while (true) {
      ++n;
      if (n >= P.max_iteration) {
        overflow = 1;
        if (tid == 0) {
          io.dbg_n[imf_count] = n;
          io.dbg_e[imf_count] = -1;
          io.dbg_z[imf_count] = -1;
        }
        break;
      }
      const int extNo = D.nm + D.nn;
      if (extNo > 2 && D.nm >= 1 && D.nn >= 1) {
        SiftStats st;
        if (D.in_lds) {
          build_envelopes<WG>(tl, ax, sh, D, N, P, wk);
          st = envelope_update<WG>(wk, tl, ax, sh, N, D, nb, n);
        } else {
          flags |= 2;
          build_envelopes<WG>(tg, ax, sh, D, N, P, wk);
          st = envelope_update<WG>(wk, tg, ax, sh, N, D, nb, n);
        }
        // No barrier before the next detection: what this sift's last steps read (block
        // sums, sh.red, sh.carry, sh.bad, sh.nk) is written next only after the
        // detection's first barrier (sh.bad by thread 0 after its last one), and its
        // own writes (masks, counts, knots) touch nothing those steps read.
        const int bad = uni(sh.bad);
        ++sifts;
        knot_sum += uni(sh.nk[0] + sh.nk[1]);
        D = detect<WG>(wk, N, sh, P);
        overflow |= D.ovf;
        peak = max(peak, max(D.tm, D.tn));
        const int extNo2 = D.nm + D.nn;
        const int nzm2 = D.nzer;
        const bool f2 = abs(extNo2 - nzm2) < 2;
        const bool f1 = uni(check_imf(st, bad) ? 1 : 0) != 0;
        extNo_last = extNo2;
        if (f1 && f2) {
          if (tid == 0) {
            io.dbg_n[imf_count] = n;
            io.dbg_e[imf_count] = extNo2;
            io.dbg_z[imf_count] = nzm2;
          }
          break;
        }
      } else {
        extNo_last = extNo;
        finished = true;
        break;
      }
    }

    // accept the IMF row, then PyEMD end_condition on S - (acc + imf).  Unless this row may
    // be a trend about to be dropped (extNo_last <= 2), the same pass makes acc += imf and
    // starts the next IMF: S - (acc + imf) is its first imf, rounding included, and its
    // sum of squares and range are the next first sift's imf_old statistics.
    const bool fuse = extNo_last > 2;
    double* row = io.out + (int64_t)imf_count * N;
    EMD_ARR(double) bs3 = wk.bsum + 3 * wk.nbcap;
    double tmax = -INFINITY, tmin = INFINITY;
    for (int blk0 = w; blk0 < nblk; blk0 += kPrefetch * NW) {
      // the global loads of kPrefetch blocks are issued before the first use
      double sv[kPrefetch], av[kPrefetch];
#pragma unroll
      for (int u = 0; u < kPrefetch; ++u) {
        const int i = (blk0 + u * NW) * 64 + lane;
        sv[u] = i < N ? io.S[i] : 0.0;
        av[u] = (i < N && have_acc) ? acc[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPrefetch; ++u) {
        const int blk = blk0 + u * NW;
        if (blk >= nblk) break;
        const int i = blk * 64 + lane;
        double la = 0.0, lo = 0.0;
        if (i < N) {
          const double v = wk.imf[i];
          row[i] = v;
          const double an = av[u] + v;
          const double tmp = sv[u] - an;
          la = fabs(tmp);
          tmax = fmax(tmax, tmp);
          tmin = fmin(tmin, tmp);
          if (fuse) {
            acc[i] = an;
            wk.imf[i] = tmp;
            lo = tmp * tmp;
          }
        }
        la = wave_tree_sum_last(la);
        if (fuse) lo = wave_tree_sum_last(lo);
        if (lane == 63) {
          wk.bsum[blk] = la;
          if (fuse) bs3[blk] = lo;
        }
      }
    }
    tmax = wave_max(tmax);
    tmin = wave_min(tmin);
    if (lane == 0) {
      sh.red[w][0] = tmax;
      sh.red[w][1] = tmin;
    }
    __syncthreads();
    {
      const double gsa = blocks_tree(wk.bsum, nblk);
      double gmx = sh.red[0][0], gmn = sh.red[0][1];
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        gmx = fmax(gmx, sh.red[q][0]);
        gmn = fmin(gmn, sh.red[q][1]);
      }
      const int done = ((gmx - gmn) < kRangeThr || gsa < kTotalPowerThr) ? 1 : 0;
      if (uni(done)) finished = true;
      if (fuse && w == 0) {
        const double sq = blocks_tree(bs3, nblk);
        if (lane == 0) {
          sh.carry[0][0] = sq;
          sh.carry[0][1] = gmx;
          sh.carry[0][2] = gmn;
        }
      }
    }
    ++imf_count;
    const bool capped = imf_count >= P.max_imfs - 1;
    if (!finished && capped) flags |= 1;
    const bool leaving = finished || capped;
    if (leaving && extNo_last <= 2) {
      --imf_count;  // the last row was a trend (PyEMD drops it); acc excludes it
      __syncthreads();
    } else if (!fuse) {
      // acc += imf, and the next IMF starts from S - acc (the same value PyEMD's
      // S - sum(IMF) gives, rounding included)
      start_imf<WG>(io, sh, wk, acc, have_acc, true);
      have_acc = true;
    } else {
      have_acc = true;  // the accept pass did both
    }
    if (leaving) {
      // after a fused accept pass the LDS holds S - acc, the residue, rounding included
      res_in_lds = fuse;
      break;  // PyEMD drops the trend row after leaving its loop
    }
  }

  // residue = S - sum(IMFs); appended unless np.allclose(residue, 0) (max|r| <= 1e-8)
  int app = 0;
  for (int i = tid; i < N; i += WG) {
    const double r = res_in_lds ? wk.imf[i] : io.S[i] - (have_acc ? acc[i] : 0.0);
    wk.imf[i] = r;
    app |= !(fabs(r) <= kResidueAtol);
  }
  const int aw = (__ballot(app) != 0ull) ? 1 : 0;
  if (lane == 0) sh.app_w[w] = aw;
  __syncthreads();
  int append = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) append |= sh.app_w[q];
  append = uni(append);
  int n_out = imf_count;
  if (append) {
    double* row = io.out + (int64_t)imf_count * N;
    for (int i = tid; i < N; i += WG) row[i] = wk.imf[i];
    n_out = imf_count + 1;
  }
  if (P.zero_fill) {
    for (int k = n_out; k < P.max_imfs; ++k) {
      double* row = io.out + (int64_t)k * N;
      for (int i = tid; i < N; i += WG) row[i] = 0.0;
    }
  }
  if (tid == 0) {
    *io.n_imfs = n_out;
    *io.overflow = overflow;
    if (io.stats) {
      io.stats[0] = sifts;
      io.stats[1] = peak;
      io.stats[2] = knot_sum;
      io.stats[3] = flags;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ SignalIO signal_io(const Batch& b, int s, int max_imfs) {
  SignalIO io;
  int64_t off, n;
  if (b.offsets) {
    off = b.offsets[s];
    n = b.offsets[s + 1] - off;
  } else {
    n = b.N;
    off = (int64_t)s * n;
  }
  io.S = b.values + off;
  io.T = b.times ? b.times + (b.times_shared ? 0 : off) : nullptr;
  io.N = (int)n;
  // default output layout: trace s's max_imfs rows start at (input offset) x max_imfs
  io.out = b.imfs + (b.out_offsets ? b.out_offsets[s] : off * max_imfs);
  io.n_imfs = b.n_imfs + s;
  io.overflow = b.overflow + s;
  io.dbg_n = b.dbg_n + (int64_t)s * max_imfs;
  io.dbg_e = b.dbg_e + (int64_t)s * max_imfs;
  io.dbg_z = b.dbg_z + (int64_t)s * max_imfs;
  io.stats = b.stats ? b.stats + (int64_t)s * 4 : nullptr;
  return io;
}

// Memory of a workgroup: the LDS work area for traces up to P.max_n samples, the LDS knot
// tile after it, and the workgroup's global knot tile (P.spill_slots slots).
template <class P>
struct WgMem {
  Work wk;
  TileT<P> tl, tg;
};

template <class P>
__device__ __forceinline__ WgMem<P> workgroup_mem(double* lds, char* spill, int64_t stride,
                                                  const SiftParams& Pr) {
  WgMem<P> m;
  const int nbcap = (Pr.max_n + 63) / 64;
  m.wk = make_work(lds, nbcap);
  m.tl = make_tile<P>(lds + (int64_t)nbcap * (kWorkBytesPerBlock / 8), Pr.tile_slots);
  char* region = spill + (int64_t)blockIdx.x * stride;
  m.tg = make_tile<P>(reinterpret_cast<double*>(region), Pr.spill_slots);
  return m;
}

// The positions of trace io's samples.
__device__ __forceinline__ IndexAxis make_axis(IndexAxis, const SignalIO&) { return {}; }
__device__ __forceinline__ TimeAxis make_axis(TimeAxis, const SignalIO& io) {
  TimeAxis a;
  a.t = io.T;
  return a;
}

// One workgroup per trace (grid = traces of this launch).  The reference launch model.
template <int WG, class AX>
__device__ __forceinline__ void plain_body(const Batch& b, const SiftParams& P,
                                           SiftShared<WG>& sh, double* dyn_lds) {
  const WgMem<typename AX::Pos> m =
      workgroup_mem<typename AX::Pos>(dyn_lds, b.spill, b.spill_stride, P);
  const int q = b.base + (int)blockIdx.x;
  const int s = b.order ? b.order[q] : q;
  if ((unsigned)s >= (unsigned)b.total) return;  // a queue entry naming no trace
  const SignalIO io = signal_io(b, s, P.max_imfs);
  decompose<WG>(io, make_axis(AX{}, io), sh, m.wk, m.tl, m.tg, P);
}

// Persistent workgroups pull queue positions from an atomic counter until the b.B queue
// entries are drained.
template <int WG, class AX>
__device__ __forceinline__ void persistent_body(const Batch& b, const SiftParams& P,
                                                SiftShared<WG>& sh, double* dyn_lds) {
  const WgMem<typename AX::Pos> m =
      workgroup_mem<typename AX::Pos>(dyn_lds, b.spill, b.spill_stride, P);
  while (true) {
    if (threadIdx.x == 0) sh.queue_idx = (int)atomicAdd(b.queue, 1u);
    __syncthreads();
    const int idx = __builtin_amdgcn_readfirstlane(sh.queue_idx);
    __syncthreads();
    if (idx >= b.B) break;
    const int q = b.base + idx;
    const int s = b.order ? b.order[q] : q;
    if ((unsigned)s >= (unsigned)b.total) continue;  // a queue entry naming no trace
    const SignalIO io = signal_io(b, s, P.max_imfs);
    decompose<WG>(io, make_axis(AX{}, io), sh, m.wk, m.tl, m.tg, P);
  }
}

template <int WG, class AX = IndexAxis>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void emd_plain_kernel(Batch b, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  plain_body<WG, AX>(b, P, sh, dyn_lds);
}

template <int WG, class AX = IndexAxis>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void emd_persistent_kernel(Batch b, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  persistent_body<WG, AX>(b, P, sh, dyn_lds);
}

// Envelope pass without sifting (diagnostic; reference envelope_batch_kernel).
template <int WG, class AX = IndexAxis>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void emd_envelope_kernel(EnvBatch e, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  using Pos = typename AX::Pos;
  const int tid = fresh_tid(), lane = tid & 63, w = uni(tid >> 6);
  constexpr int NW = WG / 64;
  const WgMem<Pos> m = workgroup_mem<Pos>(dyn_lds, e.spill, e.spill_stride, P);
  const int bidx = blockIdx.x;
  const int N = e.N;
  const double* S = e.values + (int64_t)bidx * N;
  SignalIO io{};
  io.T = e.times ? e.times + (e.times_shared ? 0 : (int64_t)bidx * N) : nullptr;
  const AX ax = make_axis(AX{}, io);
  for (int i = tid; i < N; i += WG) m.wk.imf[i] = S[i];
  __syncthreads();
  DetectResult D = detect<WG>(m.wk, N, sh, P);
  if (tid == 0) {
    e.n_extrema[2 * bidx] = D.nm;
    e.n_extrema[2 * bidx + 1] = D.nn;
  }
  const int extNo = D.nm + D.nn;
  if (extNo <= 2 || D.nm < 1 || D.nn < 1) {
    if (tid == 0) e.degenerate[bidx] = 1;
    return;
  }
  const TileT<Pos>& t = D.in_lds ? m.tl : m.tg;
  if (D.in_lds) build_envelopes<WG>(m.tl, ax, sh, D, N, P, m.wk);
  else build_envelopes<WG>(m.tg, ax, sh, D, N, P, m.wk);
  const int nb = P.nbsym;
  const int nblk = ceil_div(N, 64);
  for (int blk = w; blk < nblk; blk += NW) {
    const int i = blk * 64 + lane;
    if (i < N) {
      const Pos pi = ax.at(i);
      e.max_env[(int64_t)bidx * N + i] = envelope_at(t, sh, 0, m.wk, blk, lane, pi, nb);
      e.min_env[(int64_t)bidx * N + i] =
          envelope_at(t, sh, 1, m.wk, blk, lane, pi, 3 * nb + D.nm);
    }
  }
  for (int side = 0; side < 2; ++side) {
    double* kp = side == 0 ? e.max_kpos : e.min_kpos;
    double* kv = side == 0 ? e.max_kval : e.min_kval;
    const int k0 = sh.k0[side], nk = sh.nk[side];
    for (int k = tid; k < nk && k < e.mbuf; k += WG) {
      kp[(int64_t)bidx * e.mbuf + k] = (double)t.X[k0 + k];
      kv[(int64_t)bidx * e.mbuf + k] = t.Y[k0 + k];
    }
  }
  if (tid == 0) {
    e.degenerate[bidx] = 0;
    e.n_mirrored[2 * bidx] = sh.nk[0];
    e.n_mirrored[2 * bidx + 1] = sh.nk[1];
  }
}

}  // namespace emdk
