// emd_sift_dev.h -- batched Empirical Mode Decomposition (PyEMD EMD() defaults) for gfx950.
//
// One workgroup decomposes one trace at a time.  The per-sample state (current IMF,
// running sum of accepted IMFs) lives in VGPRs: thread t owns the contiguous chunk
// [t*c, t*c+c) with c = next_pow2(ceil(N/WG)).  Only extrema-count-sized data lives in
// LDS: the knot tile (positions, values, and the four tridiagonal arrays of the joint
// max+min envelope solve).  A tile that does not fit the LDS budget of the launch spills
// is deferred: the workgroup abandons the trace, appends it to a retry list, and a retry
// launch with a larger tile (in LDS when it fits, else in global memory) redoes it.
//
// Reference counterparts (SeismicFlow/emd-cuda kernels.py):
//   detect()          <- find_extrema_serial / find_extrema_parallel   kernels.py:115-349
//   mirror()          <- mirror_one_side                                kernels.py:360-558
//   setup/solve/eval  <- notaknot_cubic_spline_block                    kernels.py:585-736
//   stop statistics   <- check_imf_block                                kernels.py:742-821
//   decompose()       <- emd_batch_kernel body                          kernels.py:834-1080
//   envelope kernel   <- envelope_batch_kernel                          kernels.py:1100-1166
//
// Canonical arithmetic: every floating-point reduction is the perfect binary tree over
// next_pow2(N) samples in index order (in-thread pairwise tree, then the lane butterfly,
// then waves in order), the tridiagonal solve's operation order depends only on row
// indices, and the build uses -ffp-contract=off.  Results are therefore bit-identical for
// every workgroup size, chunk size, LDS tile size, spill decision, queue order and kernel
// (plain or persistent).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "emd_sift.h"

namespace emdk {

constexpr double kEnergyRatioThr = 0.2;
constexpr double kStdThr = 0.2;
constexpr double kSvarThr = 0.001;
constexpr double kTotalPowerThr = 0.005;
constexpr double kRangeThr = 0.001;
constexpr double kPowerFloor = 1e-10;
constexpr double kResidueAtol = 1e-8;

enum SplineKind { kConst = 0, kLinear = 1, kThreePoint = 2, kNotAKnot = 3 };

struct Tile {
  double* Y;   // knot values
  double* A;   // sub-diagonal (rows); second derivatives M (by slot) after the solve
  double* Bd;  // diagonal (rows)
  double* Cc;  // super-diagonal (rows)
  double* D;   // right-hand side -> solution (rows)
  int* X;      // knot positions (sample index; mirrored knots may be <0 or >N-1)
};

__device__ __forceinline__ Tile make_tile(double* base, int slots) {
  Tile t;
  t.Y = base;
  t.A = base + slots;
  t.Bd = base + 2 * slots;
  t.Cc = base + 3 * slots;
  t.D = base + 4 * slots;
  t.X = reinterpret_cast<int*>(base + 5 * slots);
  return t;
}

template <int WG>
struct SiftShared {
  static constexpr int NW = WG / 64;
  double edge_first[WG];
  double edge_last[WG];
  double red[NW][6];
  double red2[NW][4];
  unsigned scan_w[NW];
  int nz_w[NW];
  int plat_w[NW];
  int bad_w[NW];
  int app_w[NW];
  double s_first, s_last;
  // serial-path results
  int ser_tm, ser_tn, ser_nm, ser_nn, ser_bad;
  // mirror results
  int k0[2], nk[2], kind[2], row0[2], mtot;
  int bad;
  double p3[2][4];
  int queue_idx;
};

__device__ __forceinline__ int ceil_div(int a, int b) { return (a + b - 1) / b; }
__device__ __forceinline__ int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}
__device__ __forceinline__ int round_half_even_mid(int s) {
  // np.round(s / 2.0) for integer s >= 0
  int f = s >> 1;
  if ((s & 1) == 0) return f;
  return (f & 1) ? f + 1 : f;
}

// ---- wave helpers ---------------------------------------------------------------------
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    unsigned t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}
__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}
// Perfect binary tree over the 64 lanes (lane order); every lane gets the same bits.
__device__ __forceinline__ double wave_tree_sum(double v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v = fmin(v, __shfl_xor(v, d, 64));
  return v;
}
template <int NW>
__device__ __forceinline__ double waves_tree(const double* r, int stride) {
  if constexpr (NW == 1) {
    return r[0];
  } else if constexpr (NW == 2) {
    return r[0] + r[stride];
  } else {
    return (r[0] + r[stride]) + (r[2 * stride] + r[3 * stride]);
  }
}

// Streaming pairwise (perfect binary tree) accumulator over at most 2^(L-1) leaves pushed
// in index order.  push<J>() merges with the stored left siblings given by J's trailing ones.
template <int L>
struct PairAcc {
  double s[L];
  template <int J>
  __device__ __forceinline__ void push(double x) {
    int lvl = 0;
#pragma unroll
    for (int jj = J; jj & 1; jj >>= 1) {
      x = s[lvl] + x;
      ++lvl;
    }
    s[lvl] = x;
  }
  // total of the first c leaves (c a power of two, 1 .. 2^(L-1))
  __device__ __forceinline__ double total(int c) const {
    double r = s[0];
#pragma unroll
    for (int l = 1; l < L; ++l)
      if (c >= (1 << l)) r = s[l];
    return r;
  }
};
template <int C>
constexpr int levels_for() {
  int l = 1;
  while ((1 << (l - 1)) < C) ++l;
  return l;
}

template <int J, int C, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (J < C) {
    f(std::integral_constant<int, J>{});
    static_for<J + 1, C>(f);
  }
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
  int defer;       // the knots (or the plateau staging copy) do not fit this launch's tile
  int ovf;
  int exc_max, exc_min;  // this thread's exclusive offsets (-1 = unknown: serial path)
};

// Serial scan on one lane over a staged copy of the IMF: emit(idx, is_max) in index order.
template <class Emit>
__device__ __forceinline__ void serial_extrema(const double* S, int N, Emit&& emit) {
  if (N < 3) return;
  const double d_last = S[N - 1] - S[N - 2];  // d[-1]
  double dprev = 0.0;
  int run_a = -1;
  double run_dbefore = 0.0;
  bool first_run_seen = false;
  for (int k = 0; k <= N - 2; ++k) {
    double dk = S[k + 1] - S[k];
    if (k >= 1 && dprev * dk < 0.0) emit(k, dprev > 0.0);
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
          if (dk < 0.0) emit(mid, true);
        } else {
          if (dk > 0.0) emit(mid, false);
        }
      }
    }
    dprev = dk;
  }
  // a run still open here reaches the end of the signal and is dropped
}

template <int WG, int C>
__device__ __forceinline__ void write_knots_parallel(const Tile& t, const double (&v)[C], int cnt,
                                                     int i0, uint64_t mmax, uint64_t mmin,
                                                     int om, int on, int nm, int nn, int tb0,
                                                     int tb1, int& bad) {
  static_for<0, C>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < cnt) {
      if ((mmax >> j) & 1) {
        if (om < nm) {
          t.X[tb0 + om] = i0 + j;
          t.Y[tb0 + om] = v[j];
          bad |= (v[j] < 0.0);
        }
        ++om;
      }
      if ((mmin >> j) & 1) {
        if (on < nn) {
          t.X[tb1 + on] = i0 + j;
          t.Y[tb1 + on] = v[j];
          bad |= (v[j] > 0.0);
        }
        ++on;
      }
    }
  });
}

__device__ __forceinline__ void serial_write(const Tile& t, const double* S, int N, int nm, int nn, int tb0,
                             int tb1, int& bad) {
  int om = 0, on = 0;
  serial_extrema(S, N, [&](int idx, bool is_max) {
    double val = S[idx];
    if (is_max) {
      if (om < nm) {
        t.X[tb0 + om] = idx;
        t.Y[tb0 + om] = val;
        bad |= (val < 0.0);
      }
      ++om;
    } else {
      if (on < nn) {
        t.X[tb1 + on] = idx;
        t.Y[tb1 + on] = val;
        bad |= (val > 0.0);
      }
      ++on;
    }
  });
}

template <int WG, int C>
__device__ __forceinline__ DetectResult detect(const double (&v)[C], int N, int c, int i0,
                                               int cnt, SiftShared<WG>& sh, const Tile& t,
                                               const SiftParams& P) {
  constexpr int NW = WG / 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int pad = P.nbsym;

  // ---- edge exchange ------------------------------------------------------------------
  double last = v[0];
  static_for<1, C>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < cnt) last = v[j];
  });
  if (cnt > 0) {
    sh.edge_first[tid] = v[0];
    sh.edge_last[tid] = last;
    if (i0 == 0) sh.s_first = v[0];
    if (i0 + cnt == N) sh.s_last = last;
  }
  __syncthreads();
  const double prevE = (tid > 0 && cnt > 0) ? sh.edge_last[tid - 1] : 0.0;
  const double nextE = (i0 + c < N) ? sh.edge_first[tid + 1] : 0.0;

  // ---- local scan of the chunk --------------------------------------------------------
  uint64_t mmax = 0, mmin = 0;
  int nz = 0;
  int plateau = 0;
  static_for<0, C>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < cnt) {
      const int i = i0 + j;
      const double x = v[j];
      double prev;
      if constexpr (j == 0) prev = prevE; else prev = v[j - 1];
      double next;
      if constexpr (j + 1 < C) next = (j + 1 < cnt) ? v[j + 1] : nextE; else next = nextE;
      if (i <= N - 2) {
        const double d2 = next - x;
        plateau |= (d2 == 0.0);
        nz += (x * next < 0.0);
        if (i >= 1) {
          const double d1 = x - prev;
          const double p = d1 * d2;
          if (p < 0.0) {
            if (d1 > 0.0) mmax |= (uint64_t(1) << j);
            if (d1 < 0.0) mmin |= (uint64_t(1) << j);
          }
        }
      }
      nz += (x == 0.0 && (i == 0 || prev != 0.0));
    }
  });
  const unsigned cm = __popcll(mmax);
  const unsigned cn = __popcll(mmin);
  const unsigned packed = cm | (cn << 16);
  const unsigned incl = wave_incl_scan(packed, lane);
  const int nzw = wave_sum_int(nz);
  const int plw = (__ballot(plateau) != 0ull) ? 1 : 0;
  if (lane == 63) sh.scan_w[w] = incl;
  if (lane == 0) {
    sh.nz_w[w] = nzw;
    sh.plat_w[w] = plw;
  }
  __syncthreads();
  unsigned pre = 0, tot = 0;
  int nzt = 0, plt = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const unsigned sw = sh.scan_w[q];
    if (q < w) pre += sw;
    tot += sw;
    nzt += sh.nz_w[q];
    plt |= sh.plat_w[q];
  }

  DetectResult R;
  R.nzer = nzt;

  if (plt) {
    // ---- plateau present: one lane applies the full run rules to a copy of the IMF
    // staged in the tile's tridiagonal arrays (4 fp64 per slot, free during detection)
    R.exc_max = -1;
    R.exc_min = -1;
    if (N > 4 * P.tile_slots) {
      R.defer = 1;
      R.tm = R.tn = R.nm = R.nn = 0;
      R.ovf = 0;
      return R;
    }
    double* staging = t.A;
    static_for<0, C>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j < cnt) staging[i0 + j] = v[j];
    });
    __syncthreads();
    if (tid == 0) {
      int tm = 0, tn = 0;
      serial_extrema(staging, N, [&](int, bool is_max) {
        if (is_max) ++tm; else ++tn;
      });
      const int nm = min(tm, P.max_extrema), nn = min(tn, P.max_extrema);
      int bad = 0;
      if (4 * pad + nm + nn <= P.tile_slots)
        serial_write(t, staging, N, nm, nn, pad, 3 * pad + nm, bad);
      sh.ser_tm = tm;
      sh.ser_tn = tn;
      sh.ser_nm = nm;
      sh.ser_nn = nn;
      sh.ser_bad = bad;
      sh.bad = bad;
    }
    __syncthreads();
    R.tm = sh.ser_tm;
    R.tn = sh.ser_tn;
    R.nm = sh.ser_nm;
    R.nn = sh.ser_nn;
    R.defer = (4 * pad + R.nm + R.nn > P.tile_slots);
    R.ovf = (R.tm >= P.max_extrema || R.tn >= P.max_extrema);
    return R;
  }

  const unsigned excl = pre + incl - packed;
  R.exc_max = int(excl & 0xffffu);
  R.exc_min = int(excl >> 16);
  R.tm = int(tot & 0xffffu);
  R.tn = int(tot >> 16);
  R.nm = min(R.tm, P.max_extrema);
  R.nn = min(R.tn, P.max_extrema);
  R.ovf = (R.tm >= P.max_extrema || R.tn >= P.max_extrema);
  R.defer = (4 * pad + R.nm + R.nn > P.tile_slots);
  if (R.defer) return R;
  int bad = 0;
  write_knots_parallel<WG, C>(t, v, cnt, i0, mmax, mmin, R.exc_max, R.exc_min, R.nm, R.nn, pad,
                              3 * pad + R.nm, bad);
  const int bw = (__ballot(bad) != 0ull) ? 1 : 0;
  if (lane == 0) sh.bad_w[w] = bw;
  __syncthreads();
  if (tid == 0) {
    int b = 0;
    for (int q = 0; q < NW; ++q) b |= sh.bad_w[q];
    sh.bad = b;
  }
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
  int sp;        // 0: no edge sample, 1: edge sample first, 2: edge sample last
  int si;        // edge sample index
  double sv;     // edge sample value
};

__device__ __forceinline__ KDesc kd(int from, int to, int rev) {
  KDesc d;
  d.from = from;
  d.to = to;
  d.rev = rev;
  d.sp = 0;
  d.si = 0;
  d.sv = 0.0;
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
__device__ __forceinline__ int kd_idx(const KDesc& d, const Tile& t, int tb, int q) {
  const int s = kd_slot(d, q);
  return s < 0 ? d.si : t.X[tb + s];
}
__device__ __forceinline__ double kd_val(const KDesc& d, const Tile& t, int tb, int q) {
  const int s = kd_slot(d, q);
  return s < 0 ? d.sv : t.Y[tb + s];
}

__device__ __forceinline__ void finish_side(const Tile& t, int side, int tb, int n,
                                            const KDesc& L, int lsym, const KDesc& R,
                                            int rsym, int& k0, int& nk, int& kind, int& bad,
                                            double* p3) {
  // Left knots sit at positions 2*lsym - idx.  A left knot equal in position to the first
  // true knot is the same sample (same value) and is dropped, which is what PyEMD's
  // delete of equal adjacent positions does; likewise on the right.
  const int ln = kd_len(L), rn = kd_len(R);
  int nl = ln;
  if (nl > 0 && 2 * lsym - kd_idx(L, t, tb, nl - 1) == t.X[tb]) --nl;
  for (int q = 0; q < ln; ++q) {
    const int idx = kd_idx(L, t, tb, q);
    const double val = kd_val(L, t, tb, q);
    bad |= side == 0 ? (val < 0.0) : (val > 0.0);
    if (q < nl) {
      t.X[tb - nl + q] = 2 * lsym - idx;
      t.Y[tb - nl + q] = val;
    }
  }
  const int rstart = (rn > 0 && 2 * rsym - kd_idx(R, t, tb, 0) == t.X[tb + n - 1]) ? 1 : 0;
  // right knots go after the true ones (slots that kd_idx never reads)
  for (int q = 0; q < rn; ++q) {
    const int idx = kd_idx(R, t, tb, q);
    const double val = kd_val(R, t, tb, q);
    bad |= side == 0 ? (val < 0.0) : (val > 0.0);
    if (q >= rstart) {
      t.X[tb + n + q - rstart] = 2 * rsym - idx;
      t.Y[tb + n + q - rstart] = val;
    }
  }
  k0 = tb - nl;
  nk = nl + n + (rn - rstart);
  kind = nk >= 4 ? kNotAKnot : nk == 3 ? kThreePoint : nk == 2 ? kLinear : kConst;
  if (kind == kThreePoint) {
    // PyEMD cubic_spline_3pts: natural cubic spline in slope form, solved in closed form
    const double x0 = t.X[k0], x1 = t.X[k0 + 1], x2 = t.X[k0 + 2];
    const double y0 = t.Y[k0], y1 = t.Y[k0 + 1], y2 = t.Y[k0 + 2];
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
}

// Boundary mirroring (one lane).  PyEMD _prepare_points_simple with T = arange(N); the
// reference's mirror_one_side (kernels.py:360-558) follows the same branch table.
__device__ __forceinline__ void mirror(const Tile& t, int N, int nm, int nn, int nb, double S0,
                                       double SN, int* k0, int* nk, int* kind, int& bad,
                                       double (*p3)[4]) {
  const int tb0 = nb, tb1 = 3 * nb + nm;
  const int Xm0 = t.X[tb0], XmL = t.X[tb0 + nm - 1];
  const int Xn0 = t.X[tb1], XnL = t.X[tb1 + nn - 1];
  const double Ym0 = t.Y[tb0], YmL = t.Y[tb0 + nm - 1];
  const double Yn0 = t.Y[tb1], YnL = t.Y[tb1 + nn - 1];
  KDesc lmax, lmin, rmax, rmin;
  int lsym, rsym;
  if (Xm0 < Xn0) {
    if (S0 > Yn0) {
      lmax = kd(1, min(nm, nb + 1), 1);
      lmin = kd(0, min(nn, nb), 1);
      lsym = Xm0;
    } else {
      lmax = kd(0, min(nm, nb), 1);
      lmin = kd(0, min(nn, nb - 1), 1);
      lmin.sp = 2;
      lmin.si = 0;
      lmin.sv = S0;
      lsym = 0;
    }
  } else {
    if (S0 < Ym0) {
      lmax = kd(0, min(nm, nb), 1);
      lmin = kd(1, min(nn, nb + 1), 1);
      lsym = Xn0;
    } else {
      lmax = kd(0, min(nm, nb - 1), 1);
      lmax.sp = 2;
      lmax.si = 0;
      lmax.sv = S0;
      lmin = kd(0, min(nn, nb), 1);
      lsym = 0;
    }
  }
  if (XmL < XnL) {
    if (SN < YmL) {
      rmax = kd(max(nm - nb, 0), nm, 1);
      rmin = kd(max(nn - nb - 1, 0), nn - 1, 1);
      rsym = XnL;
    } else {
      rmax = kd(max(nm - nb + 1, 0), nm, 1);
      rmax.sp = 1;
      rmax.si = N - 1;
      rmax.sv = SN;
      rmin = kd(max(nn - nb, 0), nn, 1);
      rsym = N - 1;
    }
  } else {
    if (SN > YnL) {
      rmax = kd(max(nm - nb - 1, 0), nm - 1, 1);
      rmin = kd(max(nn - nb, 0), nn, 1);
      rsym = XmL;
    } else {
      rmax = kd(max(nm - nb, 0), nm, 1);
      rmin = kd(max(nn - nb + 1, 0), nn, 1);
      rmin.sp = 1;
      rmin.si = N - 1;
      rmin.sv = SN;
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
  if (2 * lsym - kd_idx(lmin, t, tb1, 0) > 0 || 2 * lsym - kd_idx(lmax, t, tb0, 0) > 0) {
    if (lsym == Xm0) lmax = kd(0, min(nm, nb), 1);
    else lmin = kd(0, min(nn, nb), 1);
    lsym = 0;
  }
  if (2 * rsym - kd_idx(rmin, t, tb1, kd_len(rmin) - 1) < N - 1 ||
      2 * rsym - kd_idx(rmax, t, tb0, kd_len(rmax) - 1) < N - 1) {
    if (rsym == XmL) rmax = kd(max(nm - nb, 0), nm, 1);
    else rmin = kd(max(nn - nb, 0), nn, 1);
    rsym = N - 1;
  }
  finish_side(t, 0, tb0, nm, lmax, lsym, rmax, rsym, k0[0], nk[0], kind[0], bad, p3[0]);
  finish_side(t, 1, tb1, nn, lmin, lsym, rmin, rsym, k0[1], nk[1], kind[1], bad, p3[1]);
}

// =======================================================================================
// Joint not-a-knot spline solve for both envelopes: setup of the reduced (n-2) systems
// (M_0 and M_{n-1} eliminated), cyclic reduction over the concatenated rows, recovery of
// the end second derivatives.  M is left in t.A indexed by knot slot.
// =======================================================================================
template <int WG>
__device__ __forceinline__ void spline_solve(const Tile& t, const SiftShared<WG>& sh) {
  const int tid = threadIdx.x;
  const int mtot = sh.mtot;
  const int row1 = sh.row0[1];
  for (int r = tid; r < mtot; r += WG) {
    const int side = (r < row1) ? 0 : 1;
    const int k = r - sh.row0[side];
    const int m = sh.nk[side] - 2;
    const int i = sh.k0[side] + k + 1;
    const double xa = t.X[i - 1], xb = t.X[i], xc = t.X[i + 1];
    const double ya = t.Y[i - 1], yb = t.Y[i], yc = t.Y[i + 1];
    const double hl = xb - xa, hr = xc - xb;
    const double rhs = 6.0 * ((yc - yb) / hr - (yb - ya) / hl);
    double a, b, cc, d;
    if (k == 0) {
      a = 0.0;
      b = hl + 2.0 * hr;
      cc = hr - hl;
      d = hr * rhs / (hl + hr);
    } else if (k == m - 1) {
      a = hl - hr;
      b = 2.0 * hl + hr;
      cc = 0.0;
      d = hl * rhs / (hl + hr);
    } else {
      a = hl;
      b = 2.0 * (hl + hr);
      cc = hr;
      d = rhs;
    }
    t.A[r] = a;
    t.Bd[r] = b;
    t.Cc[r] = cc;
    t.D[r] = d;
  }
  __syncthreads();
  if (mtot > 0) {
    int s = 1;
    for (; 2 * s <= mtot; s <<= 1) {
      for (int q = tid;; q += WG) {
        const int i = 2 * s * q + 2 * s - 1;
        if (i >= mtot) break;
        const int l = i - s, rr = i + s;
        const double alpha = -t.A[i] / t.Bd[l];
        double na = alpha * t.A[l];
        double nb = t.Bd[i] + alpha * t.Cc[l];
        double nd = t.D[i] + alpha * t.D[l];
        double nc = 0.0;
        if (rr < mtot) {
          const double gamma = -t.Cc[i] / t.Bd[rr];
          nc = gamma * t.Cc[rr];
          nb = nb + gamma * t.A[rr];
          nd = nd + gamma * t.D[rr];
        }
        t.A[i] = na;
        t.Bd[i] = nb;
        t.Cc[i] = nc;
        t.D[i] = nd;
      }
      __syncthreads();
    }
    if (tid == 0) t.D[s - 1] = t.D[s - 1] / t.Bd[s - 1];
    __syncthreads();
    for (s >>= 1; s >= 1; s >>= 1) {
      for (int q = tid;; q += WG) {
        const int i = 2 * s * q + s - 1;
        if (i >= mtot) break;
        const double xl = (i - s >= 0) ? t.D[i - s] : 0.0;
        const double xr = (i + s < mtot) ? t.D[i + s] : 0.0;
        t.D[i] = (t.D[i] - t.A[i] * xl - t.Cc[i] * xr) / t.Bd[i];
      }
      __syncthreads();
    }
    // scatter M by knot slot into A (A's row contents are dead now)
    for (int r = tid; r < mtot; r += WG) {
      const int side = (r < row1) ? 0 : 1;
      const int k = r - sh.row0[side];
      t.A[sh.k0[side] + k + 1] = t.D[r];
    }
    const int side = (tid == 0) ? 0 : (tid == 1 ? 1 : -1);
    if (side >= 0 && sh.kind[side] == kNotAKnot) {
      const int n = sh.nk[side], k0 = sh.k0[side], r0 = sh.row0[side], m = n - 2;
      const double x0 = t.X[k0], x1 = t.X[k0 + 1], x2 = t.X[k0 + 2];
      const double h0 = x1 - x0, h1 = x2 - x1;
      const double M1 = t.D[r0], M2 = t.D[r0 + 1];
      t.A[k0] = ((h0 + h1) * M1 - h0 * M2) / h1;
      const double xa = t.X[k0 + n - 3], xb = t.X[k0 + n - 2], xc = t.X[k0 + n - 1];
      const double hnm3 = xb - xa, hnm2 = xc - xb;
      const double Mb = t.D[r0 + m - 1], Ma = t.D[r0 + m - 2];
      t.A[k0 + n - 1] = ((hnm3 + hnm2) * Mb - hnm2 * Ma) / hnm3;
    }
  }
  __syncthreads();
}

// Envelope value at a sample for the closed-form spline kinds (3 knots: PyEMD's natural
// cubic; 2 knots: straight line; 1 knot: constant).
__device__ __noinline__ double eval_small(const Tile& t, int k0, int kind, const double* p3,
                                          int pos) {
  const double tt = double(pos);
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

// Cursor over one envelope's knots for increasing sample positions (not-a-knot kind keeps
// the current interval in registers; the small kinds read their 2-3 knots per sample).
struct EnvCursor {
  int k0, nk, kind, j;
  int xi0, xi1;
  double y0, y1, M0, M1, invh, hh6;
  const double* p3;
  __device__ __forceinline__ void load(const Tile& t) {
    xi0 = t.X[k0 + j];
    xi1 = t.X[k0 + j + 1];
    y0 = t.Y[k0 + j];
    y1 = t.Y[k0 + j + 1];
    M0 = t.A[k0 + j];
    M1 = t.A[k0 + j + 1];
    const double h = double(xi1 - xi0);
    invh = 1.0 / h;
    hh6 = h * h / 6.0;
  }
  __device__ __forceinline__ void init(const Tile& t, int k0_, int nk_, int kind_, int jstart,
                                       const double* p3_) {
    k0 = k0_;
    nk = nk_;
    kind = kind_;
    p3 = p3_;
    if (kind == kNotAKnot) {
      j = max(0, min(jstart, nk - 2));
      load(t);
    }
  }
  __device__ __forceinline__ double eval(const Tile& t, int pos) {
    if (kind != kNotAKnot) return eval_small(t, k0, kind, p3, pos);
    bool moved = false;
    while (j < nk - 2 && t.X[k0 + j + 1] <= pos) {
      ++j;
      moved = true;
    }
    while (j > 0 && t.X[k0 + j] > pos) {
      --j;
      moved = true;
    }
    if (moved) load(t);
    const double A = double(xi1 - pos) * invh;
    const double B = double(pos - xi0) * invh;
    return A * y0 + B * y1 + ((A * A * A - A) * M0 + (B * B * B - B) * M1) * hh6;
  }
};

template <int WG>
__device__ __forceinline__ int start_interval(const Tile& t, const SiftShared<WG>& sh, int side,
                                              int exc, int i0, int tb) {
  if (sh.kind[side] != kNotAKnot) return 0;
  const int k0 = sh.k0[side], nk = sh.nk[side];
  if (exc >= 0) return (tb - k0) + exc - 1;
  // serial-path detection: binary search for the last knot at or before i0
  int lo = 0, hi = nk - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.X[k0 + mid] <= i0) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct SiftStats {
  double sq_new, sq_diff, sq_old, std_sum, old_max, old_min;
};

// One sift: evaluate both envelopes on the chunk, subtract the mean, accumulate the
// PyEMD stop statistics.  v is updated in place.
template <int WG, int C>
__device__ __forceinline__ SiftStats envelope_update(const Tile& t, SiftShared<WG>& sh, double (&v)[C],
                                     int c, int i0, int cnt, const DetectResult& D, int nb) {
  constexpr int NW = WG / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  EnvCursor cmax, cmin;
  const int tb0 = nb, tb1 = 3 * nb + D.nm;
  cmax.init(t, sh.k0[0], sh.nk[0], sh.kind[0], start_interval(t, sh, 0, D.exc_max, i0, tb0),
            sh.p3[0]);
  cmin.init(t, sh.k0[1], sh.nk[1], sh.kind[1], start_interval(t, sh, 1, D.exc_min, i0, tb1),
            sh.p3[1]);
  PairAcc<levels_for<C>()> a_new, a_diff, a_old, a_std;
  double omax = -INFINITY, omin = INFINITY;
  static_for<0, C>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < c) {
      double ln = 0.0, ld = 0.0, lo = 0.0, ls = 0.0;
      if (j < cnt) {
        const int pos = i0 + j;
        const double emax = cmax.eval(t, pos);
        const double emin = cmin.eval(t, pos);
        const double mean = 0.5 * (emax + emin);
        const double old = v[j];
        const double nw = old - mean;
        const double diff = nw - old;
        const double r = diff / nw;
        ln = nw * nw;
        ld = diff * diff;
        lo = old * old;
        ls = r * r;
        omax = fmax(omax, old);
        omin = fmin(omin, old);
        v[j] = nw;
      }
      a_new.template push<j>(ln);
      a_diff.template push<j>(ld);
      a_old.template push<j>(lo);
      a_std.template push<j>(ls);
    }
  });
  double s0 = wave_tree_sum(a_new.total(c));
  double s1 = wave_tree_sum(a_diff.total(c));
  double s2 = wave_tree_sum(a_old.total(c));
  double s3 = wave_tree_sum(a_std.total(c));
  double mx = wave_max(omax);
  double mn = wave_min(omin);
  if (lane == 0) {
    sh.red[w][0] = s0;
    sh.red[w][1] = s1;
    sh.red[w][2] = s2;
    sh.red[w][3] = s3;
    sh.red[w][4] = mx;
    sh.red[w][5] = mn;
  }
  __syncthreads();
  SiftStats st;
  st.sq_new = waves_tree<NW>(&sh.red[0][0], 6);
  st.sq_diff = waves_tree<NW>(&sh.red[0][1], 6);
  st.sq_old = waves_tree<NW>(&sh.red[0][2], 6);
  st.std_sum = waves_tree<NW>(&sh.red[0][3], 6);
  double gmx = sh.red[0][4], gmn = sh.red[0][5];
#pragma unroll
  for (int q = 1; q < NW; ++q) {
    gmx = fmax(gmx, sh.red[q][4]);
    gmn = fmin(gmn, sh.red[q][5]);
  }
  st.old_max = gmx;
  st.old_min = gmn;
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

// Mirror + spline solve for the knots of detection D (shared by the sift and the
// envelope-only kernel).
template <int WG>
__device__ __forceinline__ void build_envelopes(const Tile& t, SiftShared<WG>& sh,
                                                const DetectResult& D, int N,
                                                const SiftParams& P) {
  if (threadIdx.x == 0) {
    int bad = sh.bad;
    mirror(t, N, D.nm, D.nn, P.nbsym, sh.s_first, sh.s_last, sh.k0, sh.nk, sh.kind, bad,
           sh.p3);
    sh.bad = bad;
    const int m0 = sh.kind[0] == kNotAKnot ? sh.nk[0] - 2 : 0;
    const int m1 = sh.kind[1] == kNotAKnot ? sh.nk[1] - 2 : 0;
    sh.row0[0] = 0;
    sh.row0[1] = m0;
    sh.mtot = m0 + m1;
  }
  __syncthreads();
  spline_solve<WG>(t, sh);
}

// =======================================================================================
// Whole decomposition of one trace (PyEMD EMD.emd default loop).
// =======================================================================================
struct SignalIO {
  const double* S;
  int N;
  double* out;      // row k at out + k*N
  int* n_imfs;
  int* overflow;
  int* dbg_n;
  int* dbg_e;
  int* dbg_z;
  int* stats;       // 4 ints or null
};

// Returns false when the trace was deferred (its knots outgrow this launch's tile); its
// outputs are then left for the retry launch to write.
template <int WG, int C>
__device__ __forceinline__ bool decompose(const SignalIO& io, SiftShared<WG>& sh,
                                          const Tile& t, const SiftParams& P, int flags0) {
  constexpr int NW = WG / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int N = io.N;
  const int c = next_pow2(max(1, ceil_div(N, WG)));
  const int i0 = tid * c;
  const int cnt = max(0, min(N - i0, c));
  const int nb = P.nbsym;

  if (tid == 0) {
    for (int k = 0; k < P.max_imfs; ++k) {
      io.dbg_n[k] = -1;
      io.dbg_e[k] = -1;
      io.dbg_z[k] = -1;
    }
  }
  // The running sum of accepted IMFs (PyEMD's np.sum(IMF, axis=0), rows added in order)
  // is kept in the last output row, which no IMF occupies before the residue is written.
  double* acc = io.out + (int64_t)(P.max_imfs - 1) * N;
  bool have_acc = false;
  double v[C];

  int imf_count = 0;
  bool finished = false;
  int extNo_last = -1;
  int overflow = 0;
  int sifts = 0, peak = 0, knot_sum = 0, flags = flags0;

  while (!finished && imf_count < P.max_imfs - 1) {
#pragma unroll
    for (int j = 0; j < C; ++j)
      v[j] = (j < cnt) ? io.S[i0 + j] - (have_acc ? acc[i0 + j] : 0.0) : 0.0;
    DetectResult D = detect<WG, C>(v, N, c, i0, cnt, sh, t, P);
    if (D.defer) return false;
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
        build_envelopes<WG>(t, sh, D, N, P);
        const SiftStats st = envelope_update<WG, C>(t, sh, v, c, i0, cnt, D, nb);
        const int bad = sh.bad;
        ++sifts;
        knot_sum += sh.nk[0] + sh.nk[1];
        __syncthreads();  // everyone has read sh.red / sh.bad / sh.nk before detect reuses them
        D = detect<WG, C>(v, N, c, i0, cnt, sh, t, P);
        if (D.defer) return false;
        overflow |= D.ovf;
        peak = max(peak, max(D.tm, D.tn));
        const int extNo2 = D.nm + D.nn;
        const int nzm2 = D.nzer;
        const bool f2 = abs(extNo2 - nzm2) < 2;
        const bool f1 = check_imf(st, bad);
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

    // accept the IMF row, then PyEMD end_condition on S - (acc + imf)
    double* row = io.out + (int64_t)imf_count * N;
    PairAcc<levels_for<C>()> a_abs;
    double tmax = -INFINITY, tmin = INFINITY;
    static_for<0, C>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if (j < c) {
        double la = 0.0;
        if (j < cnt) {
          row[i0 + j] = v[j];
          const double tmp = io.S[i0 + j] - ((have_acc ? acc[i0 + j] : 0.0) + v[j]);
          la = fabs(tmp);
          tmax = fmax(tmax, tmp);
          tmin = fmin(tmin, tmp);
        }
        a_abs.template push<j>(la);
      }
    });
    {
      const double sa = wave_tree_sum(a_abs.total(c));
      const double mx = wave_max(tmax), mn = wave_min(tmin);
      if (lane == 0) {
        sh.red2[w][0] = sa;
        sh.red2[w][1] = mx;
        sh.red2[w][2] = mn;
      }
      __syncthreads();
      const double gsa = waves_tree<NW>(&sh.red2[0][0], 4);
      double gmx = sh.red2[0][1], gmn = sh.red2[0][2];
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        gmx = fmax(gmx, sh.red2[q][1]);
        gmn = fmin(gmn, sh.red2[q][2]);
      }
      __syncthreads();
      if ((gmx - gmn) < kRangeThr) finished = true;
      else if (gsa < kTotalPowerThr) finished = true;
    }
    ++imf_count;
    const bool capped = imf_count >= P.max_imfs - 1;
    if (!finished && capped) flags |= 1;
    const bool leaving = finished || capped;
    if (leaving && extNo_last <= 2) {
      --imf_count;  // the last row was a trend (PyEMD drops it); acc excludes it
    } else {
#pragma unroll
      for (int j = 0; j < C; ++j)
        if (j < cnt) acc[i0 + j] = (have_acc ? acc[i0 + j] : 0.0) + v[j];
      have_acc = true;
    }
    if (leaving) break;  // PyEMD drops the trend row after leaving its loop
  }

  // residue = S - sum(IMFs); appended unless np.allclose(residue, 0) (max|r| <= 1e-8)
  int app = 0;
#pragma unroll
  for (int j = 0; j < C; ++j) {
    if (j < cnt) {
      const double r = io.S[i0 + j] - (have_acc ? acc[i0 + j] : 0.0);
      v[j] = r;
      app |= !(fabs(r) <= kResidueAtol);
    }
  }
  const int aw = (__ballot(app) != 0ull) ? 1 : 0;
  if (lane == 0) sh.app_w[w] = aw;
  __syncthreads();
  int append = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) append |= sh.app_w[q];
  int n_out = imf_count;
  if (append) {
    double* row = io.out + (int64_t)imf_count * N;
#pragma unroll
    for (int j = 0; j < C; ++j)
      if (j < cnt) row[i0 + j] = v[j];
    n_out = imf_count + 1;
  }
  if (P.zero_fill) {
    for (int k = n_out; k < P.max_imfs; ++k) {
      double* row = io.out + (int64_t)k * N;
#pragma unroll
      for (int j = 0; j < C; ++j)
        if (j < cnt) row[i0 + j] = 0.0;
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
  return true;
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
  io.N = (int)n;
  io.out = b.imfs + (b.out_offsets ? b.out_offsets[s] : (int64_t)s * max_imfs * n);
  io.n_imfs = b.n_imfs + s;
  io.overflow = b.overflow + s;
  io.dbg_n = b.dbg_n + (int64_t)s * max_imfs;
  io.dbg_e = b.dbg_e + (int64_t)s * max_imfs;
  io.dbg_z = b.dbg_z + (int64_t)s * max_imfs;
  io.stats = b.stats ? b.stats + (int64_t)s * 4 : nullptr;
  return io;
}

// Knot tile of this workgroup: the dynamic LDS array, or (GT) a per-workgroup global region.
template <bool GT>
__device__ __forceinline__ Tile workgroup_tile(double* lds, char* spill, int64_t stride,
                                               int slots) {
  if constexpr (GT) {
    return make_tile(reinterpret_cast<double*>(spill + (int64_t)blockIdx.x * stride), slots);
  } else {
    return make_tile(lds, slots);
  }
}

__device__ __forceinline__ void defer_trace(const Batch& b, int s) {
  const unsigned k = atomicAdd(b.defer_count, 1u);
  b.defer_list[k] = s;
}

// One workgroup per trace (grid = traces of this launch).  The reference launch model.
template <int WG, int C, bool GT>
__global__ __launch_bounds__(WG) void emd_plain_kernel(Batch b, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  const Tile t = workgroup_tile<GT>(dyn_lds, b.spill, b.spill_stride, P.tile_slots);
  const int q = b.base + (int)blockIdx.x;
  const int s = b.order ? b.order[q] : q;
  if (!decompose<WG, C>(signal_io(b, s, P.max_imfs), sh, t, P, b.retry ? 2 : 0) &&
      threadIdx.x == 0)
    defer_trace(b, s);
}

// Persistent workgroups pull queue positions from an atomic counter until the queue
// (b.B entries, or *b.count_ptr when set) is drained.
template <int WG, int C, bool GT>
__global__ __launch_bounds__(WG) void emd_persistent_kernel(Batch b, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  const Tile t = workgroup_tile<GT>(dyn_lds, b.spill, b.spill_stride, P.tile_slots);
  const int count = b.count_ptr ? (int)*b.count_ptr : b.B;
  while (true) {
    if (threadIdx.x == 0) sh.queue_idx = (int)atomicAdd(b.queue, 1u);
    __syncthreads();
    const int idx = sh.queue_idx;
    __syncthreads();
    if (idx >= count) break;
    const int q = b.base + idx;
    const int s = b.order ? b.order[q] : q;
    if (!decompose<WG, C>(signal_io(b, s, P.max_imfs), sh, t, P, b.retry ? 2 : 0) &&
        threadIdx.x == 0)
      defer_trace(b, s);
  }
}

// Envelope pass without sifting (diagnostic; reference envelope_batch_kernel).
template <int WG, int C, bool GT>
__global__ __launch_bounds__(WG) void emd_envelope_kernel(EnvBatch e, SiftParams P) {
  extern __shared__ double dyn_lds[];
  __shared__ SiftShared<WG> sh;
  const int tid = threadIdx.x;
  const Tile t = workgroup_tile<GT>(dyn_lds, e.spill, e.spill_stride, P.tile_slots);
  const int bidx = blockIdx.x;
  const int N = e.N;
  const double* S = e.values + (int64_t)bidx * N;
  const int c = next_pow2(max(1, ceil_div(N, WG)));
  const int i0 = tid * c;
  const int cnt = max(0, min(N - i0, c));
  double v[C];
#pragma unroll
  for (int j = 0; j < C; ++j) v[j] = (j < cnt) ? S[i0 + j] : 0.0;
  DetectResult D = detect<WG, C>(v, N, c, i0, cnt, sh, t, P);
  if (D.defer) {
    if (tid == 0) e.degenerate[bidx] = -1;  // tile too small: caller retries with a larger one
    return;
  }
  if (tid == 0) {
    e.n_extrema[2 * bidx] = D.nm;
    e.n_extrema[2 * bidx + 1] = D.nn;
  }
  const int extNo = D.nm + D.nn;
  if (extNo <= 2 || D.nm < 1 || D.nn < 1) {
    if (tid == 0) e.degenerate[bidx] = 1;
    return;
  }
  build_envelopes<WG>(t, sh, D, N, P);
  const int nb = P.nbsym;
  EnvCursor cmax, cmin;
  cmax.init(t, sh.k0[0], sh.nk[0], sh.kind[0], start_interval(t, sh, 0, D.exc_max, i0, nb),
            sh.p3[0]);
  cmin.init(t, sh.k0[1], sh.nk[1], sh.kind[1],
            start_interval(t, sh, 1, D.exc_min, i0, 3 * nb + D.nm), sh.p3[1]);
#pragma unroll
  for (int j = 0; j < C; ++j) {
    if (j < cnt) {
      e.max_env[(int64_t)bidx * N + i0 + j] = cmax.eval(t, i0 + j);
      e.min_env[(int64_t)bidx * N + i0 + j] = cmin.eval(t, i0 + j);
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
