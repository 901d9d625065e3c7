// emd_hilbert.h -- the Hilbert step's FFT: plan, butterflies and discrete Hilbert kernel,
// shared by the gfx950 kernel (emd_hilbert.hip) and a host emulation of it.
//
// A row pair (x1, x2) is one complex sequence u = x1 + i x2 of F = 2^LOGF points.  Each
// lane of the workgroup holds R points of it in registers; a forward pass of decimation-
// in-frequency stages (radix 16, the last one radix 2^(LOGF mod 4)) leaves U = FFT(u) in
// digit-reversed order, the pointwise product with the Hilbert multiplier -i S[k] is
// taken in that order, and the inverse stages, applied in reverse, bring the result back
// to natural order.  Between stages the points move through LDS.  Because both input
// rows are real and S is the transform of a real kernel, the result is
// H x1 + i H x2 (the imaginary parts of both analytic signals).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace emdh {

#define EMDH_HD __host__ __device__ __forceinline__

constexpr int kMinLogF = 2;   // shortest transform (rows of <= 2 samples have H x = 0)
constexpr int kMaxLogF = 14;  // longest transform: 16384 points
constexpr int kTrivial = 1;   // class of rows whose H x is zero (no transform)

struct cpx {
  double re, im;
};

EMDH_HD cpx add(cpx a, cpx b) { return {a.re + b.re, a.im + b.im}; }
EMDH_HD cpx sub(cpx a, cpx b) { return {a.re - b.re, a.im - b.im}; }
// Complex products with one fused multiply-add per part (correctly rounded on the host and
// on the device alike).
EMDH_HD cpx mul(cpx a, cpx b) {
  return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
// a * conj(b)
EMDH_HD cpx mulc(cpx a, cpx b) {
  return {fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -(a.re * b.im))};
}
// a * (DIR i): DIR = -1 multiplies by -i, +1 by i
template <int DIR>
EMDH_HD cpx rot90(cpx a) {
  return DIR < 0 ? cpx{a.im, -a.re} : cpx{-a.im, a.re};
}

// cos(2 pi m / 16); sin(2 pi m / 16) = cos(2 pi (m - 4) / 16).  m is a constant after
// unrolling, so the switch folds away.
EMDH_HD double cos16(int m) {
  constexpr double c1 = 0.92387953251128675613;  // cos(pi / 8)
  constexpr double c2 = 0.70710678118654752440;  // cos(pi / 4)
  constexpr double c3 = 0.38268343236508977173;  // sin(pi / 8)
  switch (m & 15) {
    case 0: return 1.0;
    case 1: case 15: return c1;
    case 2: case 14: return c2;
    case 3: case 13: return c3;
    case 4: case 12: return 0.0;
    case 5: case 11: return -c3;
    case 6: case 10: return -c2;
    case 7: case 9: return -c1;
    default: return -1.0;
  }
}

// a * exp(DIR 2 pi i m / 16) for a compile-time m (after unrolling); the multiples of a
// quarter turn are exact swaps and sign changes.
template <int DIR>
EMDH_HD cpx rot16(cpx a, int m) {
  m &= 15;
  if (m == 0) return a;
  if (m == 4) return rot90<DIR>(a);
  if (m == 8) return cpx{-a.re, -a.im};
  if (m == 12) return rot90<-DIR>(a);
  const double c = cos16(m), s = DIR * cos16(m - 4);
  if ((m & 3) == 2) {
    // |c| = |s| = sqrt(2)/2: (re c - im s, re s + im c) = c (re - (s/c) im, (s/c) re + im)
    const double r = s / c;  // +-1, folded
    return cpx{c * (a.re - r * a.im), c * (r * a.re + a.im)};
  }
  return cpx{fma(a.re, c, -(a.im * s)), fma(a.re, s, a.im * c)};
}

// In-place DFT of r = 2^LR points a[0..r): a[k] <- sum_n a[n] exp(DIR 2 pi i n k / r).
// r = A * B with n = n1 + A n2 and k = k2 + B k1: length-B transforms over n2, the
// twiddles exp(DIR 2 pi i n1 k2 / r), then length-A transforms over n1.
template <int DIR>
EMDH_HD void dft2(cpx* a) {
  const cpx t = a[1];
  a[1] = sub(a[0], t);
  a[0] = add(a[0], t);
}

template <int DIR>
EMDH_HD void dft4(cpx& a0, cpx& a1, cpx& a2, cpx& a3) {
  const cpx t0 = add(a0, a2), t1 = sub(a0, a2), t2 = add(a1, a3);
  const cpx t3 = rot90<DIR>(sub(a1, a3));
  a0 = add(t0, t2);
  a2 = sub(t0, t2);
  a1 = add(t1, t3);
  a3 = sub(t1, t3);
}

template <int LR, int DIR>
EMDH_HD void dft(cpx* a) {
  if constexpr (LR == 0) {
    return;
  } else if constexpr (LR == 1) {
    dft2<DIR>(a);
  } else if constexpr (LR == 2) {
    dft4<DIR>(a[0], a[1], a[2], a[3]);
  } else if constexpr (LR == 3) {
    // A = 2, B = 4: n = n1 + 2 n2, k = k2 + 4 k1
    dft4<DIR>(a[0], a[2], a[4], a[6]);
    dft4<DIR>(a[1], a[3], a[5], a[7]);
    // twiddle exp(DIR 2 pi i k2 / 8) on the n1 = 1 column = rot16 by 2 k2
    a[3] = rot16<DIR>(a[3], 2);
    a[5] = rot16<DIR>(a[5], 4);
    a[7] = rot16<DIR>(a[7], 6);
    cpx o[8];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      o[k2] = add(a[2 * k2], a[2 * k2 + 1]);
      o[k2 + 4] = sub(a[2 * k2], a[2 * k2 + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = o[k];
  } else {
    static_assert(LR == 4, "radix 16 at most");
    // A = B = 4: n = n1 + 4 n2, k = k2 + 4 k1
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4<DIR>(a[n1], a[n1 + 4], a[n1 + 8], a[n1 + 12]);
    // now a[n1 + 4 k2] holds column n1, frequency k2
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
      for (int k2 = 1; k2 < 4; ++k2) a[n1 + 4 * k2] = rot16<DIR>(a[n1 + 4 * k2], n1 * k2);
    cpx o[16];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      cpx b0 = a[4 * k2], b1 = a[4 * k2 + 1], b2 = a[4 * k2 + 2], b3 = a[4 * k2 + 3];
      dft4<DIR>(b0, b1, b2, b3);
      o[k2] = b0;
      o[k2 + 4] = b1;
      o[k2 + 8] = b2;
      o[k2 + 12] = b3;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = o[k];
  }
}

// sin and cos of pi x from two polynomials on the quarter r in [0, 1/4]:
// sin(pi r) = r S(r^2), cos(pi r) = 1 + r^2 C(r^2), S then C in the table (highest degree
// first, derived by emd-cuda_amd/csrc/poly_coefficients.py; errors below 1e-20 before
// rounding).  Arguments are reduced to the quarter in integers, so the only rounding
// before the polynomials is that of r itself.  The device kernel passes the table through
// memory (scalar registers) for the same reason as atan2's.
#define EMDH_TRIG_POLY                                                                       \
  {-2.1717401400090533e-05, 0.00046628272951211076, -0.007370429884712669,                  \
   0.08214588658005621,     -0.5992645293202978,    2.5501640398773415,                     \
   -5.16771278004997,       3.141592653589793,                                              \
   4.268505646755225e-06,   -0.0001046345989636836, 0.0019295741241061913,                  \
   -0.025806891384587305,   0.23533063035880689,    -1.3352627688545888,                    \
   4.0587121264167685,      -4.934802200544679}
constexpr int kTrigTerms = 8;

EMDH_HD void sincospi_quarter(double r, const double* T, double* s, double* c) {
  const double u = r * r;
  double p = T[0], q = T[kTrigTerms];
#pragma unroll
  for (int i = 1; i < kTrigTerms; ++i) {
    p = fma(p, u, T[i]);
    q = fma(q, u, T[kTrigTerms + i]);
  }
  *s = r * p;
  *c = fma(u, q, 1.0);
}

// sin and cos of pi num / den for 0 <= num < 2 den: x >= 1 is x - 1 with both signs
// flipped, x > 1/2 is 1 - x with the cosine's sign flipped, x > 1/4 is the cofunction of
// 1/2 - x.  `lg_den` >= 0 when den = 2^lg_den (r = ldexp, exact), else r = num / den.
EMDH_HD void sincospi_ratio(int num, int den, int lg_den, const double* T, double* s,
                            double* c) {
  double ss = 1.0, cs = 1.0;
  if (num >= den) {
    num -= den;
    ss = -1.0;
    cs = -1.0;
  }
  if (2 * num > den) {
    num = den - num;
    cs = -cs;
  }
  const bool co = 4 * num > den;
  int rn = num, rd_lg = lg_den + 0;
  int rd = den;
  if (co) {  // r = 1/2 - num/den = (den - 2 num) / (2 den)
    rn = den - 2 * num;
    rd = 2 * den;
    rd_lg = lg_den + 1;
  }
  const double r = lg_den >= 0 ? ldexp((double)rn, -rd_lg) : (double)rn / (double)rd;
  double s0, c0;
  sincospi_quarter(r, T, &s0, &c0);
  *s = ss * (co ? c0 : s0);
  *c = cs * (co ? s0 : c0);
}

// exp(-2 pi i q / 2^lg) for 0 <= q < 2^lg
EMDH_HD cpx expm(int q, int lg, const double* T) {
  double s, c;
  sincospi_ratio(2 * q, 1 << lg, lg, T, &s, &c);  // angle pi (2q / 2^lg)
  return cpx{c, -s};
}

// atan2 for the phase outputs: t = min(|x|,|y|) / max(|x|,|y|) (one correctly rounded
// division), atan(t) = t + t^3 P(t^2) on [0, 1] by Horner with fused multiply-adds, then
// the quadrant: pi/2 - a, pi - a or pi/2 + a with pi/2 and pi as two-double sums.  Within
// 1.5 ulp of the exact value; signed zeros as C's atan2 (atan2(+-0, -0) = +-pi).  P
// (highest degree first) is derived by emd-cuda_amd/csrc/atan_coefficients.py.  The
// coefficients are read from memory (the caller passes the table) so that the device
// kernel keeps them in scalar registers instead of 42 vector registers.
#define EMDH_ATAN_POLY                                                                     \
  {-1.1832505417555692e-05, 0.0001368724853148122, -0.0007518472526973822,                \
   0.002622977891906089,    -0.006575683344151699, 0.012756172907694298,                  \
   -0.020238706986393514,   0.027567942297344678,  -0.033750132001619734,                 \
   0.03872640214042632,     -0.04308119655330471,  0.04752086773576656,                   \
   -0.05261265735709454,    0.05882074956371294,   -0.06666636435777692,                  \
   0.07692305354678655,     -0.09090908969557403,  0.11111111107234799,                   \
   -0.14285714285648404,    0.19999999999999554,   -0.3333333333333333}
constexpr int kAtanTerms = 21;

EMDH_HD double atan2_poly(double y, double x, const double* P) {
  const double ax = fabs(x), ay = fabs(y);
  const bool sw = ay > ax, ng = signbit(x);
  const double mx = sw ? ay : ax, mn = sw ? ax : ay;
  const double t = mx == 0.0 ? 0.0 : mn / mx;
  const double s = t * t;
  double p = P[0];
#pragma unroll
  for (int i = 1; i < kAtanTerms; ++i) p = fma(p, s, P[i]);
  double a = fma(t * s, p, t);
  constexpr double pio2_hi = 1.5707963267948966, pio2_lo = 6.123233995736766e-17;
  constexpr double pi_hi = 3.141592653589793, pi_lo = 1.2246467991473532e-16;
  if (sw && ng) a = (pio2_hi + a) + pio2_lo;
  else if (sw) a = (pio2_hi - a) + pio2_lo;
  else if (ng) a = (pi_hi - a) + pi_lo;
  return copysign(a, y);
}

// Compile-time plan of an F = 2^LOGF point transform: stages of radix 16, the last one of
// radix 2^(LOGF mod 4); each lane holds R = 2^LOGR points (16, or 32 for the two longest
// transforms so their workgroups stay at 256 and 512 lanes) and so runs R / r butterflies
// of a radix-r stage.
template <int LOGF>
struct Plan {
  static_assert(LOGF >= kMinLogF && LOGF <= kMaxLogF, "transform length out of range");
  static constexpr int F = 1 << LOGF;
  static constexpr int LOGR = LOGF < 4 ? LOGF : (LOGF >= 13 ? 5 : 4);
  static constexpr int R = 1 << LOGR;
  static constexpr int T = F / R;                    // lanes holding points
  static constexpr int NFULL = LOGF / 4;             // stages of radix 16
  static constexpr int LOGREM = LOGF % 4;            // bits of the last stage (0: none)
  static constexpr int NS = NFULL + (LOGREM ? 1 : 0);
  static constexpr int WG = T < 64 ? 64 : T;
  // LDS exchange buffer: one double per point plus one pad double per 16
  static constexpr int XCH = F + F / 16;

  EMDH_HD static constexpr int log_radix(int s) { return s < NFULL ? 4 : LOGREM; }
  EMDH_HD static constexpr int log_block(int s) { return LOGF - 4 * s; }

  // Position of register e of lane t in stage s: the lane's butterflies are
  // beta = t + i T (i = e / r); butterfly beta of a block of L points combines the points
  // b L + q + j L/r, j = e mod r, with b = beta / (L/r) and q = beta mod (L/r).  Stage 0's
  // positions are where the lane loads its points and, after the inverse pass, where it
  // finds the results.
  EMDH_HD static int pos(int s, int t, int e) {
    const int lr = log_radix(s), lLr = log_block(s) - lr;
    const int i = e >> lr, j = e & ((1 << lr) - 1);
    const int beta = t + i * T;
    const int b = beta >> lLr, q = beta & ((1 << lLr) - 1);
    return (b << (lLr + lr)) + q + (j << lLr);
  }
  // Twiddle offset q of the butterfly of register e of lane t in stage s.
  EMDH_HD static int qoff(int s, int t, int e) {
    const int lr = log_radix(s), lLr = log_block(s) - lr;
    const int beta = t + (e >> lr) * T;
    return beta & ((1 << lLr) - 1);
  }
  // Frequency held at position p after the forward pass (digit reversal: the first
  // stage's digit is the most significant of the position, the least of the frequency).
  EMDH_HD static int freq_of(int p) {
    int k = 0, shift = 0, bits = LOGF;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int lr = log_radix(s);
      bits -= lr;
      k |= ((p >> bits) & ((1 << lr) - 1)) << shift;
      shift += lr;
    }
    return k;
  }
  // -i sgn(k) / F, the multiplier of a power-of-two transform, as S / F with
  // S = sgn(k) (0 at DC and Nyquist) for the point of register e of lane t
  EMDH_HD static double sgn_over_f(int t, int e) {
    const int k = freq_of(pos(NS - 1, t, e));
    const double sg = (k == 0 || k == F / 2) ? 0.0 : (k < F / 2 ? 1.0 : -1.0);
    return sg * (1.0 / F);
  }
  EMDH_HD static int pad(int p) { return p + (p >> 4); }
  // LDS index of register e of lane t in stage s, as the lane's base plus a per-register
  // constant: pad(pos(s, t, e)) - pad(pos(s, t, 0)) does not depend on t (a butterfly's
  // points never straddle a pad boundary differently from lane to lane), so the offsets
  // fold into the LDS instructions' immediates.
  EMDH_HD static int lds_base(int s, int t) { return pad(pos(s, t, 0)); }
  EMDH_HD static int lds_off(int s, int e) { return pad(pos(s, 0, e)) - pad(pos(s, 0, 0)); }
};

// Forward (DIF) stage s on one lane's registers: for each butterfly, the length-r DFT,
// then output k times exp(-2 pi i q k / L).
template <int LOGF, int S>
EMDH_HD void stage_fwd(cpx* a, int t, const double* T) {
  using P = Plan<LOGF>;
  constexpr int lr = P::log_radix(S), r = 1 << lr, lL = P::log_block(S);
#pragma unroll
  for (int i = 0; i < P::R / r; ++i) {
    cpx* v = a + i * r;
    dft<lr, -1>(v);
    if constexpr (lL > lr) {
      const cpx w = expm(P::qoff(S, t, i * r), lL, T);
      cpx wk = w;
#pragma unroll
      for (int k = 1; k < r; ++k) {
        v[k] = mul(v[k], wk);
        if (k + 1 < r) wk = mul(wk, w);
      }
    }
  }
}

// Inverse of stage_fwd (without the 1/r): output k times exp(+2 pi i q k / L), then the
// length-r inverse DFT.
template <int LOGF, int S>
EMDH_HD void stage_inv(cpx* a, int t, const double* T) {
  using P = Plan<LOGF>;
  constexpr int lr = P::log_radix(S), r = 1 << lr, lL = P::log_block(S);
#pragma unroll
  for (int i = 0; i < P::R / r; ++i) {
    cpx* v = a + i * r;
    if constexpr (lL > lr) {
      const cpx w = expm(P::qoff(S, t, i * r), lL, T);
      cpx wk = w;
#pragma unroll
      for (int k = 1; k < r; ++k) {
        v[k] = mulc(v[k], wk);
        if (k + 1 < r) wk = mul(wk, w);
      }
    }
    dft<lr, 1>(v);
  }
}

// Discrete Hilbert kernel of length N (N >= 3), 1 <= n <= N-1: the impulse response of
// the multiplier -i sgn(k) of an N-point DFT (k = 1..ceil(N/2)-1 positive, the DC and,
// for even N, the Nyquist bin zero):
//   N odd:  cot(pi n / 2N) / N for odd n, -tan(pi n / 2N) / N for even n;
//   N even: 2 cot(pi n / N) / N for odd n, 0 for even n.
// (The half-angle forms of (cos(pi n/N) - (-1)^n) / (N sin(pi n/N)), free of the
// cancellation of the latter at small n.)
EMDH_HD double hilbert_kernel_value(int n, int N, const double* T) {
  double s, c;
  if (N & 1) {
    sincospi_ratio(n, 2 * N, -1, T, &s, &c);  // angle pi n / 2N
    return (n & 1) ? c / (s * N) : -s / (c * N);
  }
  if (!(n & 1)) return 0.0;
  sincospi_ratio(n, N, -1, T, &s, &c);  // angle pi n / N
  return 2.0 * c / (s * N);
}

// The kernel laid out circularly on F >= 2N-1 points, so that a length-F circular
// convolution reproduces the length-N one on its first N outputs: g[p] = q[p] for
// 1 <= p <= N-1, g[p] = -q[F-p] for F-N+1 <= p <= F-1 (the kernel is odd), else 0.
EMDH_HD double circular_kernel(int p, int N, int F, const double* T) {
  if (p >= 1 && p <= N - 1) return hilbert_kernel_value(p, N, T);
  if (p >= F - N + 1 && p <= F - 1) return -hilbert_kernel_value(F - p, N, T);
  return 0.0;
}

// Transform class of a row of `len` samples: log2 of its transform length, kTrivial for
// rows whose H x is zero, 0 for rows beyond the limits.  n_fft: 0 = a DFT of the row's
// own length (taken as a circular convolution on 2^ceil(log2(2 len - 1)) points, or
// directly when len is a power of two), -1 = the smallest power of two >= len, > 0 = that
// power of two for every row.  `exact` is set for the convolution form.
// ceil(log2(v)) for 1 <= v <= 2^30 (0 for v <= 1; the loop stops at 30 for larger v)
EMDH_HD int ceil_log2(int v) {
  int l = 0;
  while (l < 30 && (1 << l) < v) ++l;
  return l;
}

EMDH_HD int transform_class(int len, int n_fft, bool* exact) {
  *exact = false;
  if (len > (1 << kMaxLogF)) return 0;
  if (n_fft > 0) {
    if (len > n_fft) return 0;
    const int l = ceil_log2(n_fft);
    return l < kMinLogF ? kTrivial : (l <= kMaxLogF ? l : 0);
  }
  if (len <= 2) return kTrivial;
  const int lp = ceil_log2(len);
  if (n_fft < 0 || (1 << lp) == len) return lp <= kMaxLogF ? lp : 0;
  *exact = true;
  const int l = ceil_log2(2 * len - 1);
  return l <= kMaxLogF ? l : 0;
}

}  // namespace emdh
