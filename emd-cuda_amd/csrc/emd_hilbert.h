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
EMDH_HD cpx mul(cpx a, cpx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// a * conj(b)
EMDH_HD cpx mulc(cpx a, cpx b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }
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
  return cpx{a.re * c - a.im * s, a.re * s + a.im * c};
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

// exp(-2 pi i q / 2^lg) for 0 <= q < 2^lg
EMDH_HD cpx expm(int q, int lg) {
  const double x = ldexp((double)q, 1 - lg);  // 2 q / 2^lg, exact
  double s, c;
#if defined(__HIP_DEVICE_COMPILE__)
  sincospi(x, &s, &c);
#else
  s = sin(M_PI * x);
  c = cos(M_PI * x);
#endif
  return cpx{c, -s};
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
};

// Forward (DIF) stage s on one lane's registers: for each butterfly, the length-r DFT,
// then output k times exp(-2 pi i q k / L).
template <int LOGF, int S>
EMDH_HD void stage_fwd(cpx* a, int t) {
  using P = Plan<LOGF>;
  constexpr int lr = P::log_radix(S), r = 1 << lr, lL = P::log_block(S);
#pragma unroll
  for (int i = 0; i < P::R / r; ++i) {
    cpx* v = a + i * r;
    dft<lr, -1>(v);
    if constexpr (lL > lr) {
      const cpx w = expm(P::qoff(S, t, i * r), lL);
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
EMDH_HD void stage_inv(cpx* a, int t) {
  using P = Plan<LOGF>;
  constexpr int lr = P::log_radix(S), r = 1 << lr, lL = P::log_block(S);
#pragma unroll
  for (int i = 0; i < P::R / r; ++i) {
    cpx* v = a + i * r;
    if constexpr (lL > lr) {
      const cpx w = expm(P::qoff(S, t, i * r), lL);
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
EMDH_HD double hilbert_kernel_value(int n, int N) {
  if (N & 1) {
    double s, c;
    const double x = (double)n / (2.0 * N);
#if defined(__HIP_DEVICE_COMPILE__)
    sincospi(x, &s, &c);
#else
    s = sin(M_PI * x);
    c = cos(M_PI * x);
#endif
    return (n & 1) ? c / (s * N) : -s / (c * N);
  }
  if (!(n & 1)) return 0.0;
  double s, c;
  const double x = (double)n / N;
#if defined(__HIP_DEVICE_COMPILE__)
  sincospi(x, &s, &c);
#else
  s = sin(M_PI * x);
  c = cos(M_PI * x);
#endif
  return 2.0 * c / (s * N);
}

// The kernel laid out circularly on F >= 2N-1 points, so that a length-F circular
// convolution reproduces the length-N one on its first N outputs: g[p] = q[p] for
// 1 <= p <= N-1, g[p] = -q[F-p] for F-N+1 <= p <= F-1 (the kernel is odd), else 0.
EMDH_HD double circular_kernel(int p, int N, int F) {
  if (p >= 1 && p <= N - 1) return hilbert_kernel_value(p, N);
  if (p >= F - N + 1 && p <= F - 1) return -hilbert_kernel_value(F - p, N);
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
