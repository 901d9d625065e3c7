// Host emulation of the Hilbert kernel's transform (emd-cuda_amd/csrc/emd_hilbert.h): the
// same plan, stage butterflies, twiddles, multiplier and kernel values as the gfx950
// kernel, with each LDS exchange done by visiting every lane in turn.  Test
// infrastructure: the CPU tests compare it with scipy.signal.hilbert and check the stage
// position maps, so the kernel's index arithmetic is exercised before it runs on a GPU.
#include <cmath>
#include <vector>

#include "../../emd-cuda_amd/csrc/emd_hilbert.h"

using namespace emdh;

namespace {

const double kTrig[2 * kTrigTerms] = EMDH_TRIG_POLY;
const double kAtan[kAtanTerms] = EMDH_ATAN_POLY;

template <int LOGF>
struct Emu {
  using P = Plan<LOGF>;
  std::vector<cpx> regs = std::vector<cpx>(P::T * P::R);
  std::vector<double> lds = std::vector<double>(P::XCH);

  cpx* lane(int t) { return &regs[(size_t)t * P::R]; }

  // LDS is poisoned before every half exchange, so a position read but never written
  // turns the result into NaN.
  template <int FROM, int TO>
  void exchange() {
    for (int part = 0; part < 2; ++part) {
      for (double& d : lds) d = NAN;
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e)
          lds[P::lds_base(FROM, t) + P::lds_off(FROM, e)] = part ? lane(t)[e].im : lane(t)[e].re;
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) {
          const double d = lds[P::lds_base(TO, t) + P::lds_off(TO, e)];
          (part ? lane(t)[e].im : lane(t)[e].re) = d;
        }
    }
  }
  template <int S>
  void forward() {
    if constexpr (S < P::NS) {
      if constexpr (S > 0) exchange<S - 1, S>();
      for (int t = 0; t < P::T; ++t) stage_fwd<LOGF, S>(lane(t), t, kTrig);
      forward<S + 1>();
    }
  }
  template <int S>
  void inverse() {
    if constexpr (S >= 0) {
      for (int t = 0; t < P::T; ++t) stage_inv<LOGF, S>(lane(t), t, kTrig);
      if constexpr (S > 0) exchange<S, S - 1>();
      inverse<S - 1>();
    }
  }

  // H x of nrows rows of len samples, as the kernel computes it for one trace
  void run(const double* x, int nrows, int len, bool exact, double* h) {
    std::vector<double> S((size_t)P::T * P::R);
    if (exact) {
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e)
          lane(t)[e] = cpx{circular_kernel(P::pos(0, t, e), len, P::F, kTrig), 0.0};
      forward<0>();
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) S[t * P::R + e] = -lane(t)[e].im * (1.0 / P::F);
    } else {
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) S[t * P::R + e] = P::sgn_over_f(t, e);
    }
    for (int r = 0; r < nrows; r += 2) {
      const bool two = r + 1 < nrows;
      const double* x1 = x + (size_t)r * len;
      const double* x2 = x1 + len;
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) {
          const int n = P::pos(0, t, e);
          lane(t)[e] = cpx{n < len ? x1[n] : 0.0, n < len && two ? x2[n] : 0.0};
        }
      forward<0>();
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) {
          const cpx v = lane(t)[e];
          const double s = S[t * P::R + e];
          lane(t)[e] = cpx{s * v.im, -s * v.re};
        }
      inverse<P::NS - 1>();
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) {
          const int n = P::pos(0, t, e);
          if (n >= len) continue;
          h[(size_t)r * len + n] = lane(t)[e].re;
          if (two) h[(size_t)(r + 1) * len + n] = lane(t)[e].im;
        }
    }
  }

  // 0 when every stage's (lane, register) -> position map is a bijection onto [0, F)
  // whose padded LDS index stays inside the exchange buffer and equals the lane's base plus
  // the register's offset, the twiddle offsets are the butterflies' q, and the digit
  // reversal is a bijection.
  static int check_maps() {
    for (int s = 0; s < P::NS; ++s) {
      std::vector<int> seen(P::F, 0);
      for (int t = 0; t < P::T; ++t)
        for (int e = 0; e < P::R; ++e) {
          const int p = P::pos(s, t, e);
          if (p < 0 || p >= P::F) return 1;
          if (P::pad(p) < 0 || P::pad(p) >= P::XCH) return 2;
          if (P::lds_base(s, t) + P::lds_off(s, e) != P::pad(p)) return 7;
          ++seen[p];
          const int lr = P::log_radix(s), lLr = P::log_block(s) - lr;
          if (P::qoff(s, t, e) != (p & ((1 << lLr) - 1))) return 3;
        }
      for (int c : seen)
        if (c != 1) return 4;
    }
    std::vector<int> seen(P::F, 0);
    for (int p = 0; p < P::F; ++p) {
      const int k = P::freq_of(p);
      if (k < 0 || k >= P::F) return 5;
      ++seen[k];
    }
    for (int c : seen)
      if (c != 1) return 6;
    return 0;
  }
};

template <int LOGF>
int run_class(const double* x, int nrows, int len, bool exact, double* h) {
  Emu<LOGF> emu;
  emu.run(x, nrows, len, exact, h);
  return 0;
}

}  // namespace

extern "C" {

// H x of nrows rows of len samples (one trace), transform chosen by n_fft as in
// emd_hilbert.  Returns the transform class (log2 of its length, 1 for H x = 0) or -1.
int hilbert_emu(const double* x, int nrows, int len, int n_fft, double* h) {
  bool exact;
  const int c = transform_class(len, n_fft, &exact);
  if (c == 0) return -1;
  if (c == kTrivial) {
    for (size_t i = 0; i < (size_t)nrows * len; ++i) h[i] = 0.0;
    return c;
  }
  switch (c) {
    case 2: run_class<2>(x, nrows, len, exact, h); break;
    case 3: run_class<3>(x, nrows, len, exact, h); break;
    case 4: run_class<4>(x, nrows, len, exact, h); break;
    case 5: run_class<5>(x, nrows, len, exact, h); break;
    case 6: run_class<6>(x, nrows, len, exact, h); break;
    case 7: run_class<7>(x, nrows, len, exact, h); break;
    case 8: run_class<8>(x, nrows, len, exact, h); break;
    case 9: run_class<9>(x, nrows, len, exact, h); break;
    case 10: run_class<10>(x, nrows, len, exact, h); break;
    case 11: run_class<11>(x, nrows, len, exact, h); break;
    case 12: run_class<12>(x, nrows, len, exact, h); break;
    case 13: run_class<13>(x, nrows, len, exact, h); break;
    case 14: run_class<14>(x, nrows, len, exact, h); break;
    default: return -1;
  }
  return c;
}

// The kernel's own atan2 and exp(-2 pi i q / 2^lg), for accuracy checks.
double hilbert_emu_atan2(double y, double x) { return atan2_poly(y, x, kAtan); }
void hilbert_emu_expm(int q, int lg, double* out) {
  const cpx w = expm(q, lg, kTrig);
  out[0] = w.re;
  out[1] = w.im;
}
double hilbert_emu_kernel_value(int n, int N) { return hilbert_kernel_value(n, N, kTrig); }

int hilbert_emu_class(int len, int n_fft, int* exact) {
  bool ex;
  const int c = transform_class(len, n_fft, &ex);
  *exact = ex ? 1 : 0;
  return c;
}

int hilbert_emu_check_maps(int logf) {
  switch (logf) {
    case 2: return Emu<2>::check_maps();
    case 3: return Emu<3>::check_maps();
    case 4: return Emu<4>::check_maps();
    case 5: return Emu<5>::check_maps();
    case 6: return Emu<6>::check_maps();
    case 7: return Emu<7>::check_maps();
    case 8: return Emu<8>::check_maps();
    case 9: return Emu<9>::check_maps();
    case 10: return Emu<10>::check_maps();
    case 11: return Emu<11>::check_maps();
    case 12: return Emu<12>::check_maps();
    case 13: return Emu<13>::check_maps();
    case 14: return Emu<14>::check_maps();
  }
  return -1;
}

}  // extern "C"
