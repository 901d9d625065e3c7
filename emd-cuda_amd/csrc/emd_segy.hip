// emd_segy.hip -- SEG-Y trace samples (big-endian, rev 1/2 sample formats) decoded to the
// fp64 rows the sift kernels read.  No reference counterpart: the reference takes fp64
// arrays already in device memory (emd.py:60-69) and describes its inputs as seismic traces
// sampled at 4 ms (emd.py:28-30).  Sample formats follow the SEG-Y rev 1 standard (binary
// header bytes 3225-3226): 1 = 4-byte IBM hexadecimal float, 2 = 4-byte two's complement,
// 3 = 2-byte two's complement, 5 = 4-byte IEEE float, 8 = 1-byte two's complement, and
// rev 2's 6 = 8-byte IEEE double.  Every conversion is exact: an IBM fraction has 24 bits
// and its exponent range (16^-64 .. 16^63) lies inside fp64's normal range.
#include "../../include/emd_gfx950.h"

#include <hip/hip_runtime.h>

#include <cstdint>

#include "emd_sift.h"

namespace emdk {

namespace {

constexpr int kDecodeThreads = 256;
constexpr int kDecodeUnroll = 4;

constexpr int sample_bytes(int fmt) {
  return fmt == 3 ? 2 : fmt == 8 ? 1 : fmt == 6 ? 8 : 4;
}

// IBM single: sign bit, 7-bit excess-64 base-16 exponent, 24-bit fraction f:
// value = f / 2^24 * 16^(e - 64) = f * 2^(4e - 280).  (w >> 22) & 0x1fc is 4e.
__device__ __forceinline__ double ibm32_to_f64(uint32_t w) {
  const double mag = ldexp((double)(w & 0xffffffu), (int)((w >> 22) & 0x1fcu) - 280);
  return (w >> 31) ? -mag : mag;
}

// IEEE binary32 -> binary64 built from the fields, so the result does not depend on the
// wave's denormal mode: subnormals m * 2^-149, normals rebias the exponent by 1023 - 127,
// infinities keep their sign and NaNs come out quiet with the fraction kept in the top bits
// (what a host float -> double conversion gives).
__device__ __forceinline__ double ieee32_to_f64(uint32_t u) {
  const uint32_t e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
  const uint64_t s = (uint64_t)(u >> 31) << 63;
  uint64_t bits;
  if (e == 0) {
    bits = s | (uint64_t)__double_as_longlong(ldexp((double)m, -149));
  } else if (e == 255) {
    bits = s | 0x7ff0000000000000ull | ((uint64_t)m << 29) | (m ? (1ull << 51) : 0ull);
  } else {
    bits = s | ((uint64_t)(e + 896u) << 52) | ((uint64_t)m << 29);
  }
  return __longlong_as_double((long long)bits);
}

// w holds the sample's bytes in file order, most significant first.
template <int FMT>
__device__ __forceinline__ double decode(uint64_t w) {
  if constexpr (FMT == 1) return ibm32_to_f64((uint32_t)w);
  else if constexpr (FMT == 2) return (double)(int32_t)(uint32_t)w;
  else if constexpr (FMT == 3) return (double)(int16_t)(uint16_t)w;
  else if constexpr (FMT == 5) return ieee32_to_f64((uint32_t)w);
  else if constexpr (FMT == 6) return __longlong_as_double((long long)w);
  else return (double)(int8_t)(uint8_t)w;
}

template <int S>
__device__ __forceinline__ uint64_t load_be_aligned(const uint8_t* p) {
  if constexpr (S == 8) return __builtin_bswap64(*(const uint64_t*)p);
  else if constexpr (S == 4) return __builtin_bswap32(*(const uint32_t*)p);
  else if constexpr (S == 2) return __builtin_bswap16(*(const uint16_t*)p);
  else return *p;
}

template <int S>
__device__ __forceinline__ uint64_t load_be_bytes(const uint8_t* p) {
  uint64_t w = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) w = (w << 8) | p[k];
  return w;
}

// One workgroup per trace (grid-stride over traces).  Each lane issues kDecodeUnroll
// samples' loads before it converts and stores them; consecutive lanes touch consecutive
// samples, so loads and the 8-byte stores coalesce.  A trace whose first sample is not
// aligned to the sample size (2-byte and 1-byte formats with odd lengths) assembles its
// samples byte by byte.
template <int FMT>
__global__ __launch_bounds__(kDecodeThreads) void segy_decode_kernel(
    const uint8_t* __restrict__ bytes, const int64_t* __restrict__ trace_pos, int64_t stride,
    int header_bytes, const int64_t* __restrict__ offsets, int N, int B,
    double* __restrict__ out) {
  constexpr int S = sample_bytes(FMT);
  constexpr int kStep = kDecodeThreads * kDecodeUnroll;
  for (int b = blockIdx.x; b < B; b += gridDim.x) {
    const uint8_t* src = bytes + (trace_pos ? trace_pos[b] : (int64_t)b * stride) + header_bytes;
    const int64_t o = offsets ? offsets[b] : (int64_t)b * N;
    const int n = offsets ? (int)(offsets[b + 1] - o) : N;
    double* dst = out + o;
    if (((uintptr_t)src & (S - 1)) == 0) {
      for (int i0 = threadIdx.x; i0 < n; i0 += kStep) {
        uint64_t w[kDecodeUnroll];
#pragma unroll
        for (int u = 0; u < kDecodeUnroll; ++u) {
          const int i = i0 + u * kDecodeThreads;
          w[u] = i < n ? load_be_aligned<S>(src + (int64_t)i * S) : 0;
        }
#pragma unroll
        for (int u = 0; u < kDecodeUnroll; ++u) {
          const int i = i0 + u * kDecodeThreads;
          if (i < n) dst[i] = decode<FMT>(w[u]);
        }
      }
    } else {
      for (int i = threadIdx.x; i < n; i += kDecodeThreads)
        dst[i] = decode<FMT>(load_be_bytes<S>(src + (int64_t)i * S));
    }
  }
}

template <int FMT>
hipError_t launch_decode(const uint8_t* bytes, const int64_t* trace_pos, int64_t stride,
                         int header_bytes, const int64_t* offsets, int N, int B, double* out,
                         hipStream_t st) {
  const int grid = B < 16384 ? B : 16384;
  hipLaunchKernelGGL(segy_decode_kernel<FMT>, dim3(grid), dim3(kDecodeThreads), 0, st, bytes,
                     trace_pos, stride, header_bytes, offsets, N, B, out);
  return hipGetLastError();
}

}  // namespace

}  // namespace emdk

extern "C" {

int emd_segy_sample_bytes(int format) {
  switch (format) {
    case 1: case 2: case 3: case 5: case 6: case 8: return emdk::sample_bytes(format);
    default: return 0;
  }
}

int emd_segy_decode(const void* bytes, const int64_t* trace_pos, int64_t trace_stride,
                    int header_bytes, const int64_t* offsets, int N, int B, int format,
                    double* out, void* stream) {
  const int S = emd_segy_sample_bytes(format);
  if (S == 0)
    return emdk::fail(EMD_ERR_VALUE,
                      "unsupported SEG-Y sample format code %d (supported: 1, 2, 3, 5, 6, 8)",
                      format);
  if (B < 0) return emdk::fail(EMD_ERR_VALUE, "batch size must be >= 0, got %d", B);
  if (header_bytes < 0)
    return emdk::fail(EMD_ERR_VALUE, "header_bytes must be >= 0, got %d", header_bytes);
  if (!offsets && N < 0) return emdk::fail(EMD_ERR_VALUE, "N must be >= 0, got %d", N);
  if (!trace_pos && offsets)
    return emdk::fail(EMD_ERR_VALUE,
                      "ragged traces (offsets given) need their byte positions (trace_pos)");
  if (!trace_pos && trace_stride < (int64_t)header_bytes + (int64_t)N * S)
    return emdk::fail(EMD_ERR_VALUE,
                      "trace_stride %lld is shorter than a header (%d) plus %d samples of %d B",
                      (long long)trace_stride, header_bytes, N, S);
  if (B == 0) return EMD_OK;
  if (!bytes || !out) return emdk::fail(EMD_ERR_VALUE, "null bytes or out");
  const uint8_t* p = (const uint8_t*)bytes;
  const hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  switch (format) {
    case 1: e = emdk::launch_decode<1>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
    case 2: e = emdk::launch_decode<2>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
    case 3: e = emdk::launch_decode<3>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
    case 5: e = emdk::launch_decode<5>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
    case 6: e = emdk::launch_decode<6>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
    default: e = emdk::launch_decode<8>(p, trace_pos, trace_stride, header_bytes, offsets, N, B, out, st); break;
  }
  return e != hipSuccess ? emdk::hip_fail(e, "SEG-Y decode launch") : EMD_OK;
}

}  // extern "C"
