// emd_costmodel.hip -- per-trace features and the FP32 cost-model MLP (12 -> 64 -> 64 -> 3,
// GELU) with its forward, fused forward+backward and Adam kernels, behind the C ABI of
// include/emd_gfx950.h.
//
// The cost model only orders the persistent kernel's queue and sizes its LDS knot tile;
// the decomposition never reads a prediction, so outputs do not depend on it.  It has no
// counterpart in the reference (whose launch is one block per signal, kernels.py:1214-1220).
//
// Parameter layout (one flat fp32 vector, kParams floats):
//   W1[64][12] | b1[64] | W2[64][64] | b2[64] | W3[3][64] | b3[3]      (W[out][in])
// Inputs are standardised in the kernels: x = (f - mu) / sd with norm = [mu(12) | sd(12)].
#include "../../include/emd_gfx950.h"

#include <hip/hip_runtime.h>

#include <cstdint>

#include "emd_sift.h"

namespace emdcm {

constexpr int kF = EMD_N_FEATURES;  // 12
constexpr int kH = 64;
constexpr int kO = 3;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kH * kF;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kO * kH;
constexpr int kParams = kB3 + kO;  // 5187
static_assert(kParams == EMD_MLP_PARAMS, "parameter count");

constexpr int kTile = 64;      // samples per workgroup
constexpr int kThreads = 256;  // 4 waves: wave g owns samples [16g, 16g + 16)
constexpr int kPer = kTile / (kThreads / kH);  // 16 samples per lane
constexpr int kLd = kH + 1;    // padded row of the activation tiles (conflict-free columns)

constexpr int kSeg = 256;      // FFT segment
constexpr int kBins = kSeg / 2 + 1;
__constant__ int kBandEdge[9] = {1, 2, 4, 8, 16, 32, 64, 96, 129};

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------------------------------------
// Features: one workgroup of 256 lanes per trace.
//   0 log N | 1 log(1 + strict extrema) | 2 log(1 + sign changes) | 3 log(rms + 1e-12) |
//   4..11 log(band energy fraction + 1e-10) of the mean |rfft|^2 over non-overlapping
//   256-sample segments (one zero-padded segment when N < 256); bands over bins
//   [1,2) [2,4) [4,8) [8,16) [16,32) [32,64) [64,96) [96,129).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int bitrev8(int v) {
  return (int)(__builtin_bitreverse32((unsigned)v) >> 24);
}

// Orders one wave's LDS accesses (its lanes' earlier writes are visible to later reads).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// waves_per_eu(5) holds the kernel to 96 VGPRs (94 used, no spills): 5 workgroups per CU
// instead of the 4 its 100 unconstrained VGPRs allowed (0.40 -> 0.36 ms for 16384 x 2001).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void features_kernel(
    const double* __restrict__ values, const int64_t* __restrict__ offsets,
    float* __restrict__ feats) {
  // each wave transforms its own segments: radix-2 FFT of 256 points in LDS, fp64
  __shared__ double re[4][kSeg], im[4][kSeg];
  __shared__ double cs[kSeg / 2], sn[kSeg / 2];
  __shared__ double pw_w[4][kBins];
  __shared__ double pw[kBins];
  __shared__ double red[4][3];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t o = offsets[blockIdx.x];
  const int n = (int)(offsets[blockIdx.x + 1] - o);
  const double* x = values + o;
  if (t < kSeg / 2) {
    double s, c;
    sincospi((double)t / (kSeg / 2), &s, &c);  // exp(-2 pi i t / 256) = c - i s
    cs[t] = c;
    sn[t] = s;
  }
  double ext = 0, zc = 0, ss = 0;
  // four strided samples per step, all their loads issued before the first use
  for (int i0 = t; i0 < n; i0 += 4 * 256) {
    double vv[4], ww[4], pp[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 256 * u;
      vv[u] = i < n ? x[i] : 0.0;
      ww[u] = i + 1 < n ? x[i + 1] : 0.0;
      pp[u] = (i >= 1 && i < n) ? x[i - 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 256 * u;
      if (i < n) {
        const double v = vv[u];
        ss += v * v;
        if (i + 1 < n) {
          const double w = ww[u];
          if (v * w < 0) zc += 1;
          if (i >= 1 && (v - pp[u]) * (w - v) < 0) ext += 1;
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    ext += __shfl_down(ext, off);
    zc += __shfl_down(zc, off);
    ss += __shfl_down(ss, off);
  }
  if (lane == 0) {
    red[wv][0] = ext;
    red[wv][1] = zc;
    red[wv][2] = ss;
  }
  __syncthreads();  // twiddles ready
  const int nseg = n >= kSeg ? n / kSeg : 1;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;  // bins lane, lane + 64, 128
  double* R = re[wv];
  double* I = im[wv];
  for (int sgi = wv; sgi < nseg; sgi += 4) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = lane + 64 * q;
      const int i = sgi * kSeg + j;
      R[bitrev8(j)] = i < n ? x[i] : 0.0;
      I[bitrev8(j)] = 0.0;
    }
    wave_sync();
    for (int half = 1; half < kSeg; half <<= 1) {
      const int step = kSeg / (2 * half);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int b = lane + 64 * q;  // butterfly 0..127
        const int k = b & (half - 1);
        const int i0 = (b - k) * 2 + k, i1 = i0 + half;
        const double wr = cs[k * step], wi = -sn[k * step];
        const double xr = R[i1], xi = I[i1];
        const double tr = xr * wr - xi * wi, ti = xr * wi + xi * wr;
        const double ur = R[i0], ui = I[i0];
        R[i0] = ur + tr;
        I[i0] = ui + ti;
        R[i1] = ur - tr;
        I[i1] = ui - ti;
      }
      wave_sync();
    }
    p0 += R[lane] * R[lane] + I[lane] * I[lane];
    p1 += R[lane + 64] * R[lane + 64] + I[lane + 64] * I[lane + 64];
    if (lane == 0) p2 += R[128] * R[128] + I[128] * I[128];
    wave_sync();
  }
  pw_w[wv][lane] = p0;
  pw_w[wv][lane + 64] = p1;
  if (lane == 0) pw_w[wv][128] = p2;
  __syncthreads();
  if (t < kBins) pw[t] = (pw_w[0][t] + pw_w[1][t] + pw_w[2][t] + pw_w[3][t]) / nseg;
  __syncthreads();
  __shared__ double band[9];
  if (t < 9) {
    double e = 0;
    const int lo = t < 8 ? kBandEdge[t] : 0, hi = t < 8 ? kBandEdge[t + 1] : kBins;
    for (int k = lo; k < hi; ++k) e += pw[k];
    band[t] = e;  // band[0..7] bands, band[8] total
  }
  __syncthreads();
  if (t == 0) {
    const double E = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    const double Z = red[0][1] + red[1][1] + red[2][1] + red[3][1];
    const double S = red[0][2] + red[1][2] + red[2][2] + red[3][2];
    float* f = feats + (int64_t)blockIdx.x * kF;
    f[0] = (float)log((double)(n > 1 ? n : 1));
    f[1] = (float)log1p(E);
    f[2] = (float)log1p(Z);
    f[3] = (float)log(sqrt(n > 0 ? S / n : 0.0) + 1e-12);
    const double tot = band[8];
    for (int b = 0; b < 8; ++b) f[4 + b] = (float)log((tot > 0 ? band[b] / tot : 0.0) + 1e-10);
  }
}

// ---------------------------------------------------------------------------------------
// MLP tiles
// ---------------------------------------------------------------------------------------
struct MlpShared {
  float w[kParams];
  float w2t[kH * kH];        // W2 transposed: [in][out]
  float x[kTile][kF + 1];
  float z1[kTile][kLd], h1[kTile][kLd], z2[kTile][kLd], h2[kTile][kLd];
  float dy[kTile][kO];
  float y[kTile][kO];
  float lossw[4];
};

__device__ __forceinline__ void load_tile(MlpShared& sh, const float* __restrict__ params,
                                          const float* __restrict__ norm,
                                          const float* __restrict__ feats, int M) {
  const int t = threadIdx.x;
  for (int p = t; p < kParams; p += kThreads) sh.w[p] = params[p];
  for (int p = t; p < kH * kH; p += kThreads) {
    const int j = p / kH, k = p % kH;  // W2[j][k] -> w2t[k][j]
    sh.w2t[k * kH + j] = params[kW2 + p];
  }
  const int s0 = blockIdx.x * kTile;
  for (int e = t; e < kTile * kF; e += kThreads) {
    const int s = e / kF, k = e % kF;
    const float f = s0 + s < M ? feats[(int64_t)(s0 + s) * kF + k] : 0.0f;
    sh.x[s][k] = (f - norm[k]) / norm[kF + k];
  }
  __syncthreads();
}

__device__ __forceinline__ void forward_tile(MlpShared& sh) {
  const int t = threadIdx.x, j = t & (kH - 1), s0 = (t >> 6) * kPer;
  {
    float w[kF];
#pragma unroll
    for (int k = 0; k < kF; ++k) w[k] = sh.w[kW1 + j * kF + k];
    const float b = sh.w[kB1 + j];
#pragma unroll 4
    for (int q = 0; q < kPer; ++q) {
      float z = b;
#pragma unroll
      for (int k = 0; k < kF; ++k) z += w[k] * sh.x[s0 + q][k];
      sh.z1[s0 + q][j] = z;
      sh.h1[s0 + q][j] = gelu(z);
    }
  }
  __syncthreads();
  {
    float acc[kPer];
    const float b = sh.w[kB2 + j];
#pragma unroll
    for (int q = 0; q < kPer; ++q) acc[q] = b;
    for (int k = 0; k < kH; ++k) {
      const float w = sh.w2t[k * kH + j];
#pragma unroll
      for (int q = 0; q < kPer; ++q) acc[q] += w * sh.h1[s0 + q][k];
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      sh.z2[s0 + q][j] = acc[q];
      sh.h2[s0 + q][j] = gelu(acc[q]);
    }
  }
  __syncthreads();
  if (t < kO * kTile) {
    const int o = t >> 6, s = t & (kTile - 1);
    float y = sh.w[kB3 + o];
    for (int k = 0; k < kH; ++k) y += sh.w[kW3 + o * kH + k] * sh.h2[s][k];
    sh.y[s][o] = y;
  }
  __syncthreads();
}

// Inference only, with a small LDS footprint (about 9 KB): 16 samples per workgroup, the
// weights read through the caches instead of staged in LDS.  The plan's forward pass runs
// while a persistent sift still holds most of every CU's LDS, so a small workgroup starts
// as soon as one sift workgroup retires.  Every sum runs in the same order as in
// forward_tile, so the outputs are the same bits.
constexpr int kTileF = 16;
constexpr int kPerF = kTileF / (kThreads / kH);  // 4 samples per lane

__global__ __launch_bounds__(kThreads) void mlp_infer_kernel(const float* __restrict__ params,
                                                            const float* __restrict__ norm,
                                                            const float* __restrict__ feats,
                                                            int M, float* __restrict__ out) {
  __shared__ float x[kTileF][kF + 1];
  __shared__ float h1[kTileF][kLd], h2[kTileF][kLd];
  __shared__ float y[kTileF][kO];
  const int t = threadIdx.x, j = t & (kH - 1), s0 = (t >> 6) * kPerF;
  const int base = blockIdx.x * kTileF;
  for (int e = t; e < kTileF * kF; e += kThreads) {
    const int s = e / kF, k = e % kF;
    const float f = base + s < M ? feats[(int64_t)(base + s) * kF + k] : 0.0f;
    x[s][k] = (f - norm[k]) / norm[kF + k];
  }
  __syncthreads();
  {
    float w[kF];
#pragma unroll
    for (int k = 0; k < kF; ++k) w[k] = params[kW1 + j * kF + k];
    const float b = params[kB1 + j];
#pragma unroll
    for (int q = 0; q < kPerF; ++q) {
      float z = b;
#pragma unroll
      for (int k = 0; k < kF; ++k) z += w[k] * x[s0 + q][k];
      h1[s0 + q][j] = gelu(z);
    }
  }
  __syncthreads();
  {
    float acc[kPerF];
    const float b = params[kB2 + j];
#pragma unroll
    for (int q = 0; q < kPerF; ++q) acc[q] = b;
    const float* w2 = params + kW2 + j * kH;  // W2[j][k]
    for (int k = 0; k < kH; ++k) {
      const float w = w2[k];
#pragma unroll
      for (int q = 0; q < kPerF; ++q) acc[q] += w * h1[s0 + q][k];
    }
#pragma unroll
    for (int q = 0; q < kPerF; ++q) h2[s0 + q][j] = gelu(acc[q]);
  }
  __syncthreads();
  if (t < kO * kTileF) {
    const int o = t / kTileF, s = t % kTileF;
    float v = params[kB3 + o];
    for (int k = 0; k < kH; ++k) v += params[kW3 + o * kH + k] * h2[s][k];
    y[s][o] = v;
  }
  __syncthreads();
  if (t < kO * kTileF) {
    const int s = t / kO, o = t % kO;
    if (base + s < M) out[(int64_t)(base + s) * kO + o] = y[s][o];
  }
}

// Loss = (1/M) sum_s sum_o tw[o] (y_so - t_so)^2 over standardised targets.  Each workgroup
// writes its partial gradient (kParams floats) and partial loss to partial[blockIdx.x].
__global__ __launch_bounds__(kThreads) void mlp_backward_kernel(
    const float* __restrict__ params, const float* __restrict__ norm,
    const float* __restrict__ feats, const float* __restrict__ targets,
    const float* __restrict__ tw, int M, float* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpShared& sh = *reinterpret_cast<MlpShared*>(smem);
  load_tile(sh, params, norm, feats, M);
  forward_tile(sh);
  const int t = threadIdx.x, j = t & (kH - 1), g = t >> 6, s0 = g * kPer;
  const int base = blockIdx.x * kTile;
  float* part = partial + (int64_t)blockIdx.x * (kParams + 1);
  const float inv_m = 1.0f / (float)M;
  // dL/dy and the loss
  float l = 0.0f;
  if (t < kO * kTile) {
    const int o = t >> 6, s = t & (kTile - 1);
    float d = 0.0f;
    if (base + s < M) {
      const float e = sh.y[s][o] - targets[(int64_t)(base + s) * kO + o];
      l = tw[o] * e * e * inv_m;
      d = 2.0f * tw[o] * e * inv_m;
    }
    sh.dy[s][o] = d;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((t & 63) == 0) sh.lossw[g] = l;
  __syncthreads();
  // dz2 = (W3^T dy) * gelu'(z2), stored over z2
#pragma unroll 4
  for (int q = 0; q < kPer; ++q) {
    const int s = s0 + q;
    float d = 0.0f;
#pragma unroll
    for (int o = 0; o < kO; ++o) d += sh.w[kW3 + o * kH + j] * sh.dy[s][o];
    sh.z2[s][j] = d * gelu_grad(sh.z2[s][j]);
  }
  __syncthreads();
  // dz1 = (W2^T dz2) * gelu'(z1), stored over z1
  {
    float acc[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) acc[q] = 0.0f;
    for (int jj = 0; jj < kH; ++jj) {
      const float w = sh.w[kW2 + jj * kH + j];  // W2[jj][j]
#pragma unroll
      for (int q = 0; q < kPer; ++q) acc[q] += w * sh.z2[s0 + q][jj];
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) sh.z1[s0 + q][j] = acc[q] * gelu_grad(sh.z1[s0 + q][j]);
  }
  __syncthreads();
  // weight gradients over the tile's samples (invalid samples carry dy = 0)
  {
    // W2[j][k] for rows j in [16g, 16g+16), column k = lane
    const int k = j;
    float acc[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) acc[q] = 0.0f;
    for (int s = 0; s < kTile; ++s) {
      const float h = sh.h1[s][k];
#pragma unroll
      for (int q = 0; q < kPer; ++q) acc[q] += sh.z2[s][s0 + q] * h;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) part[kW2 + (s0 + q) * kH + k] = acc[q];
  }
  if (t < kO * kH) {
    const int o = t >> 6, k = t & (kH - 1);
    float a = 0.0f;
    for (int s = 0; s < kTile; ++s) a += sh.dy[s][o] * sh.h2[s][k];
    part[kW3 + o * kH + k] = a;
  }
  for (int e = t; e < kH * kF; e += kThreads) {
    const int jj = e / kF, k = e % kF;
    float a = 0.0f;
    for (int s = 0; s < kTile; ++s) a += sh.z1[s][jj] * sh.x[s][k];
    part[kW1 + e] = a;
  }
  if (t < kH) {
    float a1 = 0.0f, a2 = 0.0f;
    for (int s = 0; s < kTile; ++s) {
      a1 += sh.z1[s][t];
      a2 += sh.z2[s][t];
    }
    part[kB1 + t] = a1;
    part[kB2 + t] = a2;
  } else if (t < kH + kO) {
    float a = 0.0f;
    for (int s = 0; s < kTile; ++s) a += sh.dy[s][t - kH];
    part[kB3 + t - kH] = a;
  } else if (t == kH + kO) {
    part[kParams] = sh.lossw[0] + sh.lossw[1] + sh.lossw[2] + sh.lossw[3];
  }
}

// grad[p] = sum over workgroups in index order (deterministic); grad[kParams] = loss
__global__ __launch_bounds__(256) void reduce_partial_kernel(const float* __restrict__ partial,
                                                            int nwg, float* __restrict__ grad) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p > kParams) return;
  float a = 0.0f;
  for (int w = 0; w < nwg; ++w) a += partial[(int64_t)w * (kParams + 1) + p];
  grad[p] = a;
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ params,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ grad, float lr,
                                                  float beta1, float beta2, float eps,
                                                  float c1, float c2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kParams) return;
  const float g = grad[p];
  const float mp = beta1 * m[p] + (1.0f - beta1) * g;
  const float vp = beta2 * v[p] + (1.0f - beta2) * g * g;
  m[p] = mp;
  v[p] = vp;
  params[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

}  // namespace emdcm

namespace {
using emdk::fail;
int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? EMD_OK : emdk::hip_fail(e, what);
}
int mlp_smem_attr(const void* fn) {
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(emdcm::MlpShared));
  return e == hipSuccess ? EMD_OK : emdk::hip_fail(e, "hipFuncSetAttribute(mlp kernel)");
}
}  // namespace

extern "C" {

int emd_mlp_param_count(void) { return emdcm::kParams; }

int64_t emd_mlp_workspace_floats(int M) {
  if (M <= 0) return 0;
  return (int64_t)((M + emdcm::kTile - 1) / emdcm::kTile) * (emdcm::kParams + 1);
}

int emd_features(const double* values, const int64_t* offsets, int B, float* feats,
                 void* stream) {
  if (B < 0) return fail(EMD_ERR_VALUE, "batch size must be >= 0");
  if (B == 0) return EMD_OK;
  if (!offsets || !feats) return fail(EMD_ERR_VALUE, "null offsets or features pointer");
  hipLaunchKernelGGL(emdcm::features_kernel, dim3(B), dim3(256), 0, (hipStream_t)stream,
                     values, offsets, feats);
  return launch_status("features kernel");
}

int emd_mlp_forward(const float* params, const float* norm, const float* feats, int M,
                    float* out, void* stream) {
  if (M < 0) return fail(EMD_ERR_VALUE, "sample count must be >= 0");
  if (M == 0) return EMD_OK;
  if (!params || !norm || !feats || !out) return fail(EMD_ERR_VALUE, "null pointer");
  const int nwg = (M + emdcm::kTileF - 1) / emdcm::kTileF;
  hipLaunchKernelGGL(emdcm::mlp_infer_kernel, dim3(nwg), dim3(emdcm::kThreads), 0,
                     (hipStream_t)stream, params, norm, feats, M, out);
  return launch_status("mlp forward kernel");
}

int emd_mlp_backward(const float* params, const float* norm, const float* feats,
                     const float* targets, const float* target_weight, int M, float* workspace,
                     float* grad, void* stream) {
  if (M <= 0) return fail(EMD_ERR_VALUE, "sample count must be >= 1");
  if (!params || !norm || !feats || !targets || !target_weight || !workspace || !grad)
    return fail(EMD_ERR_VALUE, "null pointer");
  int rc = mlp_smem_attr(reinterpret_cast<const void*>(&emdcm::mlp_backward_kernel));
  if (rc) return rc;
  const int nwg = (M + emdcm::kTile - 1) / emdcm::kTile;
  hipLaunchKernelGGL(emdcm::mlp_backward_kernel, dim3(nwg), dim3(emdcm::kThreads),
                     sizeof(emdcm::MlpShared), (hipStream_t)stream, params, norm, feats,
                     targets, target_weight, M, workspace);
  rc = launch_status("mlp backward kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(emdcm::reduce_partial_kernel, dim3((emdcm::kParams + 1 + 255) / 256),
                     dim3(256), 0, (hipStream_t)stream, workspace, nwg, grad);
  return launch_status("gradient reduction kernel");
}

int emd_mlp_adam(float* params, float* m, float* v, const float* grad, int step, float lr,
                 float beta1, float beta2, float eps, void* stream) {
  if (step < 1) return fail(EMD_ERR_VALUE, "Adam step counts from 1");
  if (!params || !m || !v || !grad) return fail(EMD_ERR_VALUE, "null pointer");
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(emdcm::adam_kernel, dim3((emdcm::kParams + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, params, m, v, grad, lr, beta1, beta2, eps, c1, c2);
  return launch_status("adam kernel");
}

}  // extern "C"
