// emd_dispatch.hip -- host-side selection among the compiled sift kernel instances.
#include "emd_sift_dev.h"

namespace emdk {

__global__ void reset_counter_kernel(unsigned* q) { *q = 0u; }

#define EMD_DECL_WG(WG)                        \
  void* sift_instance_##WG##_1(int which);     \
  void* sift_instance_##WG##_2(int which);     \
  void* sift_instance_##WG##_4(int which);     \
  void* sift_instance_##WG##_8(int which);     \
  void* sift_instance_##WG##_16(int which);    \
  void* sift_instance_##WG##_32(int which);    \
  void* sift_instance_##WG##_64(int which);
EMD_DECL_WG(64)
EMD_DECL_WG(128)
EMD_DECL_WG(256)
#undef EMD_DECL_WG


int64_t tile_bytes(int slots) {
  return (((int64_t)slots * kSlotBytes) + 255) / 256 * 256;
}

int static_lds_bytes(int wg) {
  switch (wg) {
    case 64: return (int)sizeof(SiftShared<64>);
    case 128: return (int)sizeof(SiftShared<128>);
    case 256: return (int)sizeof(SiftShared<256>);
  }
  return 0;
}

int chunk_for(int n, int wg) {
  int per = (n + wg - 1) / wg;
  if (per < 1) per = 1;
  int c = 1;
  while (c < per) c <<= 1;
  return c;
}

#define EMD_SWITCH_C(WG, which, chunk)                \
  switch (chunk) {                                    \
    case 1: return sift_instance_##WG##_1(which);     \
    case 2: return sift_instance_##WG##_2(which);     \
    case 4: return sift_instance_##WG##_4(which);     \
    case 8: return sift_instance_##WG##_8(which);     \
    case 16: return sift_instance_##WG##_16(which);   \
    case 32: return sift_instance_##WG##_32(which);   \
    case 64: return sift_instance_##WG##_64(which);   \
  }                                                   \
  return nullptr;

static void* instance(int which, int wg, int chunk) {
  switch (wg) {
    case 64: { EMD_SWITCH_C(64, which, chunk) }
    case 128: { EMD_SWITCH_C(128, which, chunk) }
    case 256: { EMD_SWITCH_C(256, which, chunk) }
  }
  return nullptr;
}

static void* sift_kernel_ptr(int persistent, int global_tile, int wg, int chunk) {
  return instance((persistent ? 1 : 0) + (global_tile ? 3 : 0), wg, chunk);
}
static void* envelope_kernel_ptr(int global_tile, int wg, int chunk) {
  return instance(global_tile ? 5 : 2, wg, chunk);
}

static size_t dyn_lds_bytes(int global_tile, const SiftParams& P) {
  return global_tile ? 0 : (size_t)P.tile_slots * kSlotBytes;
}

int launch_sift(int persistent, int global_tile, int wg, int chunk, int grid, const Batch& b,
                const SiftParams& P, hipStream_t stream) {
  void* fn = sift_kernel_ptr(persistent, global_tile, wg, chunk);
  if (!fn) return (int)hipErrorInvalidValue;
  if (grid <= 0) return (int)hipSuccess;
  const size_t lds = dyn_lds_bytes(global_tile, P);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return (int)e;
  Batch bb = b;
  SiftParams pp = P;
  void* args[] = {&bb, &pp};
  return (int)hipLaunchKernel(fn, dim3(grid), dim3(wg), args, lds, stream);
}

int launch_envelope(int global_tile, int wg, int chunk, const EnvBatch& e,
                    const SiftParams& P, hipStream_t stream) {
  void* fn = envelope_kernel_ptr(global_tile, wg, chunk);
  if (!fn) return (int)hipErrorInvalidValue;
  if (e.B <= 0) return (int)hipSuccess;
  const size_t lds = dyn_lds_bytes(global_tile, P);
  hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (r != hipSuccess) return (int)r;
  EnvBatch ee = e;
  SiftParams pp = P;
  void* args[] = {&ee, &pp};
  return (int)hipLaunchKernel(fn, dim3(e.B), dim3(wg), args, lds, stream);
}

int reset_queue(unsigned* q, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, q);
  return (int)hipGetLastError();
}

int max_resident_wgs(int persistent, int global_tile, int wg, int chunk, size_t dyn_lds) {
  void* fn = sift_kernel_ptr(persistent, global_tile, wg, chunk);
  if (!fn) return 0;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds) !=
      hipSuccess)
    return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, wg, dyn_lds) != hipSuccess) return 0;
  return nb;
}

}  // namespace emdk
