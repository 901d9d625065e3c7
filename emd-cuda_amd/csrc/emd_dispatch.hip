// emd_dispatch.hip -- host-side selection among the compiled sift kernel instances.
#include "emd_sift_dev.h"

namespace emdk {

__global__ void reset_counter_kernel(unsigned* q) { *q = 0u; }

void* sift_instance_64(int which);
void* sift_instance_128(int which);
void* sift_instance_256(int which);
void* sift_instance_512(int which);
void* sift_instance_1024(int which);

int64_t tile_bytes(int slots, int slot_bytes) {
  return (((int64_t)slots * slot_bytes) + 255) / 256 * 256;
}

int static_lds_bytes(int wg) {
  switch (wg) {
    case 64: return (int)sizeof(SiftShared<64>);
    case 128: return (int)sizeof(SiftShared<128>);
    case 256: return (int)sizeof(SiftShared<256>);
    case 512: return (int)sizeof(SiftShared<512>);
    case 1024: return (int)sizeof(SiftShared<1024>);
  }
  return 0;
}

size_t dyn_lds_bytes(int tile_slots, int max_n, int slot_bytes) {
  const size_t nbcap = (size_t)(max_n + 63) / 64;
  return nbcap * kWorkBlockBytes + (size_t)tile_slots * slot_bytes;
}

static void* instance(int which, int wg) {
  switch (wg) {
    case 64: return sift_instance_64(which);
    case 128: return sift_instance_128(which);
    case 256: return sift_instance_256(which);
    case 512: return sift_instance_512(which);
    case 1024: return sift_instance_1024(which);
  }
  return nullptr;
}

// timed: the instance whose positions come from a time axis
static void* sift_kernel_ptr(int persistent, int wg, int timed) {
  return instance((persistent ? 1 : 0) + (timed ? 3 : 0), wg);
}
static void* envelope_kernel_ptr(int wg, int timed) { return instance(timed ? 5 : 2, wg); }

int prepare_kernels(int lds_per_workgroup) {
  // Allow each kernel the whole per-workgroup LDS once, so launches (possibly under
  // stream capture) never change function attributes.
  for (int wg : {64, 128, 256, 512, 1024}) {
    const int dyn = lds_per_workgroup - static_lds_bytes(wg);
    for (int which = 0; which < 6; ++which) {
      hipError_t e = hipFuncSetAttribute(instance(which, wg),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
      if (e != hipSuccess) return (int)e;
    }
  }
  return (int)hipSuccess;
}

int launch_sift(int persistent, int wg, int grid, const Batch& b, const SiftParams& P,
                hipStream_t stream) {
  const int timed = b.times != nullptr;
  void* fn = sift_kernel_ptr(persistent, wg, timed);
  if (!fn) return (int)hipErrorInvalidValue;
  if (grid <= 0) return (int)hipSuccess;
  const size_t lds = dyn_lds_bytes(P.tile_slots, P.max_n, timed ? kSlotBytesTime : kSlotBytes);
  Batch bb = b;
  SiftParams pp = P;
  void* args[] = {&bb, &pp};
  return (int)hipLaunchKernel(fn, dim3(grid), dim3(wg), args, lds, stream);
}

int launch_envelope(int wg, const EnvBatch& e, const SiftParams& P, hipStream_t stream) {
  const int timed = e.times != nullptr;
  void* fn = envelope_kernel_ptr(wg, timed);
  if (!fn) return (int)hipErrorInvalidValue;
  if (e.B <= 0) return (int)hipSuccess;
  const size_t lds = dyn_lds_bytes(P.tile_slots, P.max_n, timed ? kSlotBytesTime : kSlotBytes);
  EnvBatch ee = e;
  SiftParams pp = P;
  void* args[] = {&ee, &pp};
  return (int)hipLaunchKernel(fn, dim3(e.B), dim3(wg), args, lds, stream);
}

int reset_queue(unsigned* q, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, q);
  return (int)hipGetLastError();
}

int max_resident_wgs(int persistent, int wg, size_t dyn_lds, int timed) {
  void* fn = sift_kernel_ptr(persistent, wg, timed);
  if (!fn) return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, wg, dyn_lds) != hipSuccess) return 0;
  return nb;
}

}  // namespace emdk
