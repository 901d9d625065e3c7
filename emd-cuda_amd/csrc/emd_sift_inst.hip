// emd_sift_inst.hip -- one (workgroup size, chunk) instance of the sift kernels.
// Compiled once per instance with -DEMD_WG=<64|128|256> -DEMD_C=<1..64> so the
// instances build in parallel; emd_dispatch.hip selects among them at run time.
#include "emd_sift_dev.h"

#ifndef EMD_WG
#error "EMD_WG must be defined"
#endif
#ifndef EMD_C
#error "EMD_C must be defined"
#endif

#define EMD_PASTE2(a, b, c) a##b##_##c
#define EMD_PASTE(a, b, c) EMD_PASTE2(a, b, c)
#define EMD_INST_FN EMD_PASTE(sift_instance_, EMD_WG, EMD_C)

namespace emdk {
// which: 0 plain, 1 persistent, 2 envelope (LDS knot tile); 3, 4, 5 the same kernels with
// the knot tile in global memory (built for 256-thread workgroups only).
void* EMD_INST_FN(int which) {
  switch (which) {
    case 0: return reinterpret_cast<void*>(&emd_plain_kernel<EMD_WG, EMD_C, false>);
    case 1: return reinterpret_cast<void*>(&emd_persistent_kernel<EMD_WG, EMD_C, false>);
#if EMD_WG == 256
    case 2: return reinterpret_cast<void*>(&emd_envelope_kernel<EMD_WG, EMD_C, false>);
    case 3: return reinterpret_cast<void*>(&emd_plain_kernel<EMD_WG, EMD_C, true>);
    case 4: return reinterpret_cast<void*>(&emd_persistent_kernel<EMD_WG, EMD_C, true>);
    case 5: return reinterpret_cast<void*>(&emd_envelope_kernel<EMD_WG, EMD_C, true>);
#endif
  }
  return nullptr;
}
}  // namespace emdk
