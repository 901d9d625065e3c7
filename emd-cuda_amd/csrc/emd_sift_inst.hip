// emd_sift_inst.hip -- the sift kernels for one workgroup size.
// Compiled once per size with -DEMD_WG=<64|128|256|512|1024> so the instances build in
// parallel;
// emd_dispatch.hip selects among them at run time.
#include "emd_sift_dev.h"

#ifndef EMD_WG
#error "EMD_WG must be defined"
#endif

#define EMD_PASTE2(a, b) a##b
#define EMD_PASTE(a, b) EMD_PASTE2(a, b)
#define EMD_INST_FN EMD_PASTE(sift_instance_, EMD_WG)

namespace emdk {
// which: 0 plain, 1 persistent, 2 envelope (positions = sample indices); 3-5 the same
// with positions from a time axis
void* EMD_INST_FN(int which) {
  switch (which) {
    case 0: return reinterpret_cast<void*>(&emd_plain_kernel<EMD_WG, IndexAxis>);
    case 1: return reinterpret_cast<void*>(&emd_persistent_kernel<EMD_WG, IndexAxis>);
    case 2: return reinterpret_cast<void*>(&emd_envelope_kernel<EMD_WG, IndexAxis>);
    case 3: return reinterpret_cast<void*>(&emd_plain_kernel<EMD_WG, TimeAxis>);
    case 4: return reinterpret_cast<void*>(&emd_persistent_kernel<EMD_WG, TimeAxis>);
    case 5: return reinterpret_cast<void*>(&emd_envelope_kernel<EMD_WG, TimeAxis>);
  }
  return nullptr;
}
}  // namespace emdk
