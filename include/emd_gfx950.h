/* emd_gfx950.h -- C ABI of the MI355X (gfx950) batched EMD library, libemd_gfx950.so.
 *
 * Plain C types only: device pointers are the caller's (hipMalloc'd or any torch/HIP
 * allocation on the current device), sizes are ints, streams are opaque hipStream_t
 * values passed as void* (NULL = the null stream).  Every entry point returns EMD_OK (0)
 * or a negative status; emd_last_error() gives the message of the last failure on the
 * calling thread.  A handle is not thread-safe (same as the reference launcher).
 *
 * The library never frees caller memory.  Scratch (global knot tiles, queue counter) is
 * owned by the handle and grown on demand.
 */
#ifndef EMD_GFX950_H
#define EMD_GFX950_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMD_OK 0
#define EMD_ERR_VALUE (-1)   /* bad argument: dtype/shape/size/config (Python ValueError) */
#define EMD_ERR_RUNTIME (-2) /* HIP runtime or launch failure (Python RuntimeError) */
#define EMD_ERR_MEMORY (-3)  /* device allocation failed */

/* Compile-time configuration of the reference, EMDKernelConfig (kernels.py:1170-1176):
 * max_extrema=1024, nbsym=2, max_iteration=1000, max_imfs=16, threads_per_block=128.
 * tile_slots: knot slots held in LDS per workgroup for both envelopes (0 = hold
 * 2*min(max_extrema, N) + 4*nbsym, the reference's "capacity = shared memory" rule,
 * kernels.py:1273-1275).  A sift whose knots outgrow the LDS tile keeps them in the
 * workgroup's own global-memory tile; results are bit-identical whatever the tile. */
typedef struct emd_config {
  int max_extrema;
  int nbsym;
  int max_iteration;
  int max_imfs;
  int threads_per_block;
  int tile_slots;
} emd_config;

typedef struct emd_handle emd_handle;

/* Launch shape for the ragged/persistent entry point (chosen by the autotuner). */
typedef struct emd_launch {
  int threads;     /* 64, 128 or 256 */
  int tile_slots;  /* LDS knot slots (both envelopes); 0 = as emd_config */
  int wgs_per_cu;  /* persistent kernel: resident workgroups per CU (0 = occupancy max) */
  int persistent;  /* 1 = persistent atomic-queue kernel, 0 = one workgroup per trace */
} emd_launch;

const char* emd_last_error(void);
int emd_version(void);

/* Static LDS bytes of a sift workgroup and LDS bytes per knot slot (host-side sizing for
 * EMDKernelConfig.static_shared_mem_bytes / shared_mem_bytes, kernels.py:1191-1203,
 * 1273-1275).  No device needed. */
int emd_static_lds_bytes(int threads_per_block);
int emd_lds_slot_bytes(void);

/* Device facts used by EMDKernelConfig.for_device (replaces CuPy's
 * MaxSharedMemoryPerBlockOptin query, kernels.py:1256-1262). */
int emd_device_info(int* lds_bytes_per_workgroup, int* num_cus, int* warp_size);

/* Replaces EMDBatchLauncher.__init__ (kernels.py:1294-1341): no run-time compilation;
 * validates the configuration and selects the AOT gfx950 kernels. */
int emd_create(const emd_config* cfg, emd_handle** out);
int emd_destroy(emd_handle* h);

/* Replaces EMDBatchLauncher.run -> emd_batch_kernel (kernels.py:1343-1372, 834-861):
 * one workgroup per trace, grid = B.  signals [B,N] fp64 row-major, T = arange(N)
 * (PyEMD 'simple' extrema mode always uses the sample index).  Outputs (caller
 * allocated): imfs [B,max_imfs,N] fp64 (rows beyond n_imfs are zero), n_imfs [B],
 * overflow [B], debug_n/extno/nzm [B,max_imfs] (-1 where not reached), stats [B,4]
 * (sift count, peak extrema per side, knot sum, flags; may be NULL). */
int emd_run_padded(emd_handle* h, const double* signals, int B, int N, double* imfs,
                   int* n_imfs, int* overflow, int* debug_n, int* debug_extno,
                   int* debug_nzm, int* stats, void* stream);

/* Ragged traces in CSR form: trace s is values[offsets[s] .. offsets[s+1]) (int64
 * offsets, B+1 entries, max_n >= every length).  order (NULL = FIFO) gives the queue
 * order; outputs are indexed by trace, so the order never changes them.  imfs row k of
 * trace s starts at imfs + out_offsets[s] + k*len(s) (out_offsets NULL = max_imfs *
 * offsets[s]).  No reference counterpart (the reference needs one shared N); its
 * per-trace semantics are emd_batch_kernel's. */
int emd_run_ragged(emd_handle* h, const double* values, const int64_t* offsets, int B,
                   int max_n, const int* order, const emd_launch* launch, double* imfs,
                   const int64_t* out_offsets, int* n_imfs, int* overflow, int* debug_n,
                   int* debug_extno, int* debug_nzm, int* stats, void* stream);

/* Replaces EMDBatchLauncher.run_envelope_only -> envelope_batch_kernel
 * (kernels.py:1374-1407, 1100-1166): extrema + mirror + both splines, no sifting.
 * max_env/min_env [B,N]; n_extrema [B,2]; degenerate [B]; knot positions/values
 * [B,mbuf] with mbuf = max_extrema + 2*nbsym; n_mirrored [B,2].  Outputs of degenerate
 * traces are left untouched (callers pre-fill, as the reference's zeros do). */
int emd_envelope_only(emd_handle* h, const double* signals, int B, int N, double* max_env,
                      double* min_env, int* n_extrema, int* degenerate, double* max_kpos,
                      double* max_kval, double* min_kpos, double* min_kval, int* n_mirrored,
                      void* stream);

/* Largest number of workgroups of a launch shape resident per CU (occupancy query for
 * the autotuner).  No reference counterpart (CUDA occupancy reasoning lives in
 * EMDKernelConfig.for_device, kernels.py:1205-1271). */
int emd_max_resident(emd_handle* h, const emd_launch* launch, int max_n, int* wgs_per_cu);

/* Kernel timing: with timing enabled, every emd_run_padded / emd_run_ragged records a HIP
 * event on the caller's stream right before its first sift launch and right after its
 * last one; emd_kernel_time_ms waits for the second and returns the elapsed milliseconds.
 * No reference counterpart (the reference's benchmarks time with CUDA events around the
 * whole call, README.md:167-175). */
int emd_set_timing(emd_handle* h, int enable);
int emd_kernel_time_ms(emd_handle* h, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* EMD_GFX950_H */
