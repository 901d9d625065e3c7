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
  int threads;     /* 64, 128, 256, 512 or 1024 */
  int tile_slots;  /* LDS knot slots (both envelopes); 0 = as emd_config */
  int wgs_per_cu;  /* persistent kernel: resident workgroups per CU (0 = occupancy max) */
  int persistent;  /* 1 = persistent atomic-queue kernel, 0 = one workgroup per trace */
  int traces;      /* traces the CSR offsets describe (offsets has traces + 1 entries);
                      0 = the launch's B.  A queue entry outside [0, traces) is skipped:
                      nothing is read or written for it. */
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
 * offsets).  order (NULL = FIFO over traces 0..B-1) gives the queue: the launch
 * decomposes traces order[0..B), which may index a larger CSR batch, so one batch can be
 * split into several launches (e.g. one per length bucket) that share its arrays.
 * Outputs are indexed by trace, so the order never changes them.  max_n must be >= the
 * length of every queued trace; a longer one is reported with overflow = -1 and no rows.
 * imfs row k of trace s starts at imfs + out_offsets[s] + k*len(s) (out_offsets NULL =
 * max_imfs * offsets[s]); a trace's rows at and past n_imfs[s] are scratch and hold
 * unspecified values (only the padded layout of emd_run_padded is zero-filled, like the
 * reference's zero-initialised output).  No reference counterpart (the reference needs one
 * shared N); its per-trace semantics are emd_batch_kernel's. */
int emd_run_ragged(emd_handle* h, const double* values, const int64_t* offsets, int B,
                   int max_n, const int* order, const emd_launch* launch, double* imfs,
                   const int64_t* out_offsets, int* n_imfs, int* overflow, int* debug_n,
                   int* debug_extno, int* debug_nzm, int* stats, void* stream);

/* emd_run_padded with the reference's time axis: replaces EMDBatchLauncher.run(signals, T)
 * -> emd_batch_kernel for a T that is not arange(N) (kernels.py:1343-1372; T used at
 * kernels.py:471-472, 528-552, 703).  times holds the normalised axis, (T - T[0]) /
 * min(diff T) as GPUBatchEMD computes it (emd.py:83-88), strictly increasing: one shared
 * row [N] (times_per_signal = 0, the reference's layout) or one row per signal [B,N]
 * (times_per_signal = 1, the per-signal extension kernels.py:826-828 names).  Knots sit at
 * times[idx], mirrored knots at 2*times[sym] - times[idx], and the envelopes are evaluated
 * at times[i]; extrema are still found by sample index.  Same outputs as emd_run_padded. */
int emd_run_padded_times(emd_handle* h, const double* signals, const double* times,
                         int times_per_signal, int B, int N, double* imfs, int* n_imfs,
                         int* overflow, int* debug_n, int* debug_extno, int* debug_nzm,
                         int* stats, void* stream);

/* emd_run_ragged with a time axis per trace: times is laid out like values (trace s's
 * normalised times at times[offsets[s] .. offsets[s+1])).  No reference counterpart (the
 * reference takes one shared N); its per-trace semantics are those of
 * emd_run_padded_times. */
int emd_run_ragged_times(emd_handle* h, const double* values, const double* times,
                         const int64_t* offsets, int B, int max_n, const int* order,
                         const emd_launch* launch, double* imfs, const int64_t* out_offsets,
                         int* n_imfs, int* overflow, int* debug_n, int* debug_extno,
                         int* debug_nzm, int* stats, void* stream);

/* emd_run_ragged without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, outputs and queue semantics
 * as emd_run_ragged; the caller passes every input and output.  Instead of the handle's
 * scratch (grown by hipMalloc on demand), the queue counter and the workgroups' global
 * knot tiles live in one static device array of EMD_STATIC_SCRATCH_BYTES
 * per GPU, held until the process ends and looked up by emd_create.  A launch whose
 * workgroups would need more than the array holds runs with fewer resident workgroups
 * (or shorter plain sub-launches), which never changes results; one whose single
 * workgroup region exceeds it is rejected with EMD_ERR_VALUE.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (another stream, another handle, another
 * thread) race, so the caller orders them.  Nothing is allocated or freed. */
#define EMD_STATIC_SCRATCH_BYTES (128ll << 20)
int emd_run_ragged_noalloc(emd_handle* h, const double* values, const int64_t* offsets, int B,
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

/* emd_envelope_only with the reference's time axis (run_envelope_only(signals, T),
 * kernels.py:1374-1407 with T used as in emd_run_padded_times): times [N] shared
 * (times_per_signal = 0) or [B,N].  Knot positions are reported in time units. */
int emd_envelope_only_times(emd_handle* h, const double* signals, const double* times,
                            int times_per_signal, int B, int N, double* max_env,
                            double* min_env, int* n_extrema, int* degenerate, double* max_kpos,
                            double* max_kval, double* min_kpos, double* min_kval,
                            int* n_mirrored, void* stream);

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

/* ---- Output compaction -----------------------------------------------------------------
 * row_offsets[B+1]: exclusive scan of n_imfs[s] * len(s) (len from CSR offsets, or N when
 * offsets is NULL); row_offsets[B] is the total.  emd_compact also copies each trace's
 * n_imfs rows from imfs (layout as in emd_run_padded / emd_run_ragged, out_offsets NULL =
 * default) to compact + row_offsets[s].  No reference counterpart (the reference returns
 * the padded array, kernels.py:1353-1372). */
int emd_compact_offsets(const int* n_imfs, const int64_t* offsets, int N, int B,
                        int64_t* row_offsets, void* stream);
int emd_compact(const double* imfs, const int64_t* out_offsets, int max_imfs, const int* n_imfs,
                const int64_t* offsets, int N, int B, int64_t* row_offsets, double* compact,
                void* stream);

/* ---- Captured chunks --------------------------------------------------------------------
 * A chunk is one decomposition over fixed device buffers (padded when offsets is NULL,
 * else CSR), optionally followed by compaction (row_offsets and compact both non-NULL).
 * emd_chunk_graph_create captures it once into a HIP graph with its own queue counter and
 * global-tile scratch; each emd_chunk_graph_launch replays it on `stream` against the
 * buffers' current contents (refill them between replays, e.g. by an H2D copy on the
 * same stream).  Graphs of one handle may replay concurrently on different streams.
 * No reference counterpart (the reference launches eagerly, kernels.py:1364-1371). */
typedef struct emd_chunk {
  const double* values;
  const int64_t* offsets; /* NULL: padded [B,N] */
  int N;                  /* padded length (ignored with offsets) */
  int B;
  int max_n;
  const int* order;       /* queue order or NULL */
  emd_launch launch;
  double* imfs;
  const int64_t* out_offsets;
  int* n_imfs;
  int* overflow;
  int* debug_n;
  int* debug_extno;
  int* debug_nzm;
  int* stats;             /* may be NULL */
  int64_t* row_offsets;   /* compaction: both or neither */
  double* compact;
} emd_chunk;

typedef struct emd_graph emd_graph;
int emd_chunk_graph_create(emd_handle* h, const emd_chunk* chunk, emd_graph** out);
int emd_chunk_graph_launch(emd_graph* g, void* stream);
int emd_chunk_graph_destroy(emd_graph* g);

/* ---- Cost model ----------------------------------------------------------------------
 * A small FP32 MLP predicts, per trace, total sift iterations, IMF count and peak extrema
 * from cheap features; the host uses the predictions only to order the persistent queue
 * and size the LDS knot tile, so decompositions never depend on it.  None of these has a
 * reference counterpart (the reference launches one block per signal in batch order,
 * kernels.py:1214-1220, with LDS sized from max_extrema, kernels.py:1205-1271).
 * All pointers are device pointers; all calls are asynchronous on `stream`. */
#define EMD_N_FEATURES 12
#define EMD_MLP_PARAMS 5187 /* W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] */

/* feats[B][12] (fp32) of the CSR traces (values, int64 offsets[B+1]):
 * log N, log1p(strict extrema of the raw trace), log1p(sign changes), log(rms + 1e-12),
 * then log(fraction + 1e-10) of the mean 256-point |rfft|^2 (non-overlapping segments,
 * one zero-padded segment when N < 256) in bins [1,2) [2,4) [4,8) [8,16) [16,32) [32,64)
 * [64,96) [96,129). */
int emd_features(const double* values, const int64_t* offsets, int B, float* feats,
                 void* stream);

int emd_mlp_param_count(void);
/* floats of workspace emd_mlp_backward needs for M samples */
int64_t emd_mlp_workspace_floats(int M);

/* out[M][3] = MLP((feats - norm[0:12]) / norm[12:24]) in standardised target units. */
int emd_mlp_forward(const float* params, const float* norm, const float* feats, int M,
                    float* out, void* stream);

/* Fused forward + backward of L = (1/M) sum_s sum_o target_weight[o] (y_so - targets_so)^2:
 * grad[0:EMD_MLP_PARAMS] = dL/dparams, grad[EMD_MLP_PARAMS] = L.  grad holds
 * EMD_MLP_PARAMS + 1 floats; workspace holds emd_mlp_workspace_floats(M).  Sums over
 * workgroups run in a fixed order, so results are reproducible. */
int emd_mlp_backward(const float* params, const float* norm, const float* feats,
                     const float* targets, const float* target_weight, int M, float* workspace,
                     float* grad, void* stream);

/* One Adam step (bias-corrected; step counts from 1) on params with moments m, v. */
int emd_mlp_adam(float* params, float* m, float* v, const float* grad, int step, float lr,
                 float beta1, float beta2, float eps, void* stream);

/* ---- SEG-Y trace samples ----------------------------------------------------------------
 * Decodes big-endian SEG-Y samples in device memory to fp64 rows for emd_run_padded /
 * emd_run_ragged.  Formats (binary header bytes 3225-3226): 1 IBM float, 2 int32, 3 int16,
 * 5 IEEE float, 6 IEEE double, 8 int8; every conversion is exact.  Trace b's header starts
 * at bytes + trace_pos[b] (or b * trace_stride when trace_pos is NULL) and its samples
 * header_bytes later.  Output: out[b*N + i] (offsets NULL), or out[offsets[b] + i] for
 * offsets[b+1] - offsets[b] samples (CSR; needs trace_pos).  Unaligned traces are read
 * byte by byte.  emd_segy_sample_bytes gives a format's sample size (0 = unsupported).
 * No reference counterpart: the reference takes fp64 device arrays (emd.py:60-69) of
 * seismic traces sampled at 4 ms (emd.py:28-30). */
int emd_segy_sample_bytes(int format);
int emd_segy_decode(const void* bytes, const int64_t* trace_pos, int64_t trace_stride,
                    int header_bytes, const int64_t* offsets, int N, int B, int format,
                    double* out, void* stream);

/* ---- Hilbert spectral step ------------------------------------------------------------
 * Instantaneous amplitude, phase and frequency of every IMF row from its analytic signal
 * z = x + i H x, with H x as scipy.signal.hilbert computes it:
 *   n_fft == 0:  z = hilbert(row), a DFT of the row's own length (rows up to 8192
 *                samples; power-of-two lengths up to 16384);
 *   n_fft == -1: z = hilbert(row, N=L)[:len], L the smallest power of two >= len;
 *   n_fft  > 0:  the same with L = n_fft for every row (a power of two >= max_n, <= 16384).
 * amp[n] = |z[n]|, phase[n] = atan2(Im z[n], Re z[n]) (numpy.angle), and
 * freq[n] = (phi[n+1] - phi[n]) / (2 pi dt) with phi = numpy.unwrap(numpy.angle(z)) for
 * n < len-1 (the phase increment wrapped into [-pi, pi], +-pi kept), freq[len-1] =
 * freq[len-2], and 0 for a one-sample row; dt is the sample interval (freq in cycles per
 * unit of dt).
 * Rows: trace s has n_imfs[s] rows of len(s) samples (offsets[s+1] - offsets[s], or N when
 * offsets is NULL) back to back from element row_base[s] of imfs; row_base NULL is the
 * layout of emd_run_padded (s*max_imfs*N) or emd_run_ragged (offsets[s]*max_imfs), and
 * emd_compact's row_offsets is a valid row_base.  amp, phase and freq (each may be NULL)
 * share imfs' layout; no other element is written.  With offsets, every length lies in
 * [min_n, max_n] (rows outside it are skipped); one launch per transform length the range
 * spans, one in all when every row has the same length.  Asynchronous on `stream`;
 * allocates nothing.
 * No reference counterpart: the reference stops at the IMFs of traces it describes as
 * seismic, sampled at 4 ms (emd.py:28-30); this is the usual next step on them. */
int emd_hilbert(const double* imfs, const int64_t* row_base, int max_imfs, const int* n_imfs,
                const int64_t* offsets, int N, int B, int min_n, int max_n, int n_fft, double dt,
                double* amp, double* phase, double* freq, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EMD_GFX950_H */
