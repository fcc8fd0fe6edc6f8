/*
 * gpunufft_hip.h — C ABI of the MI355X sector-NUFFT library (libgpunufft_hip.so).
 *
 * Plain pointers and sizes only; every entry point returns a status code
 * (GNUFFT_OK on success) and never exits the process. gnufft_last_error()
 * returns the calling thread's last error message.
 *
 * Data conventions (identical to the reference's):
 *   trajectory  float SoA [x0..x(M-1), y..., (z...)], coordinates in [-0.5, 0.5]
 *   k-space     complex64 (float re, im) [coil][sample], caller's sample order
 *   image/sens  complex64 [coil][z][y][x], x fastest (z absent in 2-D)
 *   grid        complex64 [coil][Gz][Gy][Gx]
 * "stream" arguments are hipStream_t values passed as void* (NULL = default
 * stream). Device pointers must be device memory of the current device.
 *
 * Each entry point below names the reference interface it replaces
 * (paths relative to the reference tree).
 */
#ifndef GPUNUFFT_HIP_H
#define GPUNUFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GNUFFT_ABI_VERSION 1

/* status codes */
enum {
  GNUFFT_OK = 0,
  GNUFFT_ERR_INVALID_ARGUMENT = 1, /* -> std::invalid_argument in the C++ layer */
  GNUFFT_ERR_OUT_OF_MEMORY = 2,    /* -> std::runtime_error (memory check) */
  GNUFFT_ERR_DEVICE = 3,           /* HIP runtime error (reference: HANDLE_ERROR exit) */
  GNUFFT_ERR_FFT = 4,              /* rocFFT error */
  GNUFFT_ERR_UNSUPPORTED = 5,
  GNUFFT_ERR_STATE = 6             /* call order violated (e.g. no trajectory yet) */
};

/* kernel lookup modes: CUDA/inc/gpuNUFFT_types.hpp:167-177 */
enum {
  GNUFFT_LOOKUP_NEAREST = 0, /* CONST_LOOKUP: nearest entry of the 1-D r^2 table */
  GNUFFT_LOOKUP_LINEAR = 1   /* TEXTURE2D_LOOKUP: emulated bilinear texture fetch */
};

/* stop stages: GpuNUFFTOutput, CUDA/inc/gpuNUFFT_types.hpp:184-193 */
enum {
  GNUFFT_STOP_CONVOLUTION = 0,
  GNUFFT_STOP_FFT = 1,
  GNUFFT_STOP_DEAPODIZATION = 2,
  GNUFFT_STOP_DENSITY_ESTIMATION = 3
};

/* memory location of caller buffers */
enum { GNUFFT_HOST = 0, GNUFFT_DEVICE = 1 };

typedef struct gnufft_plan_s* gnufft_plan;

/* Derived geometry of a plan: GpuNUFFTInfo, CUDA/inc/gpuNUFFT_types.hpp:230-311 */
typedef struct {
  int32_t dim;              /* 2 or 3 */
  uint32_t img[3];          /* image dims (z = 0 in 2-D) */
  uint32_t grid[3];         /* oversampled grid dims (z = 0 in 2-D) */
  uint32_t sectors[3];      /* sectors per dim (z = 1 in 2-D) */
  uint32_t sector_count;
  uint32_t sector_width;
  uint32_t kernel_width;
  uint32_t kernel_count;    /* entries of the lookup table the kernels index */
  uint32_t sector_pad_width;
  uint32_t sector_offset;
  uint32_t n_samples;       /* M, 0 before a trajectory is set */
  uint32_t n_sens_coils;    /* coils of the stored sensitivities, 0 = none */
  uint32_t n_chunks;        /* gridding work items (sector chunks) */
  int32_t lookup;
  int32_t has_dens;
  float osf;
  float radius_sqr;
  float dist_multiplier;
  uint64_t grid_count;      /* Gx*Gy*max(Gz,1) */
  uint64_t img_count;
} gnufft_plan_info;

/* ---- errors / version ---------------------------------------------------- */
const char* gnufft_last_error(void);
int gnufft_abi_version(void);

/* ---- device memory helpers for callers built without HIP headers ---------
 * replace the allocate/copy wrappers of CUDA/inc/cuda_utils.hpp:32-170 */
int gnufft_malloc(void** ptr, size_t bytes);
int gnufft_free(void* ptr);
int gnufft_memcpy(void* dst, const void* src, size_t bytes, int dst_loc, int src_loc,
                  void* stream);
int gnufft_zero(void* ptr, size_t bytes, void* stream); /* fill kernel, graph-capture safe */
int gnufft_stream_synchronize(void* stream);
int gnufft_mem_info(size_t* free_bytes, size_t* total_bytes); /* checkMemoryConsumption,
                                     CUDA/src/gpuNUFFT_operator_factory.cpp:628-663 */

/* ---- plan (one operator) ---------------------------------------------------
 * Replaces GpuNUFFTOperator construction + initKernel + initGpuNUFFTInfo
 * (CUDA/src/gpuNUFFT_operator.cpp:43-146) and the deapodization
 * precompute (CUDA/src/gpuNUFFT_operator_factory.cpp:359-448).
 * img_dims[2] == 0 selects 2-D. */
int gnufft_plan_create(gnufft_plan* plan, const uint32_t img_dims[3], uint32_t kernel_width,
                       uint32_t sector_width, float osf, int lookup, void* stream);
/* Frees the plan's device memory. Its rocFFT plans, if any, go to a
 * process-wide pool of at most 32 idle pairs (keyed by device, grid and
 * coil batch), from which a later plan of the same geometry takes them.
 * The batch destroy does the same. */
int gnufft_plan_destroy(gnufft_plan plan);
int gnufft_plan_get_info(gnufft_plan plan, gnufft_plan_info* info);

/* Oversampled-grid FFT implementation of a plan. The reference runs cuFFT
 * plus separate shift / scale / crop / pad / deapodization / coil kernels
 * (CUDA/src/gpuNUFFT_operator.cpp:682-850, :1143-1261).
 * GNUFFT_FFT_FUSED: one pass per axis, pruned to the image range, with those
 *   stages fused into the first and last pass (grid lengths up to 2048 whose
 *   prime factors are 2, 3, 5, 7, 11, 13, 17 or 19; the default for lengths
 *   up to 1023). A batch runs its plans on their own FFT kind.
 * GNUFFT_FFT_ROCFFT: rocFFT plus the fused pre / post kernels (any length).
 * set_fft returns GNUFFT_ERR_UNSUPPORTED for FUSED on other grid lengths. */
enum { GNUFFT_FFT_FUSED = 0, GNUFFT_FFT_ROCFFT = 1 };
int gnufft_plan_set_fft(gnufft_plan plan, int kind);
int gnufft_plan_get_fft(gnufft_plan plan, int* kind);

/* Launch shape of a plan's kernel-width-3 gridding kernels. The reference
 * launches fixed 256-thread blocks, one per sector (THREAD_BLOCK_SIZE,
 * CUDA/inc/cuda_utils.cuh:75-77; getOptimalGridDim, cuda_utils.hpp:155).
 * Here the spread and interp workgroups size their coil block (4, 2 or 1
 * coils of a padded sector in LDS) for n workgroups per CU, falling back
 * to fewer when the sector does not fit: n = 0 restores the default (3, or
 * the GNUFFT_LANE_PER_CU environment variable), 1..8 sets it for this
 * plan. Results do not depend on it; gpunufft_amd.autotune times it. */
int gnufft_plan_set_workgroups_per_cu(gnufft_plan plan, int n);

/* Sector precompute on the device: assignSectors + sort + sectorDataCount +
 * sorted trajectory/density (CUDA/src/gpuNUFFT_operator_factory.cpp:128-205,
 * :450-553; CUDA/src/gpu/precomp_kernels.cu:8-177). traj/dens live at
 * `location`; dens may be NULL. The kernels order samples by (sector,
 * index); gnufft_plan_get_tables reports the reference's std::sort order. */
int gnufft_plan_set_trajectory(gnufft_plan plan, const float* traj, uint32_t n_samples,
                               const float* dens, int location, void* stream);

/* Load precomputed tables (loadPrecomputedGpuNUFFTOperator,
 * CUDA/src/gpuNUFFT_operator_factory.cpp:576-626). traj_sorted is SoA in
 * sector order, samples in any order within a sector (the reference's
 * std::sort order or the stable one); dens_sorted and deapo may be NULL
 * (deapo NULL: computed). The tables are checked on the host first
 * (gnufft_check_precomputed); malformed tables return
 * GNUFFT_ERR_INVALID_ARGUMENT and leave the plan's previous tables freed. */
int gnufft_plan_load_precomputed(gnufft_plan plan, const float* traj_sorted,
                                 const uint32_t* data_indices, const uint32_t* sector_data_count,
                                 const uint32_t* sector_centers, const float* dens_sorted,
                                 const float* deapo, uint32_t n_samples, int location,
                                 void* stream);

/* Host-only consistency check of precomputed tables (host pointers; needs
 * no GPU). The reference loads such tables unchecked
 * (CUDA/src/gpuNUFFT_operator_factory.cpp:576-626); here
 * sector_data_count must have S+1 entries from 0, non-decreasing, ending at
 * n_samples; data_indices must be a permutation of [0, n_samples); the
 * centers must follow the standard layout (index * sw + sw / 2,
 * :207-270); and every sorted sample must lie in the sector whose range
 * holds it (CUDA/inc/precomp_utils.hpp:34-48). img_dims[2] == 0 selects 2-D. */
int gnufft_check_precomputed(const uint32_t img_dims[3], uint32_t sector_width, float osf,
                             const float* traj_sorted, const uint32_t* data_indices,
                             const uint32_t* sector_data_count, const uint32_t* sector_centers,
                             uint32_t n_samples);

/* Coil sensitivities [coil][img], kept on the device (setSens,
 * CUDA/inc/gpuNUFFT_operator.hpp:113-116). NULL / 0 coils clears them. */
int gnufft_plan_set_sens(gnufft_plan plan, const void* sens, uint32_t n_coils, int location,
                         void* stream);

/* Copy precompute tables to host buffers (any pointer may be NULL):
 * getDataIndices, getSectorDataCount, getSectorCenters, getKSpaceTraj,
 * getDens, deapodization function, getKernel
 * (CUDA/inc/gpuNUFFT_operator.hpp:135-201). For a plan whose trajectory was
 * set (not loaded), data_indices, traj_sorted and dens_sorted come in the
 * reference's own order: std::sort of (index, sector) pairs compared by
 * sector (CUDA/src/gpuNUFFT_operator_factory.cpp:56-72), built on the host
 * at this call; the kernels keep the device sort's stable order. A loaded
 * plan returns the tables it was given. */
int gnufft_plan_get_tables(gnufft_plan plan, uint32_t* data_indices,
                           uint32_t* sector_data_count, uint32_t* sector_centers,
                           float* traj_sorted, float* dens_sorted, float* deapo, float* lut);

/* Sector id of every sample in the caller's order (assignSectors output). */
int gnufft_plan_get_assigned_sectors(gnufft_plan plan, uint32_t* sectors_host);

/* ---- operators ------------------------------------------------------------
 * Adjoint: GpuNUFFTOperator::performGpuNUFFTAdj(GpuArray, GpuArray&, out),
 * CUDA/src/gpuNUFFT_operator.cpp:370-580 (stage semantics of :615-861).
 * kdata [n_coils][M] (caller order, device); out device:
 *   CONVOLUTION / DENSITY_ESTIMATION: [n_coils][grid]
 *   FFT: [n_coils][img] (cropped, scaled, before deapodization)
 *   DEAPODIZATION: [1][img] with sensitivities (coil-combined), else [n_coils][img] */
int gnufft_adjoint(gnufft_plan plan, const void* kdata, void* out, uint32_t n_coils, int stop,
                   void* stream);

/* Adjoint that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same arguments
 * and results as gnufft_adjoint; the plan (trajectory, sensitivities) must
 * be set up beforehand. The work grid and the spread sub-grids live in a
 * fixed area of gnufft_noalloc_capacity() bytes (64 MiB) per device, a
 * static device array that each device holds until the process ends; coils
 * run in batches that fit it. Every call on a device shares that device's
 * area, so calls that overlap on one device race (serialise them, e.g. on
 * one stream); each call writes every byte it reads, nothing needs
 * clearing. GNUFFT_ERR_UNSUPPORTED when one coil's
 * work memory exceeds the area, or when a stop that needs the FFT meets a
 * plan without the fused FFT (rocFFT plans allocate). */
int gnufft_adjoint_noalloc(gnufft_plan plan, const void* kdata, void* out, uint32_t n_coils,
                           int stop, void* stream);
size_t gnufft_noalloc_capacity(void);

/* Forward: performForwardGpuNUFFT(GpuArray, GpuArray&, out),
 * CUDA/src/gpuNUFFT_operator.cpp:895-1059 (stage semantics of :1092-1268).
 * img device [1][img] with sensitivities (expanded to n_coils = sens coils)
 * or [n_coils][img]; out device [n_coils][M] in the caller's order. */
int gnufft_forward(gnufft_plan plan, const void* img, void* out, uint32_t n_coils, int stop,
                   void* stream);

/* ---- per-stage timing --------------------------------------------------------
 * Replaces the reference's DEBUG stage timing (startTiming/stopTiming,
 * CUDA/src/gpuNUFFT_operator.cpp:1296-1311): when enabled, HIP events on the
 * launch stream bracket every kernel stage of the plan's calls. Reading
 * synchronises on the recorded events and returns totals since the last reset. */
enum {
  GNUFFT_T_SPREAD = 0,
  GNUFFT_T_INTERP = 1,
  GNUFFT_T_FFT = 2,
  GNUFFT_T_ZERO = 3,
  GNUFFT_T_PREPOST = 4,
  GNUFFT_T_COUNT = 5
};
typedef struct {
  double ms[GNUFFT_T_COUNT];
  uint32_t calls[GNUFFT_T_COUNT];
} gnufft_timing;
int gnufft_plan_set_timing(gnufft_plan plan, int enable);
int gnufft_plan_get_timing(gnufft_plan plan, gnufft_timing* out, int reset);

/* ---- gridding kernels alone (sorted sample order, no shift) ----------------
 * adjConvolution / forwardConvolution, CUDA/src/gpuNUFFT_operator.cpp:166-184:
 * spread: data_sorted [n_coils][M] -> grid [n_coils][grid] (grid is zeroed first)
 * interp: grid [n_coils][grid] -> data_sorted [n_coils][M] */
int gnufft_spread_sorted(gnufft_plan plan, const void* data_sorted, void* grid,
                         uint32_t n_coils, void* stream);
int gnufft_interp_sorted(gnufft_plan plan, const void* grid, void* data_sorted,
                         uint32_t n_coils, void* stream);

/* ---- batches of independent operators --------------------------------------
 * The reference runs operators one at a time, each stage a launch of its
 * own, with coils in chunks (CUDA/src/gpuNUFFT_operator.cpp:327-368, the
 * balanced sector order of CUDA/src/gpuNUFFT_operator_factory.cpp:74-126).
 * A batch runs the full adjoint or forward of many plans (any image sizes
 * and trajectory lengths; one dimension and kernel-width class) with one
 * launch per stage: spread, chunk reduction, fold, post, pre and interp
 * kernels take (plan, work item, coil block) tiles from one list balanced
 * over the XCDs; FFTs run once per distinct grid size, batched over all
 * coils of those plans.
 * n_coils[i]: coils of plan i (its sensitivity coil count when it has
 * sensitivities). The batch keeps the plans' device tables by reference:
 * plans must outlive it. A plan whose trajectory or sensitivities are set
 * again after batch_create makes batch_adjoint / batch_forward return
 * GNUFFT_ERR_STATE (create the batch again). A batch's calls may be issued
 * on different streams: each call's device work is ordered after the
 * previous call's (they share the batch's descriptors and work buffers). */
typedef struct gnufft_batch_s* gnufft_batch;
int gnufft_batch_create(gnufft_batch* batch, const gnufft_plan* plans, uint32_t n_plans,
                        const uint32_t* n_coils, void* stream);
int gnufft_batch_destroy(gnufft_batch batch);
/* kdata[i] / img[i] and out[i]: device buffers of plan i laid out as for
 * gnufft_adjoint / gnufft_forward with stop DEAPODIZATION. */
int gnufft_batch_adjoint(gnufft_batch batch, const void* const* kdata, void* const* out,
                         void* stream);
int gnufft_batch_forward(gnufft_batch batch, const void* const* img, void* const* out,
                         void* stream);

/* ---- device-resident CG-SENSE ----------------------------------------------
 * The reference's iterative reconstruction (matlab/demo/utils/cg_sense_2d.m:
 * right-hand side :46-47, system matrix A^H A + alpha I :51 and :99-102,
 * CG iterations :63-82, useMulticoil) drives its operator from the host with
 * a host round trip per product. Here the solver keeps x, r, p, A p and its
 * scalars on the device; one iteration is captured into a HIP graph (after a
 * warm-up that allocates the plan's work buffers) and replayed.
 * The plan must have sensitivities of n_coils coils. solve: kdata [n_coils][M]
 * (device, caller order) -> x_out [img] (device); asynchronous on `stream`.
 * One solver per stream at a time (it owns its vectors).
 * The captured graph holds the plan's device pointers. When the plan frees
 * or replaces any of them after the capture (set_trajectory,
 * load_precomputed, set_sens, set_fft, or work buffers grown by a call with
 * more coils), the next graph solve captures again; a plan whose sample
 * count or sensitivity coil count differs from cg_create's returns
 * GNUFFT_ERR_STATE. */
typedef struct gnufft_cg_s* gnufft_cg;
int gnufft_cg_create(gnufft_cg* cg, gnufft_plan plan, uint32_t n_coils, float alpha);
int gnufft_cg_destroy(gnufft_cg cg);
int gnufft_cg_solve(gnufft_cg cg, const void* kdata, void* x_out, uint32_t iterations,
                    int use_graph, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GPUNUFFT_HIP_H */
