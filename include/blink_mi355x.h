/*
 * blink_mi355x.h -- C ABI of the MI355X (gfx950) visibility -> dirty-image
 * library (libblink_mi355x.so).
 *
 * The reference imager (PaCER-BLINK-Project/imager) has no C ABI: its GPU
 * path is a C++ class, CPacerImagerHip (src/gpu/pacer_imager_hip.h:12-58),
 * that overrides four virtuals of CPacerImager (src/pacer_imager.h:124-141)
 * and calls free functions in src/gpu/{corrections_gpu,gridding_gpu,
 * gpu_utils}.h plus hipFFT.  Every entry point below replaces one of those
 * functions; the comment on each names the reference interface it replaces.
 *
 * Conventions (all entry points):
 *   - Plain pointers and sizes only.  Every array pointer is DEVICE memory
 *     (hipMalloc / torch) unless the comment says "host".
 *   - complex64 data is interleaved float pairs (re, im), the layout of
 *     std::complex<float> and hipfftComplex.
 *   - Visibilities: vis[t][ch][bl][pol], pol = XX, XY, YX, YY,
 *     bl = a1*(a1+1)/2 + a2 with a2 <= a1, n_bl = n_ant*(n_ant+1)/2
 *     (autos included).  Grids, counts and images: [t][ch][v][u] (N x N each).
 *   - `stream` is a hipStream_t (NULL = the null stream).  Work is enqueued
 *     asynchronously on it; nothing synchronises unless stated.
 *   - Return value: BLINK_OK (0) or an error code; blink_last_error() gives
 *     the message.  BLINK_ERR_INVALID_ARGUMENT corresponds to the reference's
 *     std::invalid_argument, BLINK_ERR_RUNTIME to std::runtime_error.
 *   - A context owns its scratch memory, FFT twiddle tables and captured
 *     graphs.  Contexts are not thread-safe: use one per host thread.  Calls
 *     on one context share its scratch, so they must not run concurrently on
 *     different streams (order them, or use one context per stream).
 *     blink_image_batch also uses two streams of the context, forked from
 *     and joined back to `stream` within the call: the second stream for the
 *     visibility compaction of sparse sets on sides up to 2048 (beside the
 *     gridding plan) and, in a call with several grid sides, for the smaller
 *     sides; the pass stream for the second passes of overlapped chunks.
 */
#ifndef BLINK_MI355X_H
#define BLINK_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLINK_API __attribute__((visibility("default")))

typedef struct blink_ctx blink_ctx;

enum blink_status {
    BLINK_OK = 0,
    BLINK_ERR_INVALID_ARGUMENT = 1,
    BLINK_ERR_RUNTIME = 2,
    BLINK_ERR_HIP = 3,
    BLINK_ERR_UNSUPPORTED = 4
};

/* Polarisation to image, same values as enum class Polarization
 * (src/gridding.hpp:4). */
enum blink_pol { BLINK_POL_XX = 0, BLINK_POL_YY = 1, BLINK_POL_I = 2, BLINK_POL_V = 3 };

/* Weighting: natural, or the reference's "U" (src/pacer_imager.cpp:421-428). */
enum blink_weighting { BLINK_WEIGHTING_NATURAL = 0, BLINK_WEIGHTING_UNIFORM = 1 };

/* Convolution kernels of blink_gridding_kernel (the implemented GridMode
 * values, src/pacer_imager_defs.h:99-137, src/pacer_imager.cpp:617-623). */
enum blink_grid_kernel {
    BLINK_KERNEL_KAISER_BESSEL = 1,          /* Kaiser-Bessel (alpha 8.6) times a sinc */
    BLINK_KERNEL_KAISER_BESSEL_NO_SINC = 2   /* Kaiser-Bessel alone */
};
#define BLINK_MAX_KERNEL_SIZE 31

/* Phase corrections fused into blink_image_batch. */
enum blink_corrections { BLINK_CORR_GEOMETRIC = 1u, BLINK_CORR_CABLE = 2u };

/* FFT direction: BACKWARD is exp(+2 pi i jk/N) (FFTW_BACKWARD / hipFFT
 * BACKWARD, src/gpu/pacer_imager_hip.cpp:95), unnormalised. */
enum blink_fft_direction { BLINK_FFT_FORWARD = -1, BLINK_FFT_BACKWARD = 1 };

/* 2-D transform of blink_image_batch.  FFT: in-house FP32 FFT.  DFT_BF16X3 /
 * DFT_FP16X3: image = D G D^T as two dense GEMMs on MFMA over the Hermitian
 * half of the grid, with every FP32 operand split into hi + lo (bf16, or fp16
 * with a power-of-two scale per data row), products hi*hi + hi*lo + lo*hi and
 * FP32 accumulation (even N <= 512). */
enum blink_transform {
    BLINK_TRANSFORM_FFT = 0,
    BLINK_TRANSFORM_DFT_BF16X3 = 1,
    BLINK_TRANSFORM_DFT_FP16X3 = 2
};
#define BLINK_MAX_DFT_PIXELS 512

/* Gridding algorithm for blink_gridding. */
enum blink_grid_algo {
    BLINK_GRID_TILED = 0,  /* per-channel row plan + LDS-privatised row tiles (default) */
    BLINK_GRID_ATOMIC = 1  /* one thread per visibility, global atomics (plain path) */
};

/* Largest grid side supported by the in-house FFT and the gridding plan
 * (sides with prime factors 2, 3, 5, 7, 11 and 13; above 4096 the batched
 * path grids with natural weighting only). */
#define BLINK_MAX_PIXELS 8192

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device`.  No reference counterpart (the
 * reference keeps this state in CPacerImagerHip members,
 * src/gpu/pacer_imager_hip.h:14-29). */
BLINK_API int blink_create(int device, blink_ctx** out);
BLINK_API void blink_destroy(blink_ctx* ctx);
/* Message of the last failed call on this context (never NULL). */
BLINK_API const char* blink_last_error(const blink_ctx* ctx);
/* Library version string. */
BLINK_API const char* blink_version(void);
/* 1 when n_pixels is supported by blink_fft2d_c2c / blink_image. */
BLINK_API int blink_fft_supported(int n_pixels);

/* ------------------------------------------------- (a) phase corrections */

/* In place: V <- V * exp(+i 2 pi w[bl] nu[ch] / c) for all 4 pol products.
 * Replaces apply_geometric_corrections_gpu (src/gpu/corrections_gpu.cpp:90-101)
 * and apply_geometric_corrections_cpu (src/corrections.cpp:7-39, using w[bl]).
 * The argument is formed and reduced modulo 2 pi in FP64, sin/cos in FP32. */
BLINK_API int blink_apply_geometric_corrections(blink_ctx* ctx, float* vis, int n_ant,
                                                int n_ch, int n_t, const float* w_m,
                                                const double* freq_hz, void* stream);

/* In place: V <- V * exp(-i 2 pi (L[a2] - L[a1]) nu[ch] / c).
 * Replaces apply_cable_lengths_corrections_gpu (src/gpu/corrections_gpu.cpp:75-87)
 * and apply_cable_lengths_corrections_cpu (src/corrections.cpp:41-74). */
BLINK_API int blink_apply_cable_corrections(blink_ctx* ctx, float* vis, int n_ant, int n_ch,
                                            int n_t, const double* cable_m,
                                            const double* freq_hz, void* stream);

/* --------------------------------------------- (b) nearest-neighbour gridding */

/* Grids every unflagged, finite (|re|,|im| < 1e12) visibility of the chosen
 * polarisation at its cell and its conjugate at the mirrored cell, adding
 * into grids[t][ch][N][N] (complex64) and counts[t][ch][N][N] (float, per
 * image).  Accumulates: the caller zeroes grids/counts first (blink_image
 * zeroes them after imaging, as the reference does).  Cells follow
 * src/pacer_imager.cpp:289,332-369 bit for bit.  weighting = UNIFORM divides
 * each cell by its count afterwards (0/0 = NaN in empty cells, as the
 * reference).  Replaces gridding_gpu (src/gpu/gridding_gpu.h:12-18) and
 * CPacerImager::gridding (src/pacer_imager.cpp:250-432). */
BLINK_API int blink_gridding(blink_ctx* ctx, const float* vis, int n_ant, int n_ch, int n_t,
                             const float* u_m, const float* v_m, const int* bl_flags,
                             const double* freq_hz, double delta_u, double delta_v,
                             double min_uv, int n_pixels, int pol, int weighting, int algo,
                             float* grids, float* counts, void* stream);

/* Convolutional gridding: CPacerImager::gridding with m_nConvolvingKernelSize
 * = kernel_size > 0 (src/pacer_imager.cpp:378-381,398-401).  Each gridded
 * visibility and its conjugate are spread over kernel_size x kernel_size
 * cells with the 1023-times oversampled kernels of makeKernels /
 * makeKaiserBesselKernel (:566-651), placed as convolve_with_kernel and
 * calc_xy_grid2 place them (:654-750; footprint cells outside the grid are
 * dropped, not wrapped); counts stay at the nearest cell, as the reference
 * keeps them.  Same inputs, accumulation and weighting as blink_gridding.
 * kernel_size 1..BLINK_MAX_KERNEL_SIZE (even sizes follow the reference's
 * kernel table except its one write past the table end). */
BLINK_API int blink_gridding_kernel(blink_ctx* ctx, const float* vis, int n_ant, int n_ch,
                                    int n_t, const float* u_m, const float* v_m,
                                    const int* bl_flags, const double* freq_hz, double delta_u,
                                    double delta_v, double min_uv, int n_pixels, int pol,
                                    int weighting, int kernel_size, int kernel_mode,
                                    float* grids, float* counts, void* stream);

/* ------------------------------------------------------ array geometry */

/* Zenith UVW of every baseline bl = a1(a1+1)/2 + a2 (a2 <= a1):
 * u = -(x[a1] - x[a2]), v, w likewise, formed in double and stored as FP32
 * metres (CAntennaPositions::CalculateUVW, no-metafits branch,
 * src/antenna_positions.cpp:81-89).  x, y, z: double[n_ant] metres. */
BLINK_API int blink_zenith_uvw(blink_ctx* ctx, const double* x, const double* y,
                               const double* z, int n_ant, float* u_m, float* v_m, float* w_m,
                               void* stream);

/* Baseline flags from per-antenna flags (non-zero = flagged): a baseline is
 * flagged when either antenna is, or when it is an auto-correlation
 * (CPacerImager::UpdateFlags, src/pacer_imager.cpp:60-86). */
BLINK_API int blink_baseline_flags(blink_ctx* ctx, const int* ant_flags, int n_ant,
                                   int* bl_flags, void* stream);

/* uv_max[0] = max |u|, uv_max[1] = max |v| over unflagged baselines, as
 * doubles (-inf when every baseline is flagged), the inputs of
 * delta_u = osf * u_max / N (CPacerImager::CalculateUVW / grid(),
 * src/pacer_imager.cpp:213-221,497-499). */
BLINK_API int blink_uv_max(blink_ctx* ctx, const float* u_m, const float* v_m,
                           const int* bl_flags, int n_bl, double* uv_max, void* stream);

/* Cell indices of every (channel, baseline): cells[(ch*n_bl + bl)*4 + k] =
 * u, v, u_conj, v_conj, or -1 for flagged / min_uv-cut baselines.  The
 * index computation of calculate_pos (src/gpu/gridding_gpu.cpp:24-46);
 * exposed so that indices can be checked bit for bit. */
BLINK_API int blink_uv_cells(blink_ctx* ctx, const float* u_m, const float* v_m,
                             const int* bl_flags, int n_bl, const double* freq_hz, int n_ch,
                             double delta_u, double delta_v, double min_uv, int n_pixels,
                             int* cells, void* stream);

/* --------------------------------------------------------------- (c) FFT */

/* Batched, unnormalised 2-D complex FFT of `batch` N x N arrays, in-house
 * LDS-staged Stockham passes (no hipFFT/rocFFT).  in == out is allowed.
 * N = 2^a 3^b 5^c 7^d 11^e 13^f, 2 <= N <= BLINK_MAX_PIXELS.  Replaces gpufftPlanMany +
 * gpufftExecC2C (src/gpu/pacer_imager_hip.cpp:83-99). */
BLINK_API int blink_fft2d_c2c(blink_ctx* ctx, const float* in, float* out, int n_pixels,
                              int batch, int direction, void* stream);

/* ------------------------------------------------- (d) shift, normalisation */

/* sums[img] = sum of values[img*image_size ...] in double (exact for the
 * integer-valued counts).  Replaces vector_sum_gpu (src/gpu/gpu_utils.cpp:19-71). */
BLINK_API int blink_vector_sum(blink_ctx* ctx, const float* values, int64_t image_size,
                               int n_images, double* sums, void* stream);

/* In place on images[n_images][ny][nx]: fft_shift (src/pacer_imager.cpp:110-128,
 * odd sizes included) and multiply by (float)(1/sums[img % n_sums]); sums may
 * be NULL (no normalisation).  Replaces fft_shift_and_norm_gpu
 * (src/gpu/gpu_utils.cpp:128-140). */
BLINK_API int blink_fft_shift_and_norm(blink_ctx* ctx, float* images, int nx, int ny,
                                       int n_images, const double* sums, int n_sums,
                                       void* stream);

/* Full imaging step on gridded data: backward FFT of grids, sum of counts,
 * multiply by (float)(1/sum), fft_shift, written to images; then, when
 * zero_grids != 0, zeroes grids and counts for the next gridding round.
 * Replaces CPacerImagerHip::image (src/gpu/pacer_imager_hip.cpp:72-127) and
 * CPacerImager::image (src/pacer_imager.cpp:134-208), without the Images
 * metadata. */
BLINK_API int blink_image(blink_ctx* ctx, float* grids, float* counts, int n_pixels,
                          int n_images, float* images, int zero_grids, void* stream);

/* avg = complex mean of n_images images (image_averaging_gpu,
 * src/gpu/gpu_utils.cpp:143-161); real_only != 0 averages the real part
 * only and leaves the imaginary part 0 (image_averaging_cpu, src/utils.cpp:4-22). */
BLINK_API int blink_image_averaging(blink_ctx* ctx, const float* images, int64_t image_size,
                                    int n_images, int real_only, float* avg, void* stream);

/* ------------------------------------------------ batched visibilities -> images */

/* One set of images that share an array, a grid size and a polarisation:
 * n_t x n_ch images from vis[t][ch][bl][4].  Every pointer is device memory. */
typedef struct blink_image_set {
    const float* vis;        /* [n_t][n_ch][n_bl][4] complex64, read only, 16-byte aligned */
    int n_ant, n_ch, n_t;
    const float* u_m;        /* [n_bl] metres */
    const float* v_m;        /* [n_bl] metres */
    const float* w_m;        /* [n_bl] metres, read when BLINK_CORR_GEOMETRIC */
    const int* bl_flags;     /* [n_bl], non-zero = flagged */
    const double* freq_hz;   /* [n_ch] */
    const double* cable_m;   /* [n_ant], read when BLINK_CORR_CABLE */
    unsigned corrections;    /* BLINK_CORR_* bits, applied on the fly (vis is not modified) */
    double delta_u, delta_v, min_uv;
    int n_pixels;            /* grid side N; odd N with the FFT transform and nearest-cell
                                gridding only (`images` then required with an average) */
    int pol;                 /* blink_pol */
    float* images;           /* [n_t][n_ch][N][N] complex64 output, 16-byte aligned.  The
                                gridded plane is Hermitian, so the images are real and every
                                transform writes imaginary parts 0 (the reference's FFT leaves
                                FP32 rounding noise there).  May be NULL when `average` is
                                given (FFT transform): the images are then not formed */
    double* count_sums;      /* optional [n_t*n_ch] output: sum of counts per image */
    int transform;           /* blink_transform: how the 2-D transform is computed */
    int weighting;           /* blink_weighting: NATURAL, or UNIFORM, the reference's "U"
                                (src/pacer_imager.cpp:421-428): every cell divided by its
                                count before the transform, so an image with any empty cell
                                is NaN everywhere, as the reference's FFT makes it */
    float* average;          /* optional [N][N] complex64 output: the mean of the set's
                                n_t x n_ch images (image_averaging_gpu,
                                src/gpu/gpu_utils.cpp:143-161; the images being real, it is
                                also image_averaging_cpu's real-part mean, src/utils.cpp:4-22).
                                With the FFT transform it is formed from the row-pass buffers
                                without the images; with a DFT transform `images` is required */
    int kernel_size;         /* 0: nearest-cell gridding; 1 .. BLINK_MAX_KERNEL_SIZE:
                                convolutional gridding, each visibility and its conjugate
                                spread over kernel_size x kernel_size cells as
                                blink_gridding_kernel spreads them (FFT transform, natural
                                weighting) */
    int kernel_mode;         /* blink_grid_kernel when kernel_size > 0 */
} blink_image_set;

/* Visibilities -> normalised, shifted dirty images (and / or their mean) for
 * every set, with the phase corrections, gridding, weighting, FFT, shift and
 * normalisation fused into a few kernels per set, the whole call captured once as a hipGraph and
 * replayed when it is made again with the same set fields (compared field
 * by field).  Calls that differ only in the vis, images, count_sums or
 * average addresses keep a graph per address set, up to two (a caller that
 * alternates between two output buffers replays both without waiting);
 * beyond that a call refreshes the captured descriptors of the least recently
 * used of them and updates that instantiated graph (hipGraphExecUpdate)
 * instead of instantiating a new one, after waiting for its previous replay.
 * Results equal
 * apply_*_corrections + gridding + image of the reference (CPacerImager::run,
 * src/pacer_imager.cpp:440-526, with average_images). */
BLINK_API int blink_image_batch(blink_ctx* ctx, const blink_image_set* sets, int n_sets,
                                void* stream);

/* Bytes of the fixed device region of blink_image_batch_noalloc (256 MiB per
 * GPU, held from library load to process exit). */
#define BLINK_NOALLOC_SCRATCH_BYTES (256u << 20)

/* blink_image_batch for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same sets, results and count sums as blink_image_batch; it allocates
 * nothing: every scratch buffer of the call (plans, compacted visibilities,
 * row-pass buffer, averages, descriptors) is cut from one fixed region of
 * BLINK_NOALLOC_SCRATCH_BYTES of device memory that the library holds on
 * each GPU it runs on, and the row-pass chunks shrink to fit it. It launches
 * its kernels directly on `stream` (no graph, no autotuning: the side's
 * tuned or default tiles) and records no profiling events. Limits:
 * BLINK_ERR_INVALID_ARGUMENT when the call's plans do not fit the region
 * (split the call: 64 channels x 10 timesteps of the 128-tile array at
 * 1024^2 fit, 768 channels of a 256-antenna array do not),
 * BLINK_ERR_UNSUPPORTED for the DFT transforms, for convolutional gridding
 * (kernel_size > 0, whose tables would be allocated) and for grid sides
 * other than 128, 180, 256, 512, 1024, 2048 and 4096 (those whose FFT tables
 * blink_create builds). Every call on a GPU shares that GPU's region,
 * whichever context makes it, and finds it as the last call left it: calls
 * that overlap on one GPU (other streams, host threads or contexts) race.
 * Each call copies its descriptors from host memory and waits for the
 * previous call's copy, so it is called eagerly, not captured into a graph. */
BLINK_API int blink_image_batch_noalloc(blink_ctx* ctx, const blink_image_set* sets, int n_sets,
                                        void* stream);

/* Releases every captured graph and scratch buffer of the context. */
BLINK_API int blink_release_cache(blink_ctx* ctx);

/* ------------------------------------------------------- measurement hooks */

/* No reference counterpart (the reference times hipFFT with hipEvents inline,
 * src/gpu/pacer_imager_hip.cpp:79-101).  enable != 0: every kernel this
 * context launches is bracketed by a pair of HIP events recorded on the
 * stream it is launched on, and blink_image_batch runs its kernels directly
 * instead of replaying a graph.  enable == 0 stops recording (the totals are
 * kept until blink_profile_reset). */
BLINK_API int blink_profile_enable(blink_ctx* ctx, int enable);
/* Number of launches and total GPU milliseconds of the kernel `kernel`
 * (e.g. "grid_rows_fft", "fft_cols") since the last reset; waits for the
 * recorded events.  Unknown names give 0 launches. */
BLINK_API int blink_profile_get(blink_ctx* ctx, const char* kernel, int64_t* launches,
                                double* total_ms);
/* Comma-separated names of every kernel recorded since the last reset
 * (written into buf, at most buf_len bytes including the terminating 0). */
BLINK_API int blink_profile_kernels(blink_ctx* ctx, char* buf, int buf_len);
BLINK_API int blink_profile_reset(blink_ctx* ctx);

/* Tile sizes of the batched kernels for grid side n_pixels: grid rows per
 * first-pass tile (grid columns u up to 2048^2, where the first pass runs
 * along v) and output lines per second-pass block (image rows up to 2048^2,
 * image columns at 4096^2).  Positive when pinned or chosen by the
 * built-in autotuner (the first blink_image_batch call that launches a new
 * side with a new batch class -- images per launch of 1, 2-3, 4-15, 16-63,
 * ... -- times every built variant unless BLINK_AUTOTUNE=0 is set in the
 * environment; the last tuned pair is reported), negative for the defaults
 * not yet tuned, 0 when the side uses the generic kernels.  Groups on the
 * direct row pass (no plan: sets with one timestep per channel and more than
 * 64 baselines per grid row, at 128^2, 256^2 and generic sides up to 186)
 * have no row tile; at 128^2 and 256^2 their column block is tuned apart
 * from these pairs and not reported here, and a pinned pair pins only the
 * column block. */
BLINK_API int blink_batch_tiles(blink_ctx* ctx, int n_pixels, int* row_tile, int* col_tile);
/* Pins the tile sizes of the batched kernels for grid side n_pixels (row_tile
 * lines per first-pass tile, col_tile output lines per second-pass block, as
 * for blink_batch_tiles; both must be
 * built variants) and turns the autotuner off for that side; 0, 0 returns the
 * side to the autotuner.  No counterpart in the reference: for callers that
 * need reproducible launch shapes, and for tests of every variant. */
BLINK_API int blink_set_batch_tiles(blink_ctx* ctx, int n_pixels, int row_tile, int col_tile);

#ifdef __cplusplus
}
#endif
#endif /* BLINK_MI355X_H */
