/*
 * blink_oracle.h -- CPU oracle for the visibility -> dirty-image path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * imager's CPU algorithm, used as the checker by tests/, by the smoke test and
 * by the CPU-baseline leg of bench.py.  Nothing in the product library links,
 * loads or calls it.
 *
 * Two numeric variants are provided:
 *   *_f32 : the reference CPU path's own precision (FP64 index maths and phase
 *           maths, FP32 visibility/grid/FFT arithmetic, FP32 storage);
 *   *_f64 : the same algorithm carried out in double precision throughout
 *           (the "FP64 oracle" image every GPU result reports its error to).
 *
 * Visibility layout (astroio Visibilities, evidenced by
 * src/pacer_imager.cpp:303-324 and src/gpu/gridding_gpu.cpp:73-97):
 *   vis[t][ch][bl][pol] complex64, pol = XX,XY,YX,YY, bl = a1(a1+1)/2 + a2
 *   with a2 <= a1.
 * Grid / image layout: [t][ch][v][u] complex64 (src/pacer_imager.cpp:268).
 */
#ifndef BLINK_ORACLE_H
#define BLINK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Polarisation selector, same order as enum class Polarization
 * (src/gridding.hpp:4): XX=0, YY=1, I=2, V=3. */
enum { ORACLE_POL_XX = 0, ORACLE_POL_YY = 1, ORACLE_POL_I = 2, ORACLE_POL_V = 3 };

/* Number of baselines including autos: n_ant(n_ant+1)/2. */
int64_t oracle_n_baselines(int n_ant);

/* Baseline -> (a1, a2), src/pacer_imager.cpp:69-70. */
void oracle_baseline_antennas(int64_t bl, int* a1, int* a2);

/* CPacerImager::UpdateFlags, src/pacer_imager.cpp:60-86: a baseline is
 * flagged when either antenna is in the list, or when it is an auto. */
void oracle_baseline_flags(int n_ant, const int* flagged_ants, int n_flagged,
                           int* bl_flags);

/* u_max, v_max = max |u|, |v| over unflagged baselines
 * (CPacerImager::CalculateUVW, src/pacer_imager.cpp:213-221).  Returns -inf
 * when every baseline is flagged, like the reference. */
void oracle_uv_max(const float* u_m, const float* v_m, const int* bl_flags,
                   int64_t n_bl, double* u_max, double* v_max);

/* delta_u = osf*u_max/N, delta_v = osf*v_max/N, with osf a float as in
 * src/pacer_imager.cpp:65,497-499. */
void oracle_delta_uv(float oversampling_factor, double u_max, double v_max,
                     int n_pixels, double* delta_u, double* delta_v);

/* CPacerImager::get_frequency_hz, src/pacer_imager.cpp:528-555. */
double oracle_frequency_hz(double frequency_resolution_mhz, int n_averaged_channels,
                           double coarse_channel_bandwidth_mhz, int coarse_channel,
                           int fine_channel, int cotter_compatible);

/* Zenith UVW for antenna positions (x,y,z) metres: u = -(x1-x2) etc.,
 * CAntennaPositions::CalculateUVW, src/antenna_positions.cpp:52-90
 * (no-metafits branch).  Outputs FP32 metres per baseline. */
void oracle_zenith_uvw(int n_ant, const double* x, const double* y, const double* z,
                       float* u_m, float* v_m, float* w_m);

/* Geometric phase correction in place, FP64 maths, result stored as FP32:
 * V <- V * exp(+i 2 pi w nu / c) for all 4 pol products.
 * Follows src/corrections.cpp:7-39 with the baseline-indexed w[bl] of the GPU
 * kernel src/gpu/corrections_gpu.cpp:61 (the CPU file's w[a1*n_ant+a2] reads
 * past the end of an n_bl-long array). */
void oracle_geometric_corrections(float* vis, int n_ant, int n_ch, int n_t,
                                  const float* w_m, const double* freq_hz);

/* Cable-length phase correction in place, src/corrections.cpp:41-74:
 * V <- V * exp(i * (-2 pi (L[a2]-L[a1]) nu / c)). */
void oracle_cable_corrections(float* vis, int n_ant, int n_ch, int n_t,
                              const double* cable_m, const double* freq_hz);

/* Cell of one visibility: the reference's nearest-neighbour index formula
 * (src/pacer_imager.cpp:289,332-369; identical on the GPU at
 * src/gpu/gridding_gpu.cpp:24-46).  Returns 0 when the visibility is outside
 * the min_uv cut (not gridded), else 1 and the wrapped (u,v) indices of the
 * cell and of its conjugate. */
int oracle_cell(float u_m, float v_m, double freq_hz, double delta_u, double delta_v,
                double min_uv, int n_pixels, int* u_idx, int* v_idx,
                int* u_idx_conj, int* v_idx_conj);

/* Per-(channel, baseline) cells: cells[(ch*n_bl+bl)*4 + {0,1,2,3}] =
 * {u_idx, v_idx, u_idx_conj, v_idx_conj}, or all -1 when the baseline is
 * flagged or cut by min_uv. */
void oracle_cells(const float* u_m, const float* v_m, const int* bl_flags, int64_t n_bl,
                  const double* freq_hz, int n_ch, double delta_u, double delta_v,
                  double min_uv, int n_pixels, int* cells);

/* CPacerImager::gridding, src/pacer_imager.cpp:250-432 (nearest-neighbour,
 * no convolution kernel).  Accumulates into grids/counts (caller zeroes),
 * per-image counts (CPU semantics).  uniform != 0 applies the "U"
 * weighting division of src/pacer_imager.cpp:421-428.  Returns the number
 * of visibilities gridded (each adds two cells). */
int64_t oracle_gridding_f32(const float* vis, int n_ant, int n_ch, int n_t,
                            const float* u_m, const float* v_m, const int* bl_flags,
                            const double* freq_hz, double delta_u, double delta_v,
                            double min_uv, int n_pixels, int pol, int uniform,
                            float* grids, float* counts);
int64_t oracle_gridding_f64(const float* vis, int n_ant, int n_ch, int n_t,
                            const float* u_m, const float* v_m, const int* bl_flags,
                            const double* freq_hz, double delta_u, double delta_v,
                            double min_uv, int n_pixels, int pol, int uniform,
                            double* grids, double* counts);

/* Gridding kernels of the convolutional mode (src/pacer_imager.cpp:566-651):
 * table[r * K + c] = _griddingKernels[r][c], r < 1023 (oversampled
 * fractional cell position), c < K.  with_sinc selects KaiserBesselKernel
 * (1) or KaiserBesselWithoutSinc (0). */
void oracle_kb_kernels(int K, int with_sinc, double* table);

/* CPacerImager::gridding with m_nConvolvingKernelSize = K > 0
 * (src/pacer_imager.cpp:378-381,398-401, convolve_with_kernel :654-714,
 * calc_xy_grid2 :716-750).  Counts stay nearest-cell.  Same arguments as
 * oracle_gridding_*; returns the number of gridded visibilities. */
int64_t oracle_gridding_kb_f32(const float* vis, int n_ant, int n_ch, int n_t,
                               const float* u_m, const float* v_m, const int* bl_flags,
                               const double* freq_hz, double delta_u, double delta_v,
                               double min_uv, int n_pixels, int pol, int uniform, int K,
                               int with_sinc, float* grids, float* counts);
int64_t oracle_gridding_kb_f64(const float* vis, int n_ant, int n_ch, int n_t,
                               const float* u_m, const float* v_m, const int* bl_flags,
                               const double* freq_hz, double delta_u, double delta_v,
                               double min_uv, int n_pixels, int pol, int uniform, int K,
                               int with_sinc, double* grids, double* counts);

/* Batched 2-D complex FFT, unnormalised, out of place (in != out).
 * sign = +1 is FFTW_BACKWARD (exp(+2 pi i jk/N)), the direction the imager
 * uses (src/pacer_imager.cpp:150-152); sign = -1 is forward.  Mixed radix
 * for any N whose prime factors are 2, 3, 5 and 7.  Returns 0 on success,
 * -1 for an unsupported N. */
int oracle_fft2d_f32(const float* in, float* out, int n_pixels, int batch, int sign);
int oracle_fft2d_f64(const double* in, double* out, int n_pixels, int batch, int sign);

/* fft_shift, src/pacer_imager.cpp:30-35,110-128: swap columns x <->
 * (x + nx/2 + nx%2) % nx for x < nx/2, then rows likewise. */
void oracle_fft_shift_f32(float* image, size_t nx, size_t ny);
void oracle_fft_shift_f64(double* image, size_t nx, size_t ny);

/* CPacerImager::image, src/pacer_imager.cpp:134-208 (without the FITS
 * metadata): batched backward FFT, then per image multiply by
 * (float)(1/sum(counts)) and fft_shift.  n_threads <= 0 uses the OpenMP
 * default. */
void oracle_image_f32(const float* grids, const float* counts, int n_pixels,
                      int n_images, float* images, int n_threads);
void oracle_image_f64(const double* grids, const double* counts, int n_pixels,
                      int n_images, double* images, int n_threads);
/* The normalisation and shift of oracle_image_f32 alone, in place on images
 * the FFT has already transformed (src/pacer_imager.cpp:158-181). */
void oracle_norm_shift_f32(float* images, const float* counts, int n_pixels, int n_images,
                           int n_threads);

/* Reference CPU pipeline end to end (CPacerImager::run with the corrections
 * switched by the flags), FP32 semantics, OpenMP over images as in
 * src/pacer_imager.cpp:158,263.  vis is corrected in place when a
 * correction flag is set.  grids/counts are scratch of n_img*N*N. */
void oracle_run_f32(float* vis, int n_ant, int n_ch, int n_t,
                    const float* u_m, const float* v_m, const float* w_m,
                    const int* bl_flags, const double* freq_hz, const double* cable_m,
                    int apply_geom, int apply_cable, double delta_u, double delta_v,
                    double min_uv, int n_pixels, int pol,
                    float* grids, float* counts, float* images, int n_threads);

/* image_averaging_cpu, src/utils.cpp:4-22: mean of the REAL parts over all
 * images (imaginary part stays 0). */
void oracle_image_averaging_real(const float* images, size_t image_size, size_t n_images,
                                 float* avg);
/* image_averaging_gpu's kernel, src/gpu/gpu_utils.cpp:143-161: complex mean. */
void oracle_image_averaging_complex(const float* images, size_t image_size, size_t n_images,
                                    float* avg);

#ifdef __cplusplus
}
#endif
#endif
