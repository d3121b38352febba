/*
 * blink_oracle.c -- CPU oracle for the visibility -> dirty-image path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, the smoke test and
 * the CPU-baseline leg of bench.py.  The product library never links, loads
 * or calls this file.  Every function cites the reference file:line whose
 * behaviour it restates.  Build flags must keep IEEE semantics: no
 * -ffast-math, and -ffp-contract=off so that no multiply-add is fused (the
 * reference's cell indices depend on the exact double rounding sequence).
 *
 * Pinning (what checks each function, beyond "the same formula typed twice"):
 *   fft_shift          : the reference KATs tests/imager_test.cpp:21-40 and
 *                        tests/gpu_utils_test.cpp:60-87;
 *   fft2d              : naive DFT in long double for small N, numpy.fft
 *                        (pocketfft) for N up to 256, Parseval, linearity;
 *   cells / gridding   : brute-force nearest-cell search, exact-half rounding
 *                        cases, conjugate symmetry of the count grid,
 *                        sum(counts) = 2 * #gridded visibilities;
 *   image              : closed-form fringe of one visibility, DC pixel =
 *                        mean Re(V), Hermitian grid -> real image;
 *   corrections        : |V| preserved, identity for w = 0 / equal cable
 *                        lengths, geometric and cable rotations commute,
 *                        closed-form phase of a unit visibility;
 *   averaging          : the KAT tests/gpu_utils_test.cpp:91-124.
 */
#include "blink_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/pacer_imager_defs.h:12 (VEL_LIGHT) and :13 (SPEED_OF_LIGHT). */
#define ORACLE_VEL_LIGHT 299792458.0
/* src/pacer_imager_defs.h:21 */
#define ORACLE_MAX_VIS 1e12

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int64_t oracle_n_baselines(int n_ant) { return ((int64_t)n_ant * (n_ant + 1)) / 2; }

/* src/pacer_imager.cpp:69-70 */
void oracle_baseline_antennas(int64_t bl, int* a1, int* a2) {
    unsigned int ant1 = (unsigned int)(-0.5 + sqrt(0.25 + 2.0 * (double)bl));
    unsigned int ant2 = (unsigned int)(bl - ((int64_t)(ant1 + 1) * ant1) / 2);
    *a1 = (int)ant1;
    *a2 = (int)ant2;
}

/* src/pacer_imager.cpp:60-86 */
void oracle_baseline_flags(int n_ant, const int* flagged_ants, int n_flagged, int* bl_flags) {
    const int64_t n_bl = oracle_n_baselines(n_ant);
    for (int64_t bl = 0; bl < n_bl; bl++) {
        int a1, a2;
        oracle_baseline_antennas(bl, &a1, &a2);
        int listed = 0;
        for (int i = 0; i < n_flagged; i++) {
            if (flagged_ants[i] == a1 || flagged_ants[i] == a2) { listed = 1; break; }
        }
        bl_flags[bl] = (listed || a1 == a2) ? 1 : 0;
    }
}

/* src/pacer_imager.cpp:213-221 */
void oracle_uv_max(const float* u_m, const float* v_m, const int* bl_flags, int64_t n_bl,
                   double* u_max, double* v_max) {
    double um = -INFINITY, vm = -INFINITY;
    for (int64_t i = 0; i < n_bl; i++) {
        if (bl_flags[i] == 0) {
            if (um < fabs(u_m[i])) um = fabs(u_m[i]);
            if (vm < fabs(v_m[i])) vm = fabs(v_m[i]);
        }
    }
    *u_max = um;
    *v_max = vm;
}

/* src/pacer_imager.cpp:497-499: oversampling_factor is a float member
 * (src/pacer_imager.h:65) promoted to double in the product. */
void oracle_delta_uv(float oversampling_factor, double u_max, double v_max, int n_pixels,
                     double* delta_u, double* delta_v) {
    *delta_u = (double)oversampling_factor * (u_max) / n_pixels;
    *delta_v = (double)oversampling_factor * (v_max) / n_pixels;
}

/* src/pacer_imager.cpp:528-555 */
double oracle_frequency_hz(double frequency_resolution_mhz, int n_averaged_channels,
                           double coarse_channel_bandwidth_mhz, int coarse_channel,
                           int fine_channel, int cotter_compatible) {
    double fine_ch_bw = frequency_resolution_mhz * n_averaged_channels;
    double coarse_ch_bw = coarse_channel_bandwidth_mhz;
    double coarse_central = coarse_channel * coarse_ch_bw;
    double mhz;
    if (cotter_compatible)
        mhz = coarse_central - coarse_ch_bw / 2.00 + fine_ch_bw * fine_channel;
    else
        mhz = coarse_central - coarse_ch_bw / 2.00 + fine_ch_bw * fine_channel + fine_ch_bw / 2.00;
    return mhz * 1e6;
}

/* src/antenna_positions.cpp:52-90, zenith branch */
void oracle_zenith_uvw(int n_ant, const double* x, const double* y, const double* z,
                       float* u_m, float* v_m, float* w_m) {
    const int64_t n_bl = oracle_n_baselines(n_ant);
    for (int64_t bl = 0; bl < n_bl; bl++) {
        int i, j;
        oracle_baseline_antennas(bl, &i, &j);
        double u = x[i] - x[j];
        double v = y[i] - y[j];
        double w = z[i] - z[j];
        u_m[bl] = (float)(-u);
        v_m[bl] = (float)(-v);
        w_m[bl] = (float)(-w);
    }
}

/* Rotate the 4 pol products of one (t, ch, bl) record by angle (double
 * maths, FP32 store), src/corrections.cpp:21-34 / :57-69. */
static void rotate4(float* rec, double angle) {
    double s = sin(angle), c = cos(angle);
    for (int p = 0; p < 4; p++) {
        double re = rec[2 * p], im = rec[2 * p + 1];
        double re_prim = re * c - im * s;
        double im_prim = im * c + re * s;
        rec[2 * p] = (float)re_prim;
        rec[2 * p + 1] = (float)im_prim;
    }
}

/* src/corrections.cpp:7-39 with w[bl] (src/gpu/corrections_gpu.cpp:61) */
void oracle_geometric_corrections(float* vis, int n_ant, int n_ch, int n_t,
                                  const float* w_m, const double* freq_hz) {
    const int64_t n_bl = oracle_n_baselines(n_ant);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < n_t; t++) {
        for (int ch = 0; ch < n_ch; ch++) {
            for (int64_t bl = 0; bl < n_bl; bl++) {
                double w = w_m[bl];
                double angle = 2.0 * M_PI * w * freq_hz[ch] / ORACLE_VEL_LIGHT;
                rotate4(vis + (((int64_t)t * n_ch + ch) * n_bl + bl) * 8, angle);
            }
        }
    }
}

/* src/corrections.cpp:41-74 */
void oracle_cable_corrections(float* vis, int n_ant, int n_ch, int n_t,
                              const double* cable_m, const double* freq_hz) {
    const int64_t n_bl = oracle_n_baselines(n_ant);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < n_t; t++) {
        for (int ch = 0; ch < n_ch; ch++) {
            for (int64_t bl = 0; bl < n_bl; bl++) {
                int a1, a2;
                oracle_baseline_antennas(bl, &a1, &a2);
                double d = cable_m[a2] - cable_m[a1];
                double angle = -2.0 * M_PI * d * freq_hz[ch] / ORACLE_VEL_LIGHT;
                rotate4(vis + (((int64_t)t * n_ch + ch) * n_bl + bl) * 8, angle);
            }
        }
    }
}

/* wrap_index, src/pacer_imager.cpp:245-248.  The reference returns side+i
 * for every negative i, which is negative (an out-of-bounds write) for
 * i < -side; a Euclidean modulo agrees with it wherever it is in bounds and
 * aliases the cell otherwise. */
static int wrap_index(int i, int side) {
    if (i >= 0) return i % side;
    int r = i % side;
    return r == 0 ? 0 : side + r;
}

/* The reference rounds with round() (half away from zero) and truncates to
 * int; a quotient that does not fit in an int (infinite or NaN when
 * delta_u is 0) is undefined behaviour there and is treated as "not gridded"
 * here. */
static int pix_of(double q, int* pix) {
    double r = round(q);
    if (!(fabs(r) < 1073741824.0)) return 0;
    *pix = (int)r;
    return 1;
}

/* src/pacer_imager.cpp:289 (wavelength), :332-369 (u, v, uv distance,
 * rounding, wrap). */
int oracle_cell(float u_m, float v_m, double freq_hz, double delta_u, double delta_v,
                double min_uv, int n_pixels, int* u_idx, int* v_idx,
                int* u_idx_conj, int* v_idx_conj) {
    double wavelength_m = ORACLE_VEL_LIGHT / freq_hz;
    double u = u_m / wavelength_m;
    double v = v_m / wavelength_m;
    double uv_distance = sqrt(u * u + v * v);
    if (!(uv_distance > min_uv)) return 0;
    int u_pix, v_pix;
    if (!pix_of(u / delta_u, &u_pix) || !pix_of(v / delta_v, &v_pix)) return 0;
    *u_idx = wrap_index(u_pix, n_pixels);
    *v_idx = wrap_index(v_pix, n_pixels);
    *u_idx_conj = wrap_index(-u_pix, n_pixels);
    *v_idx_conj = wrap_index(-v_pix, n_pixels);
    return 1;
}

void oracle_cells(const float* u_m, const float* v_m, const int* bl_flags, int64_t n_bl,
                  const double* freq_hz, int n_ch, double delta_u, double delta_v,
                  double min_uv, int n_pixels, int* cells) {
    for (int ch = 0; ch < n_ch; ch++) {
        for (int64_t bl = 0; bl < n_bl; bl++) {
            int* c = cells + ((int64_t)ch * n_bl + bl) * 4;
            c[0] = c[1] = c[2] = c[3] = -1;
            if (bl_flags[bl]) continue;
            int ui, vi, uc, vc;
            if (oracle_cell(u_m[bl], v_m[bl], freq_hz[ch], delta_u, delta_v, min_uv, n_pixels,
                            &ui, &vi, &uc, &vc)) {
                c[0] = ui; c[1] = vi; c[2] = uc; c[3] = vc;
            }
        }
    }
}

/* Polarisation selection in FP32, src/pacer_imager.cpp:301-325.  The Stokes
 * V product (0 + 0.5i)*(xy - yx) is exact in FP32: (-0.5*b, 0.5*a). */
static void select_pol(const float* rec, int pol, float* re, float* im) {
    if (pol == ORACLE_POL_XX) {
        *re = rec[0]; *im = rec[1];
    } else if (pol == ORACLE_POL_YY) {
        *re = rec[6]; *im = rec[7];
    } else if (pol == ORACLE_POL_V) {
        float dre = rec[2] - rec[4];
        float dim = rec[3] - rec[5];
        *re = 0.0f * dre - 0.5f * dim;
        *im = 0.0f * dim + 0.5f * dre;
    } else {
        *re = (rec[0] + rec[6]) / 2.0f;
        *im = (rec[1] + rec[7]) / 2.0f;
    }
}

#define GRIDDING_BODY(ACC_T)                                                                   \
    const int64_t n_bl = oracle_n_baselines(n_ant);                                            \
    const int64_t grid_size = (int64_t)n_pixels * n_pixels;                                    \
    int64_t n_gridded = 0;                                                                     \
    _Pragma("omp parallel for collapse(2) schedule(static) reduction(+:n_gridded)")            \
    for (int t = 0; t < n_t; t++) {                                                            \
        for (int ch = 0; ch < n_ch; ch++) {                                                    \
            ACC_T* g = grids + ((int64_t)t * n_ch + ch) * grid_size * 2;                       \
            ACC_T* cnt = counts + ((int64_t)t * n_ch + ch) * grid_size;                        \
            for (int64_t bl = 0; bl < n_bl; bl++) {                                            \
                if (bl_flags[bl]) continue;                                                    \
                const float* rec = vis + (((int64_t)t * n_ch + ch) * n_bl + bl) * 8;           \
                float re, im;                                                                  \
                select_pol(rec, pol, &re, &im);                                                \
                if (isnan(re) || isnan(im)) continue;                                          \
                if (!(fabs(re) < ORACLE_MAX_VIS && fabs(im) < ORACLE_MAX_VIS)) continue;       \
                int ui, vi, uc, vc;                                                            \
                if (!oracle_cell(u_m[bl], v_m[bl], freq_hz[ch], delta_u, delta_v, min_uv,      \
                                 n_pixels, &ui, &vi, &uc, &vc))                                \
                    continue;                                                                  \
                int64_t p = (int64_t)vi * n_pixels + ui;                                       \
                g[2 * p] += (ACC_T)re;                                                         \
                g[2 * p + 1] += (ACC_T)im;                                                     \
                cnt[p] += 1;                                                                   \
                int64_t q = (int64_t)vc * n_pixels + uc;                                       \
                g[2 * q] += (ACC_T)re;                                                         \
                g[2 * q + 1] -= (ACC_T)im;                                                     \
                cnt[q] += 1;                                                                   \
                n_gridded++;                                                                   \
            }                                                                                  \
            if (uniform) {                                                                     \
                for (int64_t i = 0; i < grid_size; i++) {                                      \
                    g[2 * i] = g[2 * i] / cnt[i];                                              \
                    g[2 * i + 1] = g[2 * i + 1] / cnt[i];                                      \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
    }                                                                                          \
    return n_gridded;

/* src/pacer_imager.cpp:250-432 */
int64_t oracle_gridding_f32(const float* vis, int n_ant, int n_ch, int n_t,
                            const float* u_m, const float* v_m, const int* bl_flags,
                            const double* freq_hz, double delta_u, double delta_v,
                            double min_uv, int n_pixels, int pol, int uniform,
                            float* grids, float* counts) {
    GRIDDING_BODY(float)
}

int64_t oracle_gridding_f64(const float* vis, int n_ant, int n_ch, int n_t,
                            const float* u_m, const float* v_m, const int* bl_flags,
                            const double* freq_hz, double delta_u, double delta_v,
                            double min_uv, int n_pixels, int pol, int uniform,
                            double* grids, double* counts) {
    GRIDDING_BODY(double)
}

/* ------------------------------------------------------------------------ */
/* Convolutional (Kaiser-Bessel) gridding.                                   */
/* ------------------------------------------------------------------------ */

/* Power series of I0, stopped when a term is within `precision` of the sum
 * (bessel0, src/pacer_imager.cpp:566-578). */
static double kb_bessel0(double x, double precision) {
    double d = 0.0, ds = 1.0, sum = 1.0;
    do {
        d += 2.0;
        ds *= x * x / (d * d);
        sum += ds;
    } while (ds > sum * precision);
    return sum;
}

/* makeKaiserBesselKernel (:580-599) on K * 1023 samples, then makeKernels'
 * re-ordering into 1023 rows of K (:601-651).  For even K the reference
 * writes kernel[n] (one past the end); that value is dropped here. */
void oracle_kb_kernels(int K, int with_sinc, double* table) {
    const int os = 1023;
    const size_t n = (size_t)K * os, mid = n / 2;
    double* k1 = (double*)calloc(n, sizeof(double));
    double* sinc = (double*)malloc(sizeof(double) * (mid + 1));
    const double filter_ratio = 1.0 / (double)os;
    sinc[0] = filter_ratio;
    for (size_t i = 1; i != mid + 1; i++) {
        const double x = (double)i;
        sinc[i] = with_sinc ? (sin(M_PI * filter_ratio * x) / (M_PI * x)) : filter_ratio;
    }
    const double norm = (double)os / kb_bessel0(8.6, 1e-8);
    for (size_t i = 0; i != mid + 1; i++) {
        const double term = (double)i / (double)mid;
        if (mid + i < n) k1[mid + i] = sinc[i] * kb_bessel0(8.6 * sqrt(1.0 - term * term), 1e-8) * norm;
    }
    for (size_t i = 0; i != mid; i++) k1[i] = k1[n - 1 - i];
    for (int i = 0; i < os; i++)
        for (int x = 0; x < K; x++) table[(size_t)(os - i - 1) * K + x] = k1[(size_t)x * os + i];
    free(sinc);
    free(k1);
}

/* calc_xy_grid2 with is_odd = 0: grid column of footprint coordinate p for
 * im_sign s; returns 0 when it falls outside the grid. */
static int kb_xy(int p, int s, int n_pixels, int* x) {
    const int center = n_pixels / 2;
    const int idx = s * p + n_pixels / 2;
    *x = idx < center ? center + idx : idx - center;
    return *x >= 0 && *x < n_pixels;
}

#define KB_CONVOLVE(ACC_T)                                                                     \
    {                                                                                          \
        const double u_exact = u / delta_u, v_exact = v / delta_v;                             \
        const double u_pix = round(u / delta_u), v_pix = round(v / delta_v);                   \
        int xk = (int)round((u_exact - u_pix) * 1023);                                         \
        int yk = (int)round((v_exact - v_pix) * 1023);                                         \
        xk = (xk + (1023 * 3) / 2) % 1023;                                                     \
        yk = (yk + (1023 * 3) / 2) % 1023;                                                     \
        const double* xker = table + (size_t)xk * K;                                           \
        const double* yker = table + (size_t)yk * K;                                           \
        const int mid = K / 2;                                                                 \
        int vv = (int)(v_pix - mid);                                                           \
        for (int j = 0; j != K; ++j, ++vv) {                                                   \
            const double ykv = yker[j];                                                        \
            int uu = (int)(u_pix - mid);                                                       \
            for (int i = 0; i != K; ++i, ++uu) {                                               \
                const double kv = ykv * xker[i];                                               \
                const double re_conv = re * kv;                                                \
                const double im_conv = im_sign * im * kv;                                      \
                int xg, yg;                                                                    \
                if (!kb_xy(uu, (int)im_sign, n_pixels, &xg) ||                                 \
                    !kb_xy(vv, (int)im_sign, n_pixels, &yg))                                   \
                    continue;                                                                  \
                const int64_t c = (int64_t)yg * n_pixels + xg;                                 \
                g[2 * c] = (ACC_T)((double)g[2 * c] + re_conv);                                \
                g[2 * c + 1] = (ACC_T)((double)g[2 * c + 1] + im_conv);                        \
            }                                                                                  \
        }                                                                                      \
    }

#define GRIDDING_KB_BODY(ACC_T)                                                                \
    const int64_t n_bl = oracle_n_baselines(n_ant);                                            \
    const int64_t grid_size = (int64_t)n_pixels * n_pixels;                                    \
    double* table = (double*)malloc(sizeof(double) * 1023 * (size_t)K);                        \
    oracle_kb_kernels(K, with_sinc, table);                                                    \
    int64_t n_gridded = 0;                                                                     \
    _Pragma("omp parallel for collapse(2) schedule(static) reduction(+:n_gridded)")            \
    for (int t = 0; t < n_t; t++) {                                                            \
        for (int ch = 0; ch < n_ch; ch++) {                                                    \
            ACC_T* g = grids + ((int64_t)t * n_ch + ch) * grid_size * 2;                       \
            ACC_T* cnt = counts + ((int64_t)t * n_ch + ch) * grid_size;                        \
            const double wavelength = ORACLE_VEL_LIGHT / freq_hz[ch];                          \
            for (int64_t bl = 0; bl < n_bl; bl++) {                                            \
                if (bl_flags[bl]) continue;                                                    \
                const float* rec = vis + (((int64_t)t * n_ch + ch) * n_bl + bl) * 8;           \
                float ref, imf;                                                                \
                select_pol(rec, pol, &ref, &imf);                                              \
                if (isnan(ref) || isnan(imf)) continue;                                        \
                if (!(fabs(ref) < ORACLE_MAX_VIS && fabs(imf) < ORACLE_MAX_VIS)) continue;     \
                int ui, vi, uc, vc;                                                            \
                if (!oracle_cell(u_m[bl], v_m[bl], freq_hz[ch], delta_u, delta_v, min_uv,      \
                                 n_pixels, &ui, &vi, &uc, &vc))                                \
                    continue;                                                                  \
                const double u = (double)u_m[bl] / wavelength;                                 \
                const double v = (double)v_m[bl] / wavelength;                                 \
                const float re = ref, im = imf;                                                \
                {                                                                              \
                    const double im_sign = 1.00;                                               \
                    KB_CONVOLVE(ACC_T)                                                         \
                }                                                                              \
                cnt[(int64_t)vi * n_pixels + ui] += 1;                                         \
                {                                                                              \
                    const double im_sign = -1.00;                                              \
                    KB_CONVOLVE(ACC_T)                                                         \
                }                                                                              \
                cnt[(int64_t)vc * n_pixels + uc] += 1;                                         \
                n_gridded++;                                                                   \
            }                                                                                  \
            if (uniform) {                                                                     \
                for (int64_t i = 0; i < grid_size; i++) {                                      \
                    g[2 * i] = g[2 * i] / cnt[i];                                              \
                    g[2 * i + 1] = g[2 * i + 1] / cnt[i];                                      \
                }                                                                              \
            }                                                                                  \
        }                                                                                      \
    }                                                                                          \
    free(table);                                                                               \
    return n_gridded;

int64_t oracle_gridding_kb_f32(const float* vis, int n_ant, int n_ch, int n_t,
                               const float* u_m, const float* v_m, const int* bl_flags,
                               const double* freq_hz, double delta_u, double delta_v,
                               double min_uv, int n_pixels, int pol, int uniform, int K,
                               int with_sinc, float* grids, float* counts) {
    GRIDDING_KB_BODY(float)
}

int64_t oracle_gridding_kb_f64(const float* vis, int n_ant, int n_ch, int n_t,
                               const float* u_m, const float* v_m, const int* bl_flags,
                               const double* freq_hz, double delta_u, double delta_v,
                               double min_uv, int n_pixels, int pol, int uniform, int K,
                               int with_sinc, double* grids, double* counts) {
    GRIDDING_KB_BODY(double)
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time (Cooley-Tukey), out of     */
/* place, with the sub-transform combination done as a direct DFT of the    */
/* radix.  Stands in for FFTW's fftwf_plan_many_dft(..., FFTW_BACKWARD)     */
/* (src/pacer_imager.cpp:150-152), which is not installed here.             */
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    static const int primes[] = {2, 3, 5, 7, 11, 13};
    for (int i = 0; i < 6; i++)
        if (n % primes[i] == 0) return primes[i];
    return 0;
}

static int fft_supported(int n) {
    if (n < 1) return 0;
    while (n > 1) {
        int p = smallest_factor(n);
        if (!p) return 0;
        n /= p;
    }
    return 1;
}

#define FFT_IMPL(SUFFIX, REAL)                                                                 \
    /* tw[j] = exp(sign*2*pi*i*j/N), j < N, computed in double. */                             \
    static REAL* make_twiddles_##SUFFIX(int N, int sign) {                                     \
        REAL* tw = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)N);                                \
        for (int j = 0; j < N; j++) {                                                          \
            double a = 2.0 * M_PI * (double)j / (double)N;                                     \
            tw[2 * j] = (REAL)cos(a);                                                          \
            tw[2 * j + 1] = (REAL)(sign * sin(a));                                             \
        }                                                                                      \
        return tw;                                                                             \
    }                                                                                          \
    /* out[k] = sum_j in[j*is] * w_n^(jk), w_n = tw[N/n] */                                    \
    static void fft1d_##SUFFIX(const REAL* in, size_t is, REAL* out, int n, const REAL* tw,    \
                               int N) {                                                        \
        if (n == 1) {                                                                          \
            out[0] = in[0];                                                                    \
            out[1] = in[1];                                                                    \
            return;                                                                            \
        }                                                                                      \
        int p = smallest_factor(n);                                                            \
        int m = n / p;                                                                         \
        for (int r = 0; r < p; r++) fft1d_##SUFFIX(in + 2 * r * is, is * p, out + 2 * r * m, m, tw, N); \
        const int step = N / n; /* w_n^j = tw[j*step] */                                       \
        REAL yr[16], yi[16];                                                                     \
        for (int k = 0; k < m; k++) {                                                          \
            for (int r = 0; r < p; r++) {                                                      \
                REAL a = out[2 * (r * m + k)], b = out[2 * (r * m + k) + 1];                   \
                size_t ti = ((size_t)r * k * step) % (size_t)N;                                \
                REAL c = tw[2 * ti], s = tw[2 * ti + 1];                                       \
                yr[r] = a * c - b * s;                                                         \
                yi[r] = a * s + b * c;                                                         \
            }                                                                                  \
            for (int q = 0; q < p; q++) {                                                      \
                REAL sr = 0, si = 0;                                                           \
                for (int r = 0; r < p; r++) {                                                  \
                    size_t ti = ((size_t)r * q * m * step) % (size_t)N;                        \
                    REAL c = tw[2 * ti], s = tw[2 * ti + 1];                                   \
                    sr += yr[r] * c - yi[r] * s;                                               \
                    si += yr[r] * s + yi[r] * c;                                               \
                }                                                                              \
                out[2 * (k + q * m)] = sr;                                                     \
                out[2 * (k + q * m) + 1] = si;                                                 \
            }                                                                                  \
        }                                                                                      \
    }                                                                                          \
    int oracle_fft2d_##SUFFIX(const REAL* in, REAL* out, int N, int batch, int sign) {         \
        if (!fft_supported(N) || batch < 0) return -1;                                         \
        REAL* tw = make_twiddles_##SUFFIX(N, sign >= 0 ? 1 : -1);                              \
        const int64_t gs = (int64_t)N * N;                                                     \
        /* rows */                                                                             \
        _Pragma("omp parallel for schedule(static)")                                           \
        for (int64_t ry = 0; ry < (int64_t)batch * N; ry++) {                                  \
            int64_t img = ry / N, y = ry % N;                                                  \
            fft1d_##SUFFIX(in + 2 * (img * gs + y * N), 1, out + 2 * (img * gs + y * N), N, tw, N); \
        }                                                                                      \
        /* columns */                                                                          \
        _Pragma("omp parallel")                                                                \
        {                                                                                      \
            REAL* col = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)N);                           \
            REAL* res = col + 2 * (size_t)N;                                                   \
            _Pragma("omp for schedule(static)")                                                \
            for (int64_t cx = 0; cx < (int64_t)batch * N; cx++) {                              \
                int64_t img = cx / N, x = cx % N;                                              \
                REAL* base = out + 2 * (img * gs + x);                                         \
                for (int y = 0; y < N; y++) {                                                  \
                    col[2 * y] = base[2 * (int64_t)y * N];                                     \
                    col[2 * y + 1] = base[2 * (int64_t)y * N + 1];                             \
                }                                                                              \
                fft1d_##SUFFIX(col, 1, res, N, tw, N);                                         \
                for (int y = 0; y < N; y++) {                                                  \
                    base[2 * (int64_t)y * N] = res[2 * y];                                     \
                    base[2 * (int64_t)y * N + 1] = res[2 * y + 1];                             \
                }                                                                              \
            }                                                                                  \
            free(col);                                                                         \
        }                                                                                      \
        free(tw);                                                                              \
        return 0;                                                                              \
    }

FFT_IMPL(f32, float)
FFT_IMPL(f64, double)

/* calc_fft_shift, src/pacer_imager.cpp:30-35 */
static size_t calc_fft_shift(size_t pos, size_t side) {
    size_t is_odd = side % 2;
    return (pos + side / 2 + is_odd) % side;
}

#define SHIFT_IMPL(SUFFIX, REAL)                                                               \
    void oracle_fft_shift_##SUFFIX(REAL* image, size_t nx, size_t ny) {                        \
        for (size_t y = 0; y < ny; y++) {                                                      \
            for (size_t x = 0; x < nx / 2; x++) {                                              \
                size_t src = y * nx + x;                                                       \
                size_t dst = y * nx + calc_fft_shift(x, nx);                                   \
                REAL tr = image[2 * src], ti = image[2 * src + 1];                             \
                image[2 * src] = image[2 * dst];                                               \
                image[2 * src + 1] = image[2 * dst + 1];                                       \
                image[2 * dst] = tr;                                                           \
                image[2 * dst + 1] = ti;                                                       \
            }                                                                                  \
        }                                                                                      \
        for (size_t y = 0; y < ny / 2; y++) {                                                  \
            for (size_t x = 0; x < nx; x++) {                                                  \
                size_t src = y * nx + x;                                                       \
                size_t dst = calc_fft_shift(y, ny) * nx + x;                                   \
                REAL tr = image[2 * src], ti = image[2 * src + 1];                             \
                image[2 * src] = image[2 * dst];                                               \
                image[2 * src + 1] = image[2 * dst + 1];                                       \
                image[2 * dst] = tr;                                                           \
                image[2 * dst + 1] = ti;                                                       \
            }                                                                                  \
        }                                                                                      \
    }

SHIFT_IMPL(f32, float)
SHIFT_IMPL(f64, double)

static void set_threads(int n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
}

/* src/pacer_imager.cpp:134-181: FFT, sum(counts) in double, multiply by
 * (float)(1/sum) (std::complex<float>::operator*= takes a float), shift. */
void oracle_image_f32(const float* grids, const float* counts, int n_pixels, int n_images,
                      float* images, int n_threads) {
    set_threads(n_threads);
    oracle_fft2d_f32(grids, images, n_pixels, n_images, +1);
    oracle_norm_shift_f32(images, counts, n_pixels, n_images, n_threads);
}

/* The part of src/pacer_imager.cpp:158-181 after the FFT, on transformed
 * images in place: sum(counts) in double, multiply by (float)(1/sum), shift. */
void oracle_norm_shift_f32(float* images, const float* counts, int n_pixels, int n_images,
                           int n_threads) {
    set_threads(n_threads);
    const int64_t gs = (int64_t)n_pixels * n_pixels;
    #pragma omp parallel for schedule(static)
    for (int img = 0; img < n_images; img++) {
        const float* cnt = counts + img * gs;
        float* im = images + 2 * img * gs;
        double counter_sum = 0.0;
        for (int64_t i = 0; i < gs; i++) counter_sum += cnt[i];
        double fnorm = 1.00 / counter_sum;
        float f = (float)fnorm;
        for (int64_t i = 0; i < gs; i++) {
            im[2 * i] *= f;
            im[2 * i + 1] *= f;
        }
        oracle_fft_shift_f32(im, (size_t)n_pixels, (size_t)n_pixels);
    }
}

void oracle_image_f64(const double* grids, const double* counts, int n_pixels, int n_images,
                      double* images, int n_threads) {
    set_threads(n_threads);
    oracle_fft2d_f64(grids, images, n_pixels, n_images, +1);
    const int64_t gs = (int64_t)n_pixels * n_pixels;
    #pragma omp parallel for schedule(static)
    for (int img = 0; img < n_images; img++) {
        const double* cnt = counts + img * gs;
        double* im = images + 2 * img * gs;
        double counter_sum = 0.0;
        for (int64_t i = 0; i < gs; i++) counter_sum += cnt[i];
        double fnorm = 1.00 / counter_sum;
        for (int64_t i = 0; i < gs; i++) {
            im[2 * i] *= fnorm;
            im[2 * i + 1] *= fnorm;
        }
        oracle_fft_shift_f64(im, (size_t)n_pixels, (size_t)n_pixels);
    }
}

/* src/pacer_imager.cpp:440-526 (grid + image), corrections at :463-473 */
void oracle_run_f32(float* vis, int n_ant, int n_ch, int n_t,
                    const float* u_m, const float* v_m, const float* w_m,
                    const int* bl_flags, const double* freq_hz, const double* cable_m,
                    int apply_geom, int apply_cable, double delta_u, double delta_v,
                    double min_uv, int n_pixels, int pol,
                    float* grids, float* counts, float* images, int n_threads) {
    set_threads(n_threads);
    if (apply_geom) oracle_geometric_corrections(vis, n_ant, n_ch, n_t, w_m, freq_hz);
    if (apply_cable) oracle_cable_corrections(vis, n_ant, n_ch, n_t, cable_m, freq_hz);
    const size_t n = (size_t)n_ch * n_t * n_pixels * n_pixels;
    memset(grids, 0, n * 2 * sizeof(float));
    memset(counts, 0, n * sizeof(float));
    oracle_gridding_f32(vis, n_ant, n_ch, n_t, u_m, v_m, bl_flags, freq_hz, delta_u, delta_v,
                        min_uv, n_pixels, pol, 0, grids, counts);
    oracle_image_f32(grids, counts, n_pixels, n_ch * n_t, images, n_threads);
}

/* src/utils.cpp:4-22 */
void oracle_image_averaging_real(const float* images, size_t image_size, size_t n_images,
                                 float* avg) {
    memset(avg, 0, sizeof(float) * 2 * image_size);
    for (size_t img = 0; img < n_images; img++) {
        const float* d = images + 2 * img * image_size;
        for (size_t i = 0; i < image_size; i++) avg[2 * i] = avg[2 * i] + d[2 * i];
    }
    for (size_t i = 0; i < image_size; i++) avg[2 * i] = avg[2 * i] / (float)n_images;
}

/* src/gpu/gpu_utils.cpp:143-150: tmp = data[i]; tmp += data[i + k*n] in
 * image order; tmp / n_images. */
void oracle_image_averaging_complex(const float* images, size_t image_size, size_t n_images,
                                    float* avg) {
    for (size_t i = 0; i < image_size; i++) {
        float re = images[2 * i], im = images[2 * i + 1];
        for (size_t k = 1; k < n_images; k++) {
            re += images[2 * (i + k * image_size)];
            im += images[2 * (i + k * image_size) + 1];
        }
        avg[2 * i] = re / (float)n_images;
        avg[2 * i + 1] = im / (float)n_images;
    }
}
