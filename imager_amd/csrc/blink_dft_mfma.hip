// Split-precision DFT-as-GEMM on MFMA for small grids (N <= 512).
//
// The gridded plane G is Hermitian (every visibility lands at (u, v) and its
// conjugate at (-u, -v)), so only rows v = 0 .. H = N/2 are gridded, into
// half[img][v][u] (row-major, N complex per row).  With the fftshift
// permutation p folded into the DFT matrices,
//   stage 1 (complex):  T[v][x] = sum_u  D1[x][u] G[v][u],
//                       D1[x][u] = exp(+2 pi i u p(x) / N),     v = 0 .. H;
//   stage 2 (real):     I[y][x] = sum_v  Re(D2[y][v] T[v][x]),
//                       D2[y][v] = w_v exp(+2 pi i v p(y) / N),
// with w_0 = w_H = 1 and w_v = 2 otherwise: row N - v of D1 G is the conjugate
// of row v, so the pair v, N - v contributes 2 Re(D[y][v] T[v][x]) and the
// image is real.  Stage 2 multiplies by (float)(1 / sum counts) and writes the
// imaginary part as 0.
//
// Both stages are one kernel, C[n][m] = sum_k D[n][k] A[m][k]:
//   stage 1: A[m = (img, v)][k = u] = half rows,   C stored at inter[img][x = n][v]
//   stage 2: A[m = (img, x)][k = v] = inter rows,  C stored at images[img][y = n][x]
// Rows of the MFMA result are n and columns are m, so the 32 lanes that hold
// one n store 32 consecutive m: one 256-byte segment.
//
// Precision: every operand x is split into hi = P(x) and lo = P(x - hi),
// P = bf16 or fp16, and products use hi*hi + hi*lo + lo*hi (lo*lo is below
// the FP32 rounding of the sum) with FP32 accumulation in MFMA.  D1 and D2
// are built on the host in double and split there; the data are split on the
// fly.  fp16 has 5 exponent bits, so each data row m is first scaled by a
// power of two 2^e that brings its largest component below 2^15 (exact), and
// the result column m is scaled back by 2^-e.  Complex products use
// C_re = Dr Ar + Di (-Ai) and C_im = Dr Ai + Di Ar; -Ai is the sign-flipped
// split of Ai.
#include <cmath>
#include <cstring>
#include <vector>

#include "blink_kernels.hpp"

namespace blink {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Plane order of a split DFT matrix, each [n_pad][k_pad] 16-bit values, k contiguous.
enum { kReHi = 0, kReLo, kImHi, kImLo, kDftPlanes };

template <int kPrec>
struct Prec;
template <>
struct Prec<kDftBf16> {
    using V = bf16x8;
    static constexpr bool kScale = false;
    using E = __bf16;
    __device__ static f32x16 mfma(V a, V b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Prec<kDftFp16> {
    using V = f16x8;
    static constexpr bool kScale = true;
    using E = _Float16;
    __device__ static f32x16 mfma(V a, V b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <typename V>
__device__ __forceinline__ V neg8(V x) {
    u32x4 b = __builtin_bit_cast(u32x4, x);
    b ^= 0x80008000u;
    return __builtin_bit_cast(V, b);
}

struct DftGemmArgs {
    const float2* a;          // data rows [m][lda], k < K valid
    int64_t lda;              // row stride of a (float2, even)
    int64_t m_rows;           // M
    int K, k_pad;             // valid k; k_pad a multiple of 16
    int n, n_pad;             // output rows n < N; n_pad a multiple of 64
    const uint16_t* planes;   // kDftPlanes x [n_pad][k_pad]
    int R;                    // m -> (img = m / R, r = m % R)
    int64_t img_stride, n_stride;  // C[n][m] at out[img * img_stride + n * n_stride + r]
    float2* out;
    const uint32_t* isums;    // stage 2: per-image count sums (set-relative), else null
    double* count_sums;       // stage 2: optional, isums as doubles
    int64_t img0;             // set-relative index of the first image
};

// 8 consecutive complex values of a data row from k = kb, zero beyond K or
// for rows past M.
__device__ __forceinline__ void load8(const float2* row, bool ok, int kb, int K, float2 (&xs)[8]) {
    if (ok && kb + 8 <= K) {
        const float4* p4 = (const float4*)(row + kb);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 t = p4[q];
            xs[2 * q] = make_float2(t.x, t.y);
            xs[2 * q + 1] = make_float2(t.z, t.w);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++)
            xs[e] = (ok && kb + e < K) ? row[kb + e] : make_float2(0.f, 0.f);
    }
}

// Workgroup = 8 waves sharing one slice of NS = 32 * NSUB rows of D, held in
// LDS for the workgroup's lifetime (rows padded by 16 bytes so that the 32
// lanes reading one k column hit distinct banks).  The workgroup walks the m
// tiles of 256 rows (one 32-row MFMA tile per wave) with a stride of the
// number of workgroups per slice; each wave splits its data rows once per
// k-step and applies them to its NSUB D tiles.  The next k-step's data are
// loaded before the current step's MFMAs are issued.
constexpr int kGemmThreads = 512;
constexpr int kGemmRows = 32 * (kGemmThreads / 64);  // m rows per tile

__host__ __device__ constexpr int dft_lds_ld(int k_pad) { return k_pad + 8; }

template <int kPrec, bool kComplex, int NSUB>
__global__ void __launch_bounds__(kGemmThreads) k_dft_gemm(DftGemmArgs g) {
    using S = Prec<kPrec>;
    using V = typename S::V;
    using E = typename S::E;
    constexpr int NS = 32 * NSUB;
    extern __shared__ float4 dft_lds4[];
    uint16_t* dl = (uint16_t*)dft_lds4;  // [kDftPlanes][NS][ldk]
    const int ldk = dft_lds_ld(g.k_pad);
    const int n_slices = g.n_pad / NS;
    const unsigned bid = xcd_order(blockIdx.x, gridDim.x);
    const int slice = (int)(bid % n_slices);
    const int grp = (int)(bid / n_slices);
    const int n_grp = (int)(gridDim.x / n_slices);
    const int n_base = slice * NS;

    // the D slice, 16 bytes per thread and step
    {
        const size_t plane = (size_t)g.n_pad * g.k_pad;
        const int cpr = g.k_pad / 8;  // 16-byte chunks per row
        const int per_plane = NS * cpr;
        for (int i = threadIdx.x; i < kDftPlanes * per_plane; i += kGemmThreads) {
            const int p = i / per_plane, rem = i % per_plane;
            const int r = rem / cpr, c = rem % cpr;
            *(uint4*)(dl + ((size_t)p * NS + r) * ldk + 8 * c) =
                *(const uint4*)(g.planes + p * plane + (size_t)(n_base + r) * g.k_pad + 8 * c);
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int kq = (lane >> 5) * 8;  // this lane's k offset within a step
    const size_t lplane = (size_t)NS * ldk;
    const uint16_t* dlane = dl + (size_t)(lane & 31) * ldk + kq;
    const int64_t m_tiles = (g.m_rows + kGemmRows - 1) / kGemmRows;

    for (int64_t mt = grp; mt < m_tiles; mt += n_grp) {
        const int64_t m0 = mt * kGemmRows + wave * 32;
        if (m0 >= g.m_rows) continue;  // wave-uniform: no rows for this wave
        const int64_t mrow = m0 + (lane & 31);
        const bool mok = mrow < g.m_rows;
        const float2* arow = g.a + (mok ? mrow : 0) * g.lda + kq;
        const int kv = g.K - kq;  // valid k relative to this lane's offset

        // fp16: power-of-two scale from the row's largest component
        float sc = 1.f, unsc = 1.f;
        if constexpr (S::kScale) {
            float mx = 0.f;
            for (int k0 = 0; k0 < g.k_pad; k0 += 16) {
                float2 xs[8];
                load8(arow, mok, k0, kv, xs);
#pragma unroll
                for (int e = 0; e < 8; e++) mx = fmaxf(mx, fmaxf(fabsf(xs[e].x), fabsf(xs[e].y)));
            }
            mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
            if (mx > 0.f && mx < INFINITY) {
                int ex;
                (void)frexpf(mx, &ex);                       // mx < 2^ex
                const int e = max(-126, min(126, 15 - ex));  // mx * 2^e < 2^15
                sc = __int_as_float((127 + e) << 23);
                unsc = __int_as_float((127 - e) << 23);
            }
        }

        f32x16 cre[NSUB], cim[NSUB];
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            cre[s] = f32x16{};
            cim[s] = f32x16{};
        }
        float2 xs[8];
        load8(arow, mok, 0, kv, xs);
        for (int k0 = 0; k0 < g.k_pad; k0 += 16) {
            V ar_h, ar_l, ai_h, ai_l;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                float2 x = xs[e];
                if constexpr (S::kScale) {
                    x.x *= sc;
                    x.y *= sc;
                }
                const E rh = (E)x.x, ih = (E)x.y;
                ar_h[e] = rh;
                ar_l[e] = (E)(x.x - (float)rh);
                ai_h[e] = ih;
                ai_l[e] = (E)(x.y - (float)ih);
            }
            if (k0 + 16 < g.k_pad) load8(arow, mok, k0 + 16, kv, xs);
            const V ni_h = neg8(ai_h), ni_l = neg8(ai_l);
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                const uint16_t* dp = dlane + (size_t)s * 32 * ldk + k0;
                const V dr_h = *(const V*)(dp + kReHi * lplane);
                const V dr_l = *(const V*)(dp + kReLo * lplane);
                const V di_h = *(const V*)(dp + kImHi * lplane);
                const V di_l = *(const V*)(dp + kImLo * lplane);
                cre[s] = S::mfma(dr_h, ar_h, cre[s]);
                cre[s] = S::mfma(dr_h, ar_l, cre[s]);
                cre[s] = S::mfma(dr_l, ar_h, cre[s]);
                cre[s] = S::mfma(di_h, ni_h, cre[s]);
                cre[s] = S::mfma(di_h, ni_l, cre[s]);
                cre[s] = S::mfma(di_l, ni_h, cre[s]);
                if constexpr (kComplex) {
                    cim[s] = S::mfma(dr_h, ai_h, cim[s]);
                    cim[s] = S::mfma(dr_h, ai_l, cim[s]);
                    cim[s] = S::mfma(dr_l, ai_h, cim[s]);
                    cim[s] = S::mfma(di_h, ar_h, cim[s]);
                    cim[s] = S::mfma(di_h, ar_l, cim[s]);
                    cim[s] = S::mfma(di_l, ar_h, cim[s]);
                }
            }
        }

        // acc[i] holds C[n = n_base + 32 s + 8(i/4) + 4(lane/32) + i%4][m = mrow]
        if (!mok) continue;
        const int64_t img = mrow / g.R;
        const int r = (int)(mrow % g.R);
        float scale = unsc;
        if (g.isums) {
            const uint32_t count = g.isums[g.img0 + img];
            scale *= (float)(1.0 / (double)count);
            if (g.count_sums && r == 0 && n_base == 0 && lane < 32)
                g.count_sums[g.img0 + img] = (double)count;
        }
        float2* ob = g.out + img * g.img_stride + r;
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int nn = n_base + 32 * s + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                if (nn >= g.n) continue;
                if constexpr (kComplex)
                    ob[(int64_t)nn * g.n_stride] =
                        make_float2(cre[s][i] * scale, cim[s][i] * scale);
                else
                    store_stream(&ob[(int64_t)nn * g.n_stride], make_float2(cre[s][i] * scale, 0.f));
            }
        }
    }
}

static int pad16(int k) { return (k + 15) / 16 * 16; }
static int pad64(int n) { return (n + 63) / 64 * 64; }

int dft_inter_stride(int n) { return pad16(n / 2 + 1); }

constexpr int kGemmLdsMax = 160 * 1024;

// LDS of a workgroup holding NS = 32 * nsub rows of D.
static size_t gemm_lds(int k_pad, int nsub) {
    return sizeof(uint16_t) * kDftPlanes * (size_t)(32 * nsub) * dft_lds_ld(k_pad);
}

template <int kPrec, bool kComplex>
static void gemm_kernels(const void* (&k)[2]) {
    k[0] = (const void*)k_dft_gemm<kPrec, kComplex, 1>;
    k[1] = (const void*)k_dft_gemm<kPrec, kComplex, 2>;
}

hipError_t set_dft_kernel_attributes() {
    const void* ks[8][2];
    gemm_kernels<kDftBf16, true>(ks[0]);
    gemm_kernels<kDftBf16, false>(ks[1]);
    gemm_kernels<kDftFp16, true>(ks[2]);
    gemm_kernels<kDftFp16, false>(ks[3]);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 2; j++) {
            const hipError_t e = hipFuncSetAttribute(
                ks[i][j], hipFuncAttributeMaxDynamicSharedMemorySize, kGemmLdsMax);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

template <bool kComplex>
static hipError_t launch_gemm(const DftGemmArgs& g, int prec, hipStream_t stream) {
    if (g.m_rows == 0) return hipSuccess;
    // two D tiles per wave when the slice fits, else one
    const int nsub = (g.n_pad % 64 == 0 && gemm_lds(g.k_pad, 2) <= (size_t)kGemmLdsMax) ? 2 : 1;
    const size_t lds = gemm_lds(g.k_pad, nsub);
    if (lds > (size_t)kGemmLdsMax) return hipErrorInvalidConfiguration;
    const int n_cu = device_cu_count();
    const int n_slices = g.n_pad / (32 * nsub);
    const int64_t m_tiles = (g.m_rows + kGemmRows - 1) / kGemmRows;
    const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(2, kGemmLdsMax / lds));
    const int64_t want = (n_cu * per_cu + n_slices - 1) / n_slices;  // workgroups per slice
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(m_tiles, want));
    const int64_t blocks = groups * n_slices;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks), block(kGemmThreads);
    if (prec == kDftFp16) {
        if (nsub == 2)
            hipLaunchKernelGGL((k_dft_gemm<kDftFp16, kComplex, 2>), grid, block, lds, stream, g);
        else
            hipLaunchKernelGGL((k_dft_gemm<kDftFp16, kComplex, 1>), grid, block, lds, stream, g);
    } else {
        if (nsub == 2)
            hipLaunchKernelGGL((k_dft_gemm<kDftBf16, kComplex, 2>), grid, block, lds, stream, g);
        else
            hipLaunchKernelGGL((k_dft_gemm<kDftBf16, kComplex, 1>), grid, block, lds, stream, g);
    }
    return hipGetLastError();
}

hipError_t launch_dft_rows(const float* half, const uint16_t* planes1, float* inter, int n,
                           int64_t n_img, int prec, hipStream_t stream) {
    const int h = n / 2;
    DftGemmArgs g{};
    g.a = (const float2*)half;
    g.lda = n;
    g.m_rows = n_img * (h + 1);
    g.K = n;
    g.k_pad = pad16(n);
    g.n = n;
    g.n_pad = pad64(n);
    g.planes = planes1;
    g.R = h + 1;
    g.n_stride = dft_inter_stride(n);
    g.img_stride = (int64_t)n * g.n_stride;
    g.out = (float2*)inter;
    return launch_gemm<true>(g, prec, stream);
}

hipError_t launch_dft_cols(const float* inter, const uint16_t* planes2, float* images, int n,
                           int64_t n_img, int64_t img0, const uint32_t* isums,
                           double* count_sums, int prec, hipStream_t stream) {
    const int h = n / 2;
    DftGemmArgs g{};
    g.a = (const float2*)inter;
    g.lda = dft_inter_stride(n);
    g.m_rows = n_img * n;
    g.K = h + 1;
    g.k_pad = pad16(h + 1);
    g.n = n;
    g.n_pad = pad64(n);
    g.planes = planes2;
    g.R = n;
    g.n_stride = n;
    g.img_stride = (int64_t)n * n;
    g.out = (float2*)images + img0 * (int64_t)n * n;
    g.isums = isums;
    g.count_sums = count_sums;
    g.img0 = img0;
    return launch_gemm<false>(g, prec, stream);
}

// Host: the split planes of D1 (stage 1) or D2 (stage 2) in double, rounded
// to nearest even; zero padding beyond N rows and K columns.
static uint16_t bf16_bits(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
static double bf16_value(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static uint16_t fp16_bits(double v) {
    const _Float16 h = (_Float16)v;
    uint16_t b;
    memcpy(&b, &h, 2);
    return b;
}
static double fp16_value(uint16_t b) {
    _Float16 h;
    memcpy(&h, &b, 2);
    return (double)h;
}

void make_dft_planes(int n, int stage, int prec, std::vector<uint16_t>& planes) {
    const int h = n / 2;
    const int K = stage == 1 ? n : h + 1;
    const int k_pad = pad16(K), n_pad = pad64(n);
    const size_t plane = (size_t)n_pad * k_pad;
    planes.assign(kDftPlanes * plane, 0);
    auto bits = [&](double v) { return prec == kDftFp16 ? fp16_bits(v) : bf16_bits(v); };
    auto value = [&](uint16_t b) { return prec == kDftFp16 ? fp16_value(b) : bf16_value(b); };
    for (int r = 0; r < n; r++) {
        const int pr = r < h ? r + h : r - h;  // fftshift: output r takes frequency index pr
        for (int k = 0; k < K; k++) {
            const long long idx = ((long long)k * pr) % n;
            const double a = 2.0 * M_PI * (double)idx / (double)n;
            const double w = (stage == 1 || k == 0 || k == h) ? 1.0 : 2.0;
            const double c = w * std::cos(a), s = w * std::sin(a);
            const uint16_t ch = bits(c), sh = bits(s);
            const size_t at = (size_t)r * k_pad + k;
            planes[kReHi * plane + at] = ch;
            planes[kReLo * plane + at] = bits(c - value(ch));
            planes[kImHi * plane + at] = sh;
            planes[kImLo * plane + at] = bits(s - value(sh));
        }
    }
}

size_t dft_planes_bytes(int n, int stage) {
    const int K = stage == 1 ? n : n / 2 + 1;
    return sizeof(uint16_t) * kDftPlanes * (size_t)pad64(n) * pad16(K);
}

}  // namespace blink
