// Split-precision DFT-as-GEMM on MFMA for small grids (N <= 512).
//
// With the fftshift permutation p folded into the DFT matrix,
//   D[y][v] = exp(+2 pi i v p(y) / N),   image = (1/sum counts) * D G D^T.
// Both factors are products "data times D^T":
//   stage 1: T^T = (G D^T)^T,   A[m = (img, v)][k = u] = G,      out[img][x][v]
//   stage 2: I   = (T^T D^T)^T, A[m = (img, x)][k = v] = T^T,    out[img][y][x]
// so one kernel computes C'[n][m] = sum_k D'[n][k] A[m][k] with D'[n][k] =
// exp(+2 pi i k p(n) / N) and stores C'[n][m] at out[(img * N + n) * N + r],
// m = img * N + r -- rows of the MFMA output are n, columns are m, so a
// wave's store of 32 consecutive r is one contiguous 256-byte segment.
//
// Precision: every operand x is split into x_hi = bf16(x) and
// x_lo = bf16(x - x_hi); products use hi*hi + hi*lo + lo*hi (the lo*lo term
// is below FP32 rounding of the sum) with FP32 accumulation in MFMA.  D' is
// built on the host in double and split there; the data are split on the fly.
// Complex: C_re = Dr Ar - Di Ai, C_im = Dr Ai + Di Ar, the minus sign taken
// from a pre-negated copy of the imaginary planes.  12 MFMA 32x32x16 per
// k-step of 16 per wave.
#include <cmath>
#include <cstring>
#include <vector>

#include "blink_kernels.hpp"

namespace blink {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Plane order of the split DFT matrix, each [n_pad][k_pad] bf16, k contiguous.
enum { kDrH = 0, kDrL, kDiH, kDiL, kNiH, kNiL, kDftPlanes };

struct DftArgs {
    const float2* a;          // data rows [m][k], k < K contiguous (row stride K)
    const uint16_t* planes;   // kDftPlanes x [n_pad][k_pad]
    float2* out;
    int n;                    // grid side N (= K = output side)
    int k_pad, n_pad;         // padded to 16 / 64
    int64_t m_rows;           // images * N
    const uint32_t* isums;    // per-image count sums for the final scale, or null
    const double* sums;
    int64_t img0;             // global index of the first image (for the sums)
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// One workgroup = 4 waves, a 64 (n) x 64 (m) output tile; wave (wn, wm)
// computes the 32 x 32 sub-tile at (wn*32, wm*32).
__global__ void __launch_bounds__(256) k_dft_gemm(DftArgs g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wn = wave & 1, wm = wave >> 1;
    const int n_blocks = g.n_pad / 64;
    const int nb = blockIdx.x % n_blocks;
    const int64_t mb = blockIdx.x / n_blocks;
    const int n0 = nb * 64 + wn * 32;
    const int64_t m0 = mb * 64 + wm * 32;
    const int kq = (lane >> 5) * 8;            // this lane's k offset within a step
    const int nrow = n0 + (lane & 31);         // MFMA A-operand row (n)
    const int64_t mcol = m0 + (lane & 31);     // MFMA B-operand column (m)
    const bool m_ok = mcol < g.m_rows;
    const float2* arow = g.a + (m_ok ? mcol : 0) * (int64_t)g.n;
    const size_t plane = (size_t)g.n_pad * g.k_pad;
    const uint16_t* drow = g.planes + (size_t)nrow * g.k_pad;

    f32x16 cre = {}, cim = {};
    for (int k0 = 0; k0 < g.k_pad; k0 += 16) {
        const int kb = k0 + kq;
        // D' operands: 8 consecutive k of row n from each plane (16 bytes)
        bf16x8 dr_h = *(const bf16x8*)(drow + kDrH * plane + kb);
        bf16x8 dr_l = *(const bf16x8*)(drow + kDrL * plane + kb);
        bf16x8 di_h = *(const bf16x8*)(drow + kDiH * plane + kb);
        bf16x8 di_l = *(const bf16x8*)(drow + kDiL * plane + kb);
        bf16x8 ni_h = *(const bf16x8*)(drow + kNiH * plane + kb);
        bf16x8 ni_l = *(const bf16x8*)(drow + kNiL * plane + kb);
        // data operands: 8 consecutive k of row m, split into hi / lo
        bf16x8 ar_h, ar_l, ai_h, ai_l;
        float2 xs[8];
        if (m_ok && kb + 8 <= g.n) {
            const float4* p4 = (const float4*)(arow + kb);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float4 t = p4[q];
                xs[2 * q] = make_float2(t.x, t.y);
                xs[2 * q + 1] = make_float2(t.z, t.w);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++)
                xs[e] = (m_ok && kb + e < g.n) ? arow[kb + e] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float2 x = xs[e];
            __bf16 h, l;
            split_bf16(x.x, h, l);
            ar_h[e] = h;
            ar_l[e] = l;
            split_bf16(x.y, h, l);
            ai_h[e] = h;
            ai_l[e] = l;
        }
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_h, ar_h, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_h, ar_l, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_l, ar_h, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ni_h, ai_h, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ni_h, ai_l, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ni_l, ai_h, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_h, ai_h, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_h, ai_l, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dr_l, ai_h, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(di_h, ar_h, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(di_h, ar_l, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(di_l, ar_h, cim, 0, 0, 0);
    }
    // acc[i] holds C'[n = n0 + 8(i/4) + 4(lane/32) + i%4][m = m0 + lane%32]
    if (!m_ok) return;
    const int64_t img = mcol / g.n;
    const int r = (int)(mcol % g.n);
    float s = 1.0f;
    if (g.isums) s = (float)(1.0 / (double)g.isums[g.img0 + img]);
    else if (g.sums) s = (float)(1.0 / g.sums[g.img0 + img]);
    float2* obase = g.out + img * (int64_t)g.n * g.n + r;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int nn = n0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (nn < g.n) obase[(int64_t)nn * g.n] = make_float2(cre[i] * s, cim[i] * s);
    }
}

hipError_t launch_dft_gemm(const float* a, const uint16_t* planes, float* out, int n, int64_t n_img,
                           int64_t img0, const uint32_t* isums, const double* sums,
                           hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    DftArgs g{};
    g.a = (const float2*)a;
    g.planes = planes;
    g.out = (float2*)out;
    g.n = n;
    g.k_pad = (n + 15) / 16 * 16;
    g.n_pad = (n + 63) / 64 * 64;
    g.m_rows = n_img * n;
    g.isums = isums;
    g.sums = sums;
    g.img0 = img0;
    const int64_t m_blocks = (g.m_rows + 63) / 64;
    const int64_t blocks = m_blocks * (g.n_pad / 64);
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_dft_gemm, dim3((unsigned)blocks), dim3(256), 0, stream, g);
    return hipGetLastError();
}

// Host: D'[n][k] = exp(+2 pi i k p(n) / N), p = fftshift permutation, split
// into bf16 hi/lo in double precision; planes padded with zeros.
static uint16_t bf16_bits(double v) {
    // round-to-nearest-even of the float value to bf16
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

void make_dft_planes(int n, std::vector<uint16_t>& planes) {
    const int k_pad = (n + 15) / 16 * 16, n_pad = (n + 63) / 64 * 64;
    const size_t plane = (size_t)n_pad * k_pad;
    planes.assign(kDftPlanes * plane, 0);
    const int h = n / 2, o = n & 1;
    for (int r = 0; r < n; r++) {
        const int pr = r < h ? r + h + o : (o && r == h ? r : r - h - o);
        for (int k = 0; k < n; k++) {
            const long long idx = ((long long)k * pr) % n;
            const double a = 2.0 * M_PI * (double)idx / (double)n;
            const double c = std::cos(a), s = std::sin(a);
            const uint16_t ch = bf16_bits(c), sh = bf16_bits(s), nh = bf16_bits(-s);
            const size_t at = (size_t)r * k_pad + k;
            planes[kDrH * plane + at] = ch;
            planes[kDrL * plane + at] = bf16_bits(c - bf16_value(ch));
            planes[kDiH * plane + at] = sh;
            planes[kDiL * plane + at] = bf16_bits(s - bf16_value(sh));
            planes[kNiH * plane + at] = nh;
            planes[kNiL * plane + at] = bf16_bits(-s - bf16_value(nh));
        }
    }
}

size_t dft_planes_bytes(int n) {
    const int k_pad = (n + 15) / 16 * 16, n_pad = (n + 63) / 64 * 64;
    return sizeof(uint16_t) * kDftPlanes * (size_t)n_pad * k_pad;
}

}  // namespace blink
