// Device-side helpers shared by the gfx950 kernels: complex arithmetic, the
// reference's cell-index formula (bit-exact), phase rotation with FP64
// argument reduction, and the fft_shift index permutation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blink {

// src/pacer_imager_defs.h:12
constexpr double kVelLight = 299792458.0;
// src/pacer_imager_defs.h:21 (compared in double, as fabs(float) < 1e12 is)
constexpr double kMaxVis = 1e12;
constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// i * a
__device__ __forceinline__ float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }

// wrap_index (src/pacer_imager.cpp:245-248, src/gpu/gridding_gpu.cpp:18-21)
// as a Euclidean modulo: identical wherever the reference stays in bounds
// (i >= -side), and aliased instead of out of bounds below that.
__device__ __forceinline__ int wrap_index(int i, int side) {
    // one conditional add or subtract for -side <= i < 2 side (every cell of
    // an oversampled grid), the integer division otherwise
    if (i >= -side && i < 2 * side) return i < 0 ? i + side : (i >= side ? i - side : i);
    int r = i % side;
    return r < 0 ? r + side : r;
}

// Cell of one visibility: src/pacer_imager.cpp:289 (wavelength = c/nu),
// :332-340 (u,v in wavelengths), :350,362 (uv distance cut), :364-369
// (round half away from zero, truncate, wrap).  Every operation is a single
// IEEE double operation with no contraction, so the indices equal the
// reference's bit for bit.  Returns false when the visibility is not gridded
// (min_uv cut, or a quotient too large for an int, which the reference
// leaves undefined).
__device__ __forceinline__ bool cell_of(float u_m, float v_m, double wavelength, double delta_u,
                                        double delta_v, double min_uv, int n_pixels, int& u_pix,
                                        int& v_pix) {
#pragma clang fp contract(off)
    const double u = (double)u_m / wavelength;
    const double v = (double)v_m / wavelength;
    const double uv_distance = sqrt(u * u + v * v);
    if (!(uv_distance > min_uv)) return false;
    const double ru = round(u / delta_u);
    const double rv = round(v / delta_v);
    if (!(fabs(ru) < 1073741824.0) || !(fabs(rv) < 1073741824.0)) return false;
    u_pix = (int)ru;
    v_pix = (int)rv;
    return true;
}

// Wavelength of a channel, src/pacer_imager.cpp:289.
__device__ __forceinline__ double wavelength_of(double freq_hz) {
#pragma clang fp contract(off)
    return kVelLight / freq_hz;
}

// Fraction of a turn in [-0.5, 0.5]: x - rint(x) is exact in double.
__device__ __forceinline__ double reduce_turns(double turns) { return turns - rint(turns); }

// exp(i * 2 pi * turns) with the argument reduced in FP64 and sin/cos in FP32.
// The FP32 rounding error of the reduced angle, d = theta - (float)theta, is
// folded back with a first-order correction, so the result is accurate to a
// few FP32 ulp for any |turns| < 2^52.
__device__ __forceinline__ float2 unit_phasor(double turns) {
    const double th = reduce_turns(turns) * kTwoPi;
    const float thf = (float)th;
    const float d = (float)(th - (double)thf);
    float s, c;
    sincosf(thf, &s, &c);
    return make_float2(c - d * s, s + d * c);
}

// Geometric correction angle in turns: w * nu / c (angle 2 pi w nu / c,
// src/gpu/corrections_gpu.cpp:61).
__device__ __forceinline__ double geometric_turns(float w_m, double freq_hz) {
    return ((double)w_m * freq_hz) / kVelLight;
}

// Cable correction angle in turns: -(L[a2] - L[a1]) nu / c
// (src/gpu/corrections_gpu.cpp:27-28).
__device__ __forceinline__ double cable_turns(double delta_len_m, double freq_hz) {
    return -(delta_len_m * freq_hz) / kVelLight;
}

// Baseline -> antennas, the formula of src/pacer_imager.cpp:69-70.
__device__ __forceinline__ void baseline_antennas(unsigned bl, unsigned& a1, unsigned& a2) {
    a1 = (unsigned)(-0.5 + sqrt(0.25 + 2.0 * (double)bl));
    a2 = bl - ((a1 + 1u) * a1) / 2u;
}

// fft_shift destination of index p on a side of n (src/pacer_imager.cpp:30-35,
// 110-128): the swap p <-> p + n/2 + n%2 for p < n/2.  For even n this is
// (p + n/2) % n; for odd n the middle element stays.
__device__ __forceinline__ int shift_index(int p, int n) {
    const int h = n / 2, o = n & 1;
    if (p < h) return p + h + o;
    if (o && p == h) return p;
    return p - h - o;
}

// Polarisation combination in FP32 (src/pacer_imager.cpp:301-325):
// rec points at the 4 complex pol products of one (t, ch, bl).
__device__ __forceinline__ float2 select_pol(const float4 xx_xy, const float4 yx_yy, int pol) {
    if (pol == 0) return make_float2(xx_xy.x, xx_xy.y);           // XX
    if (pol == 1) return make_float2(yx_yy.z, yx_yy.w);           // YY
    if (pol == 3) {                                               // V = (i/2)(XY - YX)
        const float dre = xx_xy.z - yx_yy.x;
        const float dim = xx_xy.w - yx_yy.y;
        return make_float2(0.0f * dre - 0.5f * dim, 0.0f * dim + 0.5f * dre);
    }
    return make_float2((xx_xy.x + yx_yy.z) / 2.0f, (xx_xy.y + yx_yy.w) / 2.0f);  // I
}

// Workgroup id in XCD-contiguous order.  Workgroups are dispatched to the 8
// XCDs round-robin (id % 8); the returned logical ids give XCD x one
// contiguous range, so neighbouring logical ids share an L2.  A bijection of
// [0, nwg) for any nwg.
__device__ __forceinline__ unsigned xcd_order(unsigned bid, unsigned nwg) {
    const unsigned q = nwg >> 3, r = nwg & 7u, x = bid & 7u;
    return x * q + min(x, r) + (bid >> 3);
}

// Streaming (non-temporal) accesses for data touched once per pass: the
// row-pass buffer and the images.  They bypass the normal cache allocation,
// leaving L2 to the visibilities and tables that are reused.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
    const f32x4_t t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, (f32x4_t*)p);
}
__device__ __forceinline__ void store_stream(float2* p, float2 v) {
    const f32x2_t t = {v.x, v.y};
    __builtin_nontemporal_store(t, (f32x2_t*)p);
}

// select_pol from the one or two products the polarisation needs: x0 = XX
// (XX, I), YY (YY) or XY (V); x1 = YY (I) or YX (V).  Same arithmetic.
__device__ __forceinline__ float2 select_pol2(float2 x0, float2 x1, int pol) {
    if (pol <= 1) return x0;                                       // XX, YY
    if (pol == 3) {                                                // V = (i/2)(XY - YX)
        const float dre = x0.x - x1.x;
        const float dim = x0.y - x1.y;
        return make_float2(0.0f * dre - 0.5f * dim, 0.0f * dim + 0.5f * dre);
    }
    return make_float2((x0.x + x1.x) / 2.0f, (x0.y + x1.y) / 2.0f);  // I
}

__device__ __forceinline__ bool vis_ok(float2 v) {
    return !isnan(v.x) && !isnan(v.y) && (double)fabsf(v.x) < kMaxVis &&
           (double)fabsf(v.y) < kMaxVis;
}

}  // namespace blink
