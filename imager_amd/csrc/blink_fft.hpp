// In-house LDS-staged mixed-radix FFT for gfx950.
//
// One-dimensional transforms of length N are done by a "team" of lanes
// (TEAM <= 64: lanes of one wavefront, several teams per wave; TEAM > 64:
// the whole workgroup) on rows held in LDS.  Each pass is a radix-R Stockham
// autosort step (Govindaraju et al., "High Performance Discrete Fourier
// Transforms on Graphics Processors", SC'08): butterfly j reads
// x[j + r*N/R], multiplies by the twiddle w^(r*(j mod Ns)) of the current
// span Ns*R, does a radix-R DFT and writes
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  Passes ping-pong between two LDS rows,
// so no index reversal is needed and every butterfly needs only R registers.
// Radices 2, 3, 4, 5, 8 and 16, and the primes 7, 11 and 13 as direct
// sums, cover N = 2^a 3^b 5^c 7^d 11^e 13^f.
//
// Twiddles come from a table tw[m] = exp(+2 pi i m / N), m < N, built on the
// host in double and rounded to FP32 (accurate to 1/2 ulp); the forward
// direction uses the conjugate.
#pragma once

#include "blink_device.hpp"

namespace blink {

constexpr int kMaxFftPasses = 16;

struct FftPlan {
    int n;                       // transform length
    int n_passes;
    int radix[kMaxFftPasses];    // radix of each pass
    int span[kMaxFftPasses];     // Ns before each pass (product of previous radices)
};

// Synchronisation of the lanes working on one row.
template <bool kBlockTeam>
__device__ __forceinline__ void team_sync() {
    if constexpr (kBlockTeam) {
        __syncthreads();
    } else {
        // One wavefront executes its LDS operations in order; the fences keep
        // the compiler from moving LDS accesses across this point.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Radix-R DFTs with sign S: y[k] = sum_r x[r] exp(S 2 pi i r k / R).
template <int S>
__device__ __forceinline__ void dft2(float2* x) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <int S>
__device__ __forceinline__ float2 rot_s(float2 a) {  // S * i * a
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int S>
__device__ __forceinline__ void dft4(float2* x) {
    const float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    const float2 c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    const float2 id = rot_s<S>(d);
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, id);
    x[3] = csub(b, id);
}

template <int S>
__device__ __forceinline__ void dft3(float2* x) {
    constexpr float kC = -0.5f;
    constexpr float kS = 0.86602540378443864676f;  // sqrt(3)/2
    const float2 t = cadd(x[1], x[2]);
    const float2 d = csub(x[1], x[2]);
    const float2 m = make_float2(x[0].x + kC * t.x, x[0].y + kC * t.y);
    const float2 r = rot_s<S>(cscale(d, kS));
    x[0] = cadd(x[0], t);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
}

template <int S>
__device__ __forceinline__ void dft5(float2* x) {
    constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
    constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
    constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
    const float2 t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    const float2 t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
    const float2 a1 = make_float2(x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y);
    const float2 b1 = rot_s<S>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    const float2 b2 = rot_s<S>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, b1);
    x[4] = csub(a1, b1);
    x[2] = cadd(a2, b2);
    x[3] = csub(a2, b2);
}

// exp(S 2 pi i k / 16), k < 8 (folds to constants once loops are unrolled).
template <int S>
__device__ __forceinline__ float2 root16(int k) {
    float c, sn;
    switch (k) {
        case 0: c = 1.0f; sn = 0.0f; break;
        case 1: c = 0.92387953251128675613f; sn = 0.38268343236508977173f; break;
        case 2: c = 0.70710678118654752440f; sn = 0.70710678118654752440f; break;
        case 3: c = 0.38268343236508977173f; sn = 0.92387953251128675613f; break;
        case 4: c = 0.0f; sn = 1.0f; break;
        case 5: c = -0.38268343236508977173f; sn = 0.92387953251128675613f; break;
        case 6: c = -0.70710678118654752440f; sn = 0.70710678118654752440f; break;
        default: c = -0.92387953251128675613f; sn = 0.38268343236508977173f; break;
    }
    return make_float2(c, S > 0 ? sn : -sn);
}

template <int S>
__device__ __forceinline__ void dft8(float2* x) {
    float2 e[4] = {x[0], x[2], x[4], x[6]};
    float2 o[4] = {x[1], x[3], x[5], x[7]};
    dft4<S>(e);
    dft4<S>(o);
    o[1] = cmul(o[1], root16<S>(2));
    o[2] = rot_s<S>(o[2]);
    o[3] = cmul(o[3], root16<S>(6));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = cadd(e[k], o[k]);
        x[k + 4] = csub(e[k], o[k]);
    }
}

template <int S>
__device__ __forceinline__ void dft16(float2* x) {
    float2 e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        e[k] = x[2 * k];
        o[k] = x[2 * k + 1];
    }
    dft8<S>(e);
    dft8<S>(o);
#pragma unroll
    for (int k = 1; k < 8; k++) o[k] = (k == 4) ? rot_s<S>(o[k]) : cmul(o[k], root16<S>(k));
#pragma unroll
    for (int k = 0; k < 8; k++) {
        x[k] = cadd(e[k], o[k]);
        x[k + 8] = csub(e[k], o[k]);
    }
}

// Radix-R DFT for a prime R (7, 11, 13) as a direct sum,
// y[k] = sum_r x[r] w^(r k) with w = exp(S 2 pi i / R) = tw[n / R] (the
// conjugate for S < 0): R^2 complex products, for the rare sides that need it.
template <int S, int R>
__device__ __forceinline__ void dft_prime(float2* x, const float2* __restrict__ tw, int n) {
    const int step = n / R;
    float2 y[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        float2 acc = x[0];
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = tw[((r * k) % R) * step];
            if (S < 0) w.y = -w.y;
            acc = cadd(acc, cmul(x[r], w));
        }
        y[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = y[k];
}

// One Stockham pass of compile-time radix R over a row: src -> dst.
template <int S, int R>
__device__ __forceinline__ void stockham_pass(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int n, int ns, const float2* __restrict__ tw,
                                              int lane, int team) {
    const int nb = n / R;
    const int tstep = n / (ns * R);  // twiddle index step per (r * k)
    for (int j = lane; j < nb; j += team) {
        float2 v[R];
        const int k = j % ns;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
        if (ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = tw[r * k * tstep];
                if (S < 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        if constexpr (R == 2) dft2<S>(v);
        else if constexpr (R == 3) dft3<S>(v);
        else if constexpr (R == 4) dft4<S>(v);
        else if constexpr (R == 5) dft5<S>(v);
        else if constexpr (R == 8) dft8<S>(v);
        else if constexpr (R == 7 || R == 11 || R == 13) dft_prime<S, R>(v, tw, n);
        else dft16<S>(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[base + r * ns] = v[r];
    }
}

// Transforms the row in buf (LDS, length plan.n) using tmp (LDS, same length)
// as the ping-pong partner.  Returns the buffer holding the result (buf or
// tmp).  All lanes of the team must call it.
template <int S, bool kBlockTeam>
__device__ float2* fft_row(float2* buf, float2* tmp, const FftPlan& plan,
                           const float2* __restrict__ tw, int lane, int team) {
    float2* src = buf;
    float2* dst = tmp;
    for (int p = 0; p < plan.n_passes; p++) {
        const int R = plan.radix[p];
        const int ns = plan.span[p];
        switch (R) {
            case 2: stockham_pass<S, 2>(src, dst, plan.n, ns, tw, lane, team); break;
            case 3: stockham_pass<S, 3>(src, dst, plan.n, ns, tw, lane, team); break;
            case 4: stockham_pass<S, 4>(src, dst, plan.n, ns, tw, lane, team); break;
            case 5: stockham_pass<S, 5>(src, dst, plan.n, ns, tw, lane, team); break;
            case 7: stockham_pass<S, 7>(src, dst, plan.n, ns, tw, lane, team); break;
            case 8: stockham_pass<S, 8>(src, dst, plan.n, ns, tw, lane, team); break;
            case 11: stockham_pass<S, 11>(src, dst, plan.n, ns, tw, lane, team); break;
            case 13: stockham_pass<S, 13>(src, dst, plan.n, ns, tw, lane, team); break;
            default: stockham_pass<S, 16>(src, dst, plan.n, ns, tw, lane, team); break;
        }
        team_sync<kBlockTeam>();
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

}  // namespace blink
