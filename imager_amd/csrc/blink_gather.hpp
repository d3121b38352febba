// Device helpers shared by the gridding kernels: gathering plan entries
// (polarisation combination, fused phase correction, NaN / MAX_VIS test,
// conjugation) and a block-wide integer sum.
#pragma once

#include "blink_kernels.hpp"

namespace blink {

// U plan entries of one image gathered by one thread: slots base + u*stride
// (u < U) of the entry range [e0, e1).  The three phases are separate so
// that a kernel can issue the loads early and do other work (zeroing its LDS
// tile, copying tables) while they are in flight:
//   load_entries  the U entry words (out-of-range slots load entry e0,
//                 always valid, and are discarded later),
//   load_records  the U visibility records and phase inputs they name,
//   consume       pol combination, phase correction, NaN / MAX_VIS test,
//                 conjugation; sink(entry_index, u_idx, value) per gridded
//                 visibility, returning the sum of the sink's return values.
template <int U>
struct GatherBatch {
    int idx[U];
    uint32_t ev[U];
    float2 x0[U], x1[U];  // the one or two pol products the polarisation needs
    double cf[U];         // phase coefficient of the baseline

    __device__ __forceinline__ void load_entries(const uint32_t* __restrict__ ent, int base,
                                                 int e0, int e1, int stride) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = base + u * stride;
            idx[u] = i < e1 ? i : -1;
            ev[u] = ent[i < e1 ? i : e0];
        }
    }

    // rec: the image's records as float2 [bl][XX, XY, YX, YY]
    __device__ __forceinline__ void load_records(const GatherArgs& g, const float2* rec) {
        // XX: XX; YY: YY; I: XX and YY; V: XY and YX
        const int p0 = g.pol == 1 ? 3 : (g.pol == 3 ? 1 : 0);
        const int p1 = g.pol == 3 ? 2 : 3;
        const bool two = g.pol >= 2;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const unsigned bl = ev[u] >> 14;
            x0[u] = rec[4 * (size_t)bl + p0];
            x1[u] = two ? rec[4 * (size_t)bl + p1] : make_float2(0.f, 0.f);
            cf[u] = g.corrections ? g.phase_coef[bl] : 0.0;
        }
    }

    // the U compacted visibilities (cimg: the image's [n_bl] values, NaN where
    // not gridded) instead of the records
    __device__ __forceinline__ void load_compact(const float2* cimg) {
#pragma unroll
        for (int u = 0; u < U; u++) x0[u] = cimg[ev[u] >> 14];
    }

    template <class Sink>
    __device__ __forceinline__ int consume_compact(Sink sink) const {
        int total = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (idx[u] < 0) continue;
            float2 v = x0[u];
            if (isnan(v.x)) continue;
            if (ev[u] & (1u << 13)) v.y = -v.y;
            total += sink(idx[u], (int)(ev[u] & 0x1fffu), v);
        }
        return total;
    }

    template <class Sink>
    __device__ __forceinline__ int consume(const GatherArgs& g, double nu, Sink sink) const {
        int total = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (idx[u] < 0) continue;
            float2 v = select_pol2(x0[u], x1[u], g.pol);
            if (g.corrections) v = cmul(v, unit_phasor(((cf[u] * nu) / kVelLight)));
            if (!vis_ok(v)) continue;
            if (ev[u] & (1u << 13)) v.y = -v.y;
            total += sink(idx[u], (int)(ev[u] & 0x1fffu), v);
        }
        return total;
    }
};

// Phase coefficient of baseline bl (metres): w - (L[a2] - L[a1]), each term
// when its correction is on; exp(+2 pi i coef nu / c) is the product of the
// geometric (src/gpu/corrections_gpu.cpp:61-63) and cable (:27-28) rotations.
__device__ __forceinline__ double phase_coefficient(unsigned corrections, const float* w_m,
                                                    const double* cable_m, unsigned bl) {
    double c = 0.0;
    if (corrections & 1u) c = (double)w_m[bl];
    if (corrections & 2u) {
        unsigned a1, a2;
        baseline_antennas(bl, a1, a2);
        c = c - (cable_m[a2] - cable_m[a1]);
    }
    return c;
}

// Gathers the plan entries [e0, e1) of one image, `stride` threads apart, U
// entries per thread at a time (U independent load chains in flight).
template <int U, class Sink>
__device__ __forceinline__ int gather_range(const GatherArgs& g, int64_t img, int ch,
                                            const uint32_t* __restrict__ ent, int e0, int e1,
                                            int tid, int stride, Sink sink) {
    int total = 0;
    const double nu = g.freq_hz[ch];
    const float2* rec = (const float2*)g.vis + 4 * (img * g.n_bl);
    for (int base = e0 + tid; base < e1; base += stride * U) {
        GatherBatch<U> gb;
        gb.load_entries(ent, base, e0, e1, stride);
        gb.load_records(g, rec);
        total += gb.consume(g, nu, sink);
    }
    return total;
}

// Block-wide sum of one int per thread; the result is valid in thread 0.
__device__ __forceinline__ int block_sum_int(int x, int* scratch) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = x;
    __syncthreads();
    int t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += scratch[w];
    return t;
}

}  // namespace blink
