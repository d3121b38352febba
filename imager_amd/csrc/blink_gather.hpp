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
    float4 a[U], b[U];
    float wv[U];
    double ca[U], cb[U];

    __device__ __forceinline__ void load_entries(const uint32_t* __restrict__ ent, int base,
                                                 int e0, int e1, int stride) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = base + u * stride;
            idx[u] = i < e1 ? i : -1;
            ev[u] = ent[i < e1 ? i : e0];
        }
    }

    __device__ __forceinline__ void load_records(const GatherArgs& g, const float4* vis) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const unsigned bl = ev[u] >> 14;
            a[u] = g.pol != 1 ? vis[2 * bl] : make_float4(0.f, 0.f, 0.f, 0.f);
            b[u] = g.pol != 0 ? vis[2 * bl + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
            wv[u] = (g.corrections & 1u) ? g.w_m[bl] : 0.f;
            ca[u] = cb[u] = 0.0;
            if (g.corrections & 2u) {
                unsigned a1, a2;
                baseline_antennas(bl, a1, a2);
                ca[u] = g.cable_m[a1];
                cb[u] = g.cable_m[a2];
            }
        }
    }

    template <class Sink>
    __device__ __forceinline__ int consume(const GatherArgs& g, double nu, Sink sink) const {
        int total = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (idx[u] < 0) continue;
            float2 v = select_pol(a[u], b[u], g.pol);
            if (g.corrections) {
                double turns = 0.0;
                if (g.corrections & 1u) turns += reduce_turns(geometric_turns(wv[u], nu));
                if (g.corrections & 2u) turns += reduce_turns(cable_turns(cb[u] - ca[u], nu));
                v = cmul(v, unit_phasor(turns));
            }
            if (!vis_ok(v)) continue;
            if (ev[u] & (1u << 13)) v.y = -v.y;
            total += sink(idx[u], (int)(ev[u] & 0x1fffu), v);
        }
        return total;
    }
};

// Gathers the plan entries [e0, e1) of one image, `stride` threads apart, U
// entries per thread at a time (U independent load chains in flight).
template <int U, class Sink>
__device__ __forceinline__ int gather_range(const GatherArgs& g, int64_t img, int ch,
                                            const uint32_t* __restrict__ ent, int e0, int e1,
                                            int tid, int stride, Sink sink) {
    int total = 0;
    const double nu = g.freq_hz[ch];
    const float4* vis = g.vis + 2 * (img * g.n_bl);
    for (int base = e0 + tid; base < e1; base += stride * U) {
        GatherBatch<U> gb;
        gb.load_entries(ent, base, e0, e1, stride);
        gb.load_records(g, vis);
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
