// Device helpers shared by the gridding kernels: gathering one plan entry
// (polarisation combination, fused phase correction, NaN / MAX_VIS test,
// conjugation) and a block-wide integer sum.
#pragma once

#include "blink_kernels.hpp"

namespace blink {

// Gathers the visibility of one plan entry (pol combination, optional phase
// correction), or returns false when it is not finite / above MAX_VIS.
__device__ __forceinline__ bool gather_entry(const GatherArgs& g, int64_t img, int ch, uint32_t e,
                                             int& u_idx, float2& val) {
    const unsigned bl = e >> 14;
    u_idx = (int)(e & 0x1fffu);
    const int64_t rec = img * g.n_bl + bl;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
    if (g.pol != 1) a = g.vis[2 * rec];
    if (g.pol != 0) b = g.vis[2 * rec + 1];
    float2 v = select_pol(a, b, g.pol);
    if (g.corrections) {
        const double nu = g.freq_hz[ch];
        double turns = 0.0;
        if (g.corrections & 1u) turns += reduce_turns(geometric_turns(g.w_m[bl], nu));
        if (g.corrections & 2u) {
            unsigned a1, a2;
            baseline_antennas(bl, a1, a2);
            turns += reduce_turns(cable_turns(g.cable_m[a2] - g.cable_m[a1], nu));
        }
        v = cmul(v, unit_phasor(turns));
    }
    if (!vis_ok(v)) return false;
    if (e & (1u << 13)) v.y = -v.y;
    val = v;
    return true;
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
