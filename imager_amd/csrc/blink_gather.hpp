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

// Per-channel constants of the plan's cell computation: the wavelength, the
// reciprocals of the wavelength and of the cell sizes (div_rn), and the
// products k = 1 / (wavelength * cell size) of the one-product rounding
// (round_quotient_fast); `fast` when all of them are in div_rn_range.
struct CellUnit {
    double wl, r_wl, r_du, r_dv, ku, kv;
    bool fast;
    __device__ __forceinline__ CellUnit(double freq_hz, const SetDesc& sd) {
#pragma clang fp contract(off)
        wl = wavelength_of(freq_hz);
        r_wl = 1.0 / wl;
        r_du = 1.0 / sd.du;
        r_dv = 1.0 / sd.dv;
        ku = 1.0 / (wl * sd.du);
        kv = 1.0 / (wl * sd.dv);
        fast = div_rn_range(wl) && div_rn_range(sd.du) && div_rn_range(sd.dv);
    }
};

// cell_of with the min_uv test reduced to a NaN test when min_uv is -inf
// (sqrt(u^2 + v^2) > -inf holds for every non-NaN u, v) and the divisions by
// the wavelength and the cell sizes done with div_rn and per-unit reciprocals:
// the same quotients, so the same cells.  Operands outside div_rn_range (a
// zero, huge, infinite or NaN coordinate or cell size) take the IEEE division.
// Almost every cell is decided first by one product per coordinate
// (round_quotient_fast), which gives the same integers.
__device__ __forceinline__ bool plan_cell(float u_m, float v_m, const CellUnit& cu,
                                          const SetDesc& sd, int& up, int& vp) {
#pragma clang fp contract(off)
    if (sd.min_uv == -INFINITY) {
        const double um = (double)u_m, vm = (double)v_m;
        if (cu.fast && round_quotient_fast(um, cu.ku, up) && round_quotient_fast(vm, cu.kv, vp))
            return true;
        const bool fast = cu.fast && div_rn_range(um) && div_rn_range(vm);
        double u, v, qu, qv;
        if (fast) {
            u = div_rn(um, cu.wl, cu.r_wl);
            v = div_rn(vm, cu.wl, cu.r_wl);
            qu = div_rn(u, sd.du, cu.r_du);
            qv = div_rn(v, sd.dv, cu.r_dv);
        } else {
            u = um / cu.wl;
            v = vm / cu.wl;
            qu = u / sd.du;
            qv = v / sd.dv;
        }
        if (isnan(u) || isnan(v)) return false;
        const double ru = round(qu);
        const double rv = round(qv);
        if (!(fabs(ru) < 1073741824.0) || !(fabs(rv) < 1073741824.0)) return false;
        up = (int)ru;
        vp = (int)rv;
        return true;
    }
    return cell_of(u_m, v_m, cu.wl, sd.du, sd.dv, sd.min_uv, sd.n_pixels, up, vp);
}

// The direct row pass's walk over all baselines of one image (no plan), `nt`
// threads, U baselines per thread and round.  Every load of a round (flag,
// u, v, the record or two, the phase coefficient) is issued together, and
// the next round's loads are issued before this round's arithmetic, so that
// the memory traffic of one round overlaps the work of the previous one (all
// workgroups walk in step, so without it the whole chip alternates between
// load bursts and arithmetic).  Each baseline's cell is the plan's
// (plan_cell, transposed when uv_swap); the corrected visibility is added at
// its cell and its conjugate at the mirrored cell (N - r, N - c) when those
// fall in the plan rows r0 .. r0 + nr - 1, through at(row - r0, position),
// an LDS float2.  Returns the count the plan path gives: entries in the
// self-conjugate rows (0, and N/2 for even N) count once, others twice.
// kWin: the rows and at() cover a window of the plane that every cell is
// expected to fall in; at() returns nullptr for a position outside it, and a
// cell outside the window's rows or positions sets `outside` (the caller then
// discards the walk).  `top`, when given, is raised to the highest row (from
// r0) that holds a cell of the thread's gridded baselines (whatever their
// values: the same rows for every image of a channel).
template <int U>
struct WalkBatch {
    int fl[U];
    float uu[U], vv[U];
    float2 x0[U], x1[U];
    double cf[U];
    // Branch-free: indices past the end read the last baseline and are
    // flagged, and the second pol product and the phase coefficient are read
    // whether or not they are used (the same 32-byte record; the coefficient
    // table always exists), so that the compiler can keep the loads of the
    // next round in flight while this round is consumed.
    __device__ __forceinline__ void load(const SetDesc& sd, const float2* rec, int p0, int p1,
                                         int b0, int tid, int nt) {
        const int n_bl = sd.g.n_bl;
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int b = b0 + tid + k * nt;
            const int bl = min(b, n_bl - 1);
            const int f = sd.flags[bl];
            fl[k] = b < n_bl ? f : 1;
            uu[k] = sd.u_m[bl];
            vv[k] = sd.v_m[bl];
            x0[k] = rec[4 * (size_t)bl + p0];
            x1[k] = rec[4 * (size_t)bl + p1];
            cf[k] = sd.g.phase_coef[bl];
        }
    }
};

template <int U, bool kWin = false, class At>
__device__ __forceinline__ int direct_walk(const SetDesc& sd, int64_t img, int ch, int n, int r0,
                                           int nr, int tid, int nt, At at,
                                           bool* outside = nullptr, int* top = nullptr) {
    const GatherArgs& g = sd.g;
    const double nu = g.freq_hz[ch];
    const CellUnit cu(nu, sd);
    const float2* rec = (const float2*)g.vis + 4 * (img * g.n_bl);
    // XX: XX; YY: YY; I: XX and YY; V: XY and YX
    const int p0 = g.pol == 1 ? 3 : (g.pol == 3 ? 1 : 0);
    const int p1 = g.pol == 3 ? 2 : 3;
    const int n_bl = g.n_bl;
    int n_ok = 0;
    if (n_bl <= 0) return 0;
    WalkBatch<U> cur;
    cur.load(sd, rec, p0, p1, 0, tid, nt);
    for (int b0 = 0; b0 < n_bl; b0 += nt * U) {
        WalkBatch<U> nxt;
        nxt.load(sd, rec, p0, p1, b0 + nt * U, tid, nt);  // past the end: flagged
#pragma unroll
        for (int k = 0; k < U; k++) {
            int up, vp;
            if (cur.fl[k] || !plan_cell(cur.uu[k], cur.vv[k], cu, sd, up, vp)) continue;
            const int r = wrap_index(sd.uv_swap ? up : vp, n);
            const int c = wrap_index(sd.uv_swap ? vp : up, n);
            const int r2 = r == 0 ? 0 : n - r;
            const bool in1 = (unsigned)(r - r0) < (unsigned)nr;
            const bool in2 = (unsigned)(r2 - r0) < (unsigned)nr;
            if (!in1 && !in2) {
                if constexpr (kWin) *outside = true;
                continue;
            }
            // (before the value test: the rows depend on the cells only, the
            // same for every image of a channel)
            if (top) *top = max(*top, max(in1 ? r - r0 : -1, in2 ? r2 - r0 : -1));
            float2 v = select_pol2(cur.x0[k], cur.x1[k], g.pol);
            if (g.corrections) v = cmul(v, unit_phasor(((cur.cf[k] * nu) / kVelLight)));
            if (!vis_ok(v)) continue;
            if (in1) {
                float2* d = at(r - r0, c);
                if constexpr (kWin)
                    if (!d) {
                        *outside = true;
                        continue;
                    }
                atomicAdd(&d->x, v.x);
                atomicAdd(&d->y, v.y);
                n_ok += (r == 0 || 2 * r == n) ? 1 : 2;
            }
            if (in2) {
                float2* d = at(r2 - r0, c == 0 ? 0 : n - c);
                if constexpr (kWin)
                    if (!d) {
                        *outside = true;
                        continue;
                    }
                atomicAdd(&d->x, v.x);
                atomicAdd(&d->y, -v.y);
                n_ok += (r2 == 0 || 2 * r2 == n) ? 1 : 2;
            }
        }
        cur = nxt;
    }
    return n_ok;
}

// Column (or row) of footprint coordinate p for sign s, as calc_xy_grid2
// computes it (src/pacer_imager.cpp:716-750): index s*p + N/2, then the
// half swap; coordinates that land outside the grid are dropped.
__device__ __forceinline__ bool kb_cell(int p, int s, int n, int& x) {
    const int c = n / 2;
    const int idx = s * p + c;
    x = idx < c ? c + idx : idx - c;
    return x >= 0 && x < n;
}

// The direct row passes' walk for convolutional gridding (transposed plan:
// rows are u): every gridded baseline of the image adds its K x K footprint
// around its nearest cell with value V and, with conj(V), around the mirrored
// cell (as k_grid_rows_kb adds them per plan entry: the kernel rows from the
// 1023-times oversampled fractional position, weights in FP64), keeping the
// cells of plan rows r0 .. r0 + nr - 1 (of 0 .. N/2) through at(row - r0,
// position).  kWin: rows 0 .. nr - 1 are a window that every cell is expected
// to fall in; a cell of rows nr .. N/2, or a position at() has no slot for
// (nullptr), sets *outside.  Returns the nearest-cell counts as direct_walk
// does (two per gridded visibility, split by row).
template <bool kWin, class At>
__device__ __forceinline__ int kb_walk(const SetDesc& sd, int64_t img, int ch, int n, int r0,
                                       int nr, int tid, int nt, At at, bool* outside) {
#pragma clang fp contract(off)
    const GatherArgs& g = sd.g;
    const double nu = g.freq_hz[ch];
    const double wl = wavelength_of(nu);
    const float2* rec = (const float2*)g.vis + 4 * (img * g.n_bl);
    const int p0 = g.pol == 1 ? 3 : (g.pol == 3 ? 1 : 0);
    const int p1 = g.pol == 3 ? 2 : 3;
    const int K = sd.kb_size, mid = K / 2, h = n / 2;
    int n_ok = 0;
    for (int bl = tid; bl < g.n_bl; bl += nt) {
        if (sd.flags[bl]) continue;
        float2 v = select_pol2(rec[4 * (size_t)bl + p0],
                               g.pol >= 2 ? rec[4 * (size_t)bl + p1] : make_float2(0.f, 0.f),
                               g.pol);
        if (g.corrections) v = cmul(v, unit_phasor(((g.phase_coef[bl] * nu) / kVelLight)));
        if (!vis_ok(v)) continue;
        int up, vp;
        if (!cell_of(sd.u_m[bl], sd.v_m[bl], wl, sd.du, sd.dv, sd.min_uv, n, up, vp)) continue;
        // the nearest cell's plan row and its mirror's: the counts
        const int q1 = wrap_index(up, n), q2 = q1 == 0 ? 0 : n - q1;
        if ((unsigned)(q1 - r0) < (unsigned)nr && q1 <= h) n_ok += (q1 == 0 || 2 * q1 == n) ? 1 : 2;
        if ((unsigned)(q2 - r0) < (unsigned)nr && q2 <= h) n_ok += (q2 == 0 || 2 * q2 == n) ? 1 : 2;
        const double ue = (double)sd.u_m[bl] / wl / sd.du;
        const double ve = (double)sd.v_m[bl] / wl / sd.dv;
        int kx = (int)round((ue - (double)up) * 1023.0);
        int ky = (int)round((ve - (double)vp) * 1023.0);
        kx = (kx + 1534) % 1023;
        ky = (ky + 1534) % 1023;
        const double* xk = sd.kb_table + (int64_t)kx * K;
        const double* yk = sd.kb_table + (int64_t)ky * K;
        for (int sgn = 1; sgn >= -1; sgn -= 2) {
            for (int c = 0; c < K; c++) {
                int x;
                if (!kb_cell(up - mid + c, sgn, n, x) || x > h) continue;
                if ((unsigned)(x - r0) >= (unsigned)nr) {
                    if constexpr (kWin) *outside = true;
                    continue;
                }
                const double wx = xk[c];
                for (int j = 0; j < K; j++) {
                    int y;
                    if (!kb_cell(vp - mid + j, sgn, n, y)) continue;
                    float2* d = at(x - r0, y);
                    if (!d) {
                        *outside = true;
                        continue;
                    }
                    const double w = yk[j] * wx;
                    atomicAdd(&d->x, (float)((double)v.x * w));
                    atomicAdd(&d->y, (float)((double)sgn * (double)v.y * w));
                }
            }
        }
    }
    return n_ok;
}

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
// Largest x over the workgroup, returned to every thread (scratch: one int
// per wavefront).
__device__ __forceinline__ int block_max_int(int x, int* scratch) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = x;
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) x = max(x, scratch[w]);
    __syncthreads();  // scratch may be reused
    return x;
}

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
