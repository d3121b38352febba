// Batched visibilities -> images kernels for power-of-two grid sides, built on
// the register-resident FFT (blink_fft_reg.hpp).
//
// The plan is transposed (SetDesc::uv_swap): its rows are grid columns u and an
// entry's position is v, so the first pass transforms along v and the second
// along u, the image's fast axis, and writes whole image rows.
//
//   k_grid_rows_reg : (tile of plan rows, image) -> plan entries gathered (pol,
//                     phase correction, NaN/MAX_VIS) and accumulated in LDS,
//                     each non-empty plan row transformed and written to the
//                     column-blocked row-pass buffer (whole lines);
//   k_cols_reg      : (block of image rows, image) -> the block's sequence pairs
//                     staged through LDS from one contiguous run of the buffer
//                     (plan rows with no entries read as zero), transformed,
//                     scaled by (float)(1/sum counts) and stored fft-shifted into
//                     images[img][y][x], two whole image rows per pair.
//
// Hermitian half: every visibility is gridded at (u, v) and conjugated at
// (-u, -v), so the grid satisfies G[-v][-u] = conj(G[v][u]).  With p the plan
// row coordinate (the grid column u) and q the other (v), the
// first-pass transforms R_p(s) = sum_q G w^(qs) satisfy R_{-p}(s) =
// conj(R_p(s)).  Only plan rows p = 0 .. N/2 are gridded and transformed; the
// second pass reads R_p(s) for p <= N/2 and uses conj(R_{N-p}(s)) for the
// others.  The count sum stays exact: an entry in a plan row 0 < p < N/2 stands
// for itself and for its conjugate partner in row N - p, so it counts twice.
#include "blink_fft_reg.hpp"
#include "blink_gather.hpp"
#include "blink_kernels.hpp"

#include <algorithm>

namespace blink {

// Row stride (float2) of an LDS tile of `rows` rows: at least the padded row,
// and congruent to 32 / rows modulo 32, so that a transposed walk (rows
// fastest, then columns) puts the 32 lanes of a half-wave on 32 distinct
// 8-byte bank pairs.
__host__ __device__ constexpr int tile_stride(int n, int rows) {
    const int want = 32 / rows;
    const int p = padded_len(n);
    return p + ((want - p % 32) + 32) % 32;
}

// Row-pass buffer layout of the register kernels (N = 128 .. 2048): plan
// rows v = 0 .. N/2 of each image (grid columns u, with the transposed plan),
// cut into blocks of kInterBlock columns, each block stored row after row:
// element (v, x) of image li at
//   li (N/2 + 1) N + (x / W) (N/2 + 1) W + v W + x % W,   W = kInterBlock.
// A transformed row is written as N / W segments of W * 8 = 256 bytes (whole
// lines), and a column block of up to W columns reads one contiguous run.
constexpr int kInterBlock = kInterBlockCols;
__device__ __forceinline__ int64_t inter_index(int n, int64_t li, int v, int x) {
    const int64_t h1 = n / 2 + 1;
    return li * h1 * n + (int64_t)(x / kInterBlock) * h1 * kInterBlock + (int64_t)v * kInterBlock +
           x % kInterBlock;
}

// Lanes per row: 16 values per lane up to N = 1024, 32 at N = 2048.
template <int N>
struct BatchLanes;
template <> struct BatchLanes<128> { static constexpr int L = 8; };
template <> struct BatchLanes<256> { static constexpr int L = 16; };
template <> struct BatchLanes<512> { static constexpr int L = 32; };
template <> struct BatchLanes<1024> { static constexpr int L = 64; };
template <> struct BatchLanes<2048> { static constexpr int L = 64; };

// (N, RB): lanes per row L and RB rows per row tile (or columns per column
// block); one team of L lanes per row, RB * L threads per workgroup.
template <int N, int RB_>
struct BatchShape {
    static constexpr int L = BatchLanes<N>::L;
    static constexpr int RB = RB_;
};

// LDS of one workgroup: RB tile rows (row-tile rows, or column teams) and a
// copy of the twiddle table; with uniform weighting, a count per tile cell.
template <int N, int RB, bool kUniform = false>
constexpr size_t batch_lds_bytes() {
    return sizeof(float2) * ((size_t)RB * tile_stride(N, RB) + N) +
           (kUniform ? sizeof(unsigned) * (size_t)RB * N : 0);
}

// Threads of a row-tile workgroup: the RB teams of L lanes that run the row
// FFTs, and at least 256 for the gather of plan entries, which is latency
// bound: small sides with many baselines (128^2: 8 lanes per row) would
// otherwise leave each workgroup a few dozen threads for thousands of entries.
template <int N, int RB>
constexpr int rows_reg_threads() {
    return RB * BatchLanes<N>::L > 256 ? RB * BatchLanes<N>::L : 256;
}

// kUniform: the "U" weighting of src/pacer_imager.cpp:421-428 -- each cell is
// divided by its count before the transform, so an empty cell becomes
// 0 / 0 = NaN, as in the reference (the transform then spreads it over the
// row, and the second pass over the image).
template <int N, int RB_, bool kMulti, bool kUniform>
__global__ void __launch_bounds__((rows_reg_threads<N, RB_>()))
k_grid_rows_reg(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
                int64_t gi0, float2* __restrict__ inter, const float2* __restrict__ tw) {
    constexpr int L = BatchLanes<N>::L, RB = RB_, TS = tile_stride(N, RB_);
    constexpr int NT = rows_reg_threads<N, RB_>();
    using F = RegFft<N, L>;
    extern __shared__ float4 rows_reg_lds4[];
    float2* tile = (float2*)rows_reg_lds4;  // [RB][TS]
    __shared__ int rs[RB + 1];   // first entry of each tile row, then the end
    __shared__ int rrow[RB];     // grid row of each tile row
    __shared__ int red[NT / 64 + 1];

    // A tile is RB consecutive NON-EMPTY rows of the plan's row list, so no
    // workgroup carries rows without entries; tiles past the last non-empty
    // row find an empty entry range and leave.
    constexpr int H = N / 2;
    constexpr int n_rb = (H + 1 + RB - 1) / RB;  // enough tiles for rows 0 .. N/2
    const unsigned nwg = gridDim.x;
    unsigned bid = blockIdx.x;
    if ((nwg & 7u) == 0) bid = (bid & 7u) * (nwg >> 3) + (bid >> 3);  // XCD-contiguous images
    const int tile_id = (int)(bid % n_rb);
    const int li = (int)(bid / n_rb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const GatherArgs& g = sd.g;
    const int img = (int)(gi - sd.img_base);  // image within its set
    const int ch = img % g.n_ch;
    const int s0 = tile_id * RB;  // first list slot of this tile
    const int tid = threadIdx.x;
    const int* rl = sd.row_list + (int64_t)ch * row_list_len(N);
    const int* rst = rl + H + 1;  // first entries per slot, then the end
    const int e0 = rst[s0], e1 = rst[min(s0 + RB, H + 1)];  // uniform
    if (e0 == e1) return;                                  // past the last non-empty row

    // The first batch of gathers, the twiddle table and the row starts are
    // requested before the LDS set-up, so their latency overlaps the zeroing
    // of the tile; only LDS work waits at the barrier.
    constexpr int U = 4;
    const uint32_t* ent = sd.entries + (int64_t)ch * 2 * g.n_bl;
    const float2* rec = (const float2*)g.vis + 4 * ((int64_t)img * g.n_bl);
    const float2* cimg = sd.compact ? sd.compact + (int64_t)img * g.n_bl : nullptr;  // uniform
    GatherBatch<U> gb;
    gb.load_entries(ent, e0 + tid, e0, e1, NT);
    constexpr int TWK = (N + NT - 1) / NT;
    float2 twv[TWK];
#pragma unroll
    for (int k = 0; k < TWK; k++)
        if (tid + k * NT < N) twv[k] = tw[tid + k * NT];
    const int rsv = tid <= RB ? rst[min(s0 + tid, H + 1)] : 0;
    const int rrv = tid < RB ? rl[min(s0 + tid, H)] : 0;
    // the gather's copies, uniform: row starts of the tile, and a bit per tile
    // row whose plan row is 0 or N/2 (entries there count once, else twice)
    int rsr[RB + 1];
    unsigned edge = 0;
#pragma unroll
    for (int q = 0; q <= RB; q++) rsr[q] = rst[min(s0 + q, H + 1)];
#pragma unroll
    for (int q = 0; q < RB; q++) {
        const int pr = rl[min(s0 + q, H)];
        if (pr == 0 || pr == H) edge |= 1u << q;
    }
    const double nu = g.freq_hz[ch];
    if (cimg) gb.load_compact(cimg);
    else gb.load_records(g, rec);
    float2* twl = tile + RB * TS;  // twiddle table copy [N]
#pragma unroll
    for (int k = 0; k < TWK; k++)
        if (tid + k * NT < N) twl[tid + k * NT] = twv[k];
    if (tid <= RB) rs[tid] = rsv;
    if (tid < RB) rrow[tid] = rrv;
    for (int i = tid; i < RB * TS; i += NT) tile[i] = make_float2(0.f, 0.f);
    unsigned* cnt = (unsigned*)(twl + N);  // [RB][N] cell counts (kUniform)
    if constexpr (kUniform)
        for (int i = tid; i < RB * N; i += NT) cnt[i] = 0u;
    __syncthreads();
    auto sink = [&](int e, int u_idx, float2 v) {
        // tile rows are consecutive non-empty plan rows (rows past the last one
        // start at the end): the row of entry e counts the later starts <= e
        int r = 0;
#pragma unroll
        for (int k = 1; k < RB; k++)
            if (rsr[k] <= e) r = k;
        float2* c = &tile[r * TS + pad_index(u_idx)];
        atomicAdd(&c->x, v.x);
        atomicAdd(&c->y, v.y);
        if constexpr (kUniform) atomicAdd(&cnt[r * N + u_idx], 1u);
        return ((edge >> r) & 1u) ? 1 : 2;
    };
    int n_ok = cimg ? gb.consume_compact(sink) : gb.consume(g, nu, sink);
    for (int base = e0 + tid + NT * U; base < e1; base += NT * U) {
        GatherBatch<U> gm;
        gm.load_entries(ent, base, e0, e1, NT);
        if (cimg) {
            gm.load_compact(cimg);
            n_ok += gm.consume_compact(sink);
        } else {
            gm.load_records(g, rec);
            n_ok += gm.consume(g, nu, sink);
        }
    }
    const int tot = block_sum_int(n_ok, red);
    if (tid == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    if constexpr (kUniform) {
        for (int i = tid; i < RB * N; i += NT) {
            float2* c = &tile[(i / N) * TS + pad_index(i % N)];
            const float k = (float)cnt[i];
            *c = make_float2(c->x / k, c->y / k);
        }
        __syncthreads();
    }

    // each team transforms its row and stores it from registers, row-major
    // (consecutive lanes write consecutive elements: whole 128-byte lines);
    // rows without entries are not written
    const int team = tid / L, lane = tid % L;
    float2* row = tile + team * TS;
    if (team < RB && rs[team + 1] != rs[team]) {
        float2 v[F::V];
        F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
        team_sync<false>();
        F::template run<1>(v, row, lane, TwTable{twl});
        const int vrow = rrow[team];
        if (sd.cached_inter)
            F::store(v, lane, [&](int i, float2 x) { inter[inter_index(N, li, vrow, i)] = x; });
        else
            F::store(v, lane, [&](int i, float2 x) { store_stream(&inter[inter_index(N, li, vrow, i)], x); });
    }
}

// Direct row pass for dense sets whose images each have a plan of their own
// (one timestep per channel): no plan is built.  One workgroup per image
// holds its whole half plane (plan rows 0 .. N/2) in LDS, walks every
// baseline of the image in order (direct_walk: flags, u, v and the records
// coalesced, each record line read once), computes the cells as the plan's
// count pass does and accumulates each visibility and its conjugate; then it
// transforms and stores rows 0 .. the highest row holding a cell, and writes
// the image's row starts so that the second pass reads the rows above as
// empty (the cells lie within |cell| <= round(N / (osf wavelength)), about
// 0.3 N at osf 2).  Built where the half plane fits in LDS with
// room for a second workgroup per CU (N = 128); tiles of part of the rows
// would each walk all baselines, and measured slower than the plan path
// (768 channels x 256^2, 32896 baselines: two tiles per image 0.70 against
// 0.66 ms per step).  The count sums start at zero (k_direct_prep).
template <int N, int NT, bool kMulti>
__global__ void __launch_bounds__(NT)
k_grid_rows_direct(SetDesc first, const SetDesc* __restrict__ sets,
                   const int* __restrict__ img_set, int64_t gi0, float2* __restrict__ inter,
                   const float2* __restrict__ tw) {
    constexpr int L = BatchLanes<N>::L, TS = padded_len(N), H = N / 2, TEAMS = NT / L;
    using F = RegFft<N, L>;
    extern __shared__ float4 rows_direct_lds4[];
    float2* tile = (float2*)rows_direct_lds4;  // [H + 1][TS]
    float2* twl = tile + (H + 1) * TS;         // twiddle table copy [N]
    __shared__ int red[NT / 64 + 1];

    const int li = (int)xcd_order(blockIdx.x, gridDim.x);  // XCD-contiguous images
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const int img = (int)(gi - sd.img_base);  // image within its set
    const int ch = img % sd.g.n_ch;
    const int tid = threadIdx.x;
    for (int i = tid; i < N; i += NT) twl[i] = tw[i];
    for (int i = tid; i < (H + 1) * TS; i += NT) tile[i] = make_float2(0.f, 0.f);
    __syncthreads();
    int top = -1;
    const int n_ok = direct_walk<2>(
        sd, img, ch, N, 0, H + 1, tid, NT,
        [&](int r, int c) { return &tile[r * TS + pad_index(c)]; }, nullptr, &top);
    const int n_rows = block_max_int(top, red) + 1;
    const int tot = block_sum_int(n_ok, red);
    if (tid == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();

    // rows 0 .. n_rows - 1, one team of L lanes per row
    const int team = tid / L, lane = tid % L;
    for (int r = team; r < n_rows; r += TEAMS) {
        float2* row = tile + r * TS;
        float2 v[F::V];
        F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
        team_sync<false>();
        F::template run<1>(v, row, lane, TwTable{twl});
        if (sd.cached_inter)
            F::store(v, lane, [&](int i, float2 x) { inter[inter_index(N, li, r, i)] = x; });
        else
            F::store(v, lane, [&](int i, float2 x) { store_stream(&inter[inter_index(N, li, r, i)], x); });
    }
    int* rsg = sd.row_start + (int64_t)ch * (N + 1);
    for (int i = tid; i <= N; i += NT) rsg[i] = min(i, n_rows);
}

// Largest of a and b over the workgroup (every thread gets both); NaN
// operands are ignored (fmaxf).
__device__ __forceinline__ void block_max2(float& a, float& b, float* scratch) {
    for (int o = 32; o > 0; o >>= 1) {
        a = fmaxf(a, __shfl_xor(a, o, 64));
        b = fmaxf(b, __shfl_xor(b, o, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    if (lane == 0) scratch[wid] = a, scratch[nw + wid] = b;
    __syncthreads();
    for (int w = 0; w < nw; w++) a = fmaxf(a, scratch[w]), b = fmaxf(b, scratch[nw + w]);
}

// Lanes per row of the windowed direct row pass's transforms: 8 values per
// lane (radix 8, 8, 4: two exchanges), and a team keeps two rows in
// registers (64 lanes, four values each, measured 2 % slower at 768 x 256^2).
constexpr int kWinLanes = 32;

// Transforms local rows 0 .. nr - 1 of an LDS tile of stride s and stores
// them as plan rows r0 .. of image li of the row-pass buffer.  kFull: whole
// rows; else each row holds positions 0 .. e_c and N - e_c .. N - 1 (at
// e_c + 1 .. 2 e_c), the others zero.  Each team of kWinLanes lanes takes two
// rows at a time into registers and uses their LDS (2 s >= padded_len(N)
// elements) as the transform's scratch.
template <int N, int NT, bool kFull>
__device__ __forceinline__ void win_transform_rows(float2* tile, int r0, int nr, int s, int e_c,
                                                   int li, float2* __restrict__ inter,
                                                   const float2* twl, bool cached) {
    constexpr int L = kWinLanes, TEAMS = NT / L;
    using F = RegFft<N, L>;
    const int team = (int)threadIdx.x / L, lane = (int)threadIdx.x % L;
    const int c_shift = N - 2 * e_c - 1;
    auto at = [&](const float2* row, int i) {
        if constexpr (kFull) return row[i];
        const int k = i <= e_c ? i : (i >= N - e_c ? i - c_shift : -1);
        return k >= 0 ? row[k] : make_float2(0.f, 0.f);
    };
    auto put = [&](int r, const float2(&v)[F::V]) {
        if (cached)
            F::store(v, lane, [&](int i, float2 x) { inter[inter_index(N, li, r, i)] = x; });
        else
            F::store(v, lane,
                     [&](int i, float2 x) { store_stream(&inter[inter_index(N, li, r, i)], x); });
    };
    for (int p = team; 2 * p < nr; p += TEAMS) {
        float2* base = tile + 2 * p * s;
        const bool two = 2 * p + 1 < nr;
        float2 v0[F::V], v1[F::V];
        F::load(v0, lane, [&](int i) { return at(base, i); });
        if (two) F::load(v1, lane, [&](int i) { return at(base + s, i); });
        team_sync<false>();
        F::template run<1>(v0, base, lane, TwTable{twl});
        put(r0 + 2 * p, v0);
        if (two) {
            team_sync<false>();
            F::template run<1>(v1, base, lane, TwTable{twl});
            put(r0 + 2 * p + 1, v1);
        }
    }
}

// Direct row pass at sides whose half plane does not fit in LDS (N = 256).
// A baseline's cell is round((x / wavelength) / cell size) per coordinate,
// and |cell| grows with |x|, so the cells of an image lie within E_u, E_v of
// the origin, the cells of the set's largest unflagged |u| and |v| at the
// image's wavelength (partial maxima from the direct set-up launch).  When
// plan rows 0 .. E_r and positions -E_c .. E_c (a window, E < N/2) fit in
// LDS, the workgroup holds only those, walks every baseline once, transforms
// the window's rows and stores them; the image's row starts (row p holds
// entries when start[p + 1] != start[p]) tell the second pass that rows above
// E_r are empty.  Otherwise (a window too large, or several timesteps per
// channel, whose images share their row starts) the workgroup only sets
// start[N] = -1, which the second pass never reads, and leaves the image to
// k_grid_rows_direct_blocks, launched next.  A cell outside the window (not
// expected: the bound is exact) does the same.  kKb: convolutional gridding
// (kb_walk), the window widened by K/2 on every side for the footprints.
// `cap`: tile elements in LDS.
template <int N, int NT, bool kMulti, bool kKb>
__global__ void __launch_bounds__(NT)
k_grid_rows_direct_win(SetDesc first, const SetDesc* __restrict__ sets,
                       const int* __restrict__ img_set, int64_t gi0, float2* __restrict__ inter,
                       const float2* __restrict__ tw, int cap) {
    constexpr int H = N / 2;
    constexpr int kMinStride = (padded_len(N) + 1) / 2;  // two rows hold the scratch
    extern __shared__ float4 rows_win_lds4[];
    float2* tile = (float2*)rows_win_lds4;  // [cap]
    float2* twl = tile + cap;               // twiddle table copy [N]
    __shared__ int red[NT / 64];
    __shared__ float redm[2 * (NT / 64)];

    const int li = (int)xcd_order(blockIdx.x, gridDim.x);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const int img = (int)(gi - sd.img_base);
    const int ch = img % sd.g.n_ch;
    const int tid = threadIdx.x;
    int* rsg = sd.row_start + (int64_t)ch * (N + 1);

    // the window: largest |u|, |v| of the unflagged baselines, then their cells
    float mu = 0.f, mv = 0.f;
    if (tid < kUvMaxParts) mu = sd.uv_max[2 * tid], mv = sd.uv_max[2 * tid + 1];
    block_max2(mu, mv, redm);
    double er, ec;
    {
#pragma clang fp contract(off)
        const double wl = wavelength_of(sd.g.freq_hz[ch]);
        const double eu = fabs(round(((double)mu / wl) / sd.du));
        const double ev = fabs(round(((double)mv / wl) / sd.dv));
        const double margin = kKb ? (double)(sd.kb_size / 2) : 0.0;
        er = (sd.uv_swap ? eu : ev) + margin;
        ec = (sd.uv_swap ? ev : eu) + margin;
    }
    const bool fits = sd.n_img == sd.g.n_ch && er < H && ec < H &&
                      (((int)er + 2) & ~1) * (int64_t)max(2 * (int)ec + 1, kMinStride) <= cap;
    if (!fits) {  // uniform
        if (tid == 0) rsg[N] = -1;
        return;
    }
    const int e_c = (int)ec, n_win = (int)er + 1, s = max(2 * e_c + 1, kMinStride);
    const int c_shift = N - 2 * e_c - 1;  // positions N - e_c .. N - 1 -> e_c + 1 ..
    for (int i = tid; i < N; i += NT) twl[i] = tw[i];
    for (int i = tid; i < ((n_win + 1) & ~1) * s; i += NT) tile[i] = make_float2(0.f, 0.f);
    __syncthreads();
    bool outside = false;
    auto at = [&](int r, int c) -> float2* {
        const int j = c <= e_c ? c : (c >= N - e_c ? c - c_shift : -1);
        return j >= 0 ? &tile[r * s + j] : nullptr;
    };
    int n_ok;
    if constexpr (kKb)
        n_ok = kb_walk<true>(sd, img, ch, N, 0, n_win, tid, NT, at, &outside);
    else
        n_ok = direct_walk<2, true>(sd, img, ch, N, 0, n_win, tid, NT, at, &outside);
    if (__syncthreads_or(outside ? 1 : 0)) {
        if (tid == 0) rsg[N] = -1;
        return;
    }
    win_transform_rows<N, NT, false>(tile, 0, n_win, s, e_c, li, inter, twl,
                                     sd.cached_inter != 0);
    const int tot = block_sum_int(n_ok, red);
    if (tid == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    // row starts of the image's channel: rows 0 .. n_win - 1 hold entries
    for (int i = tid; i <= N; i += NT) rsg[i] = min(i, n_win);
}

// The windowed direct row pass's other images (start[N] = -1): blocks of rb
// whole rows, each walking every baseline, then every row stored and the
// row starts of rows 0 .. N/2 set (start[N] stays -1, so that the other
// images of a channel with several timesteps see it too).  Workgroups take
// the chunk's images in turn and return at once for the windowed ones.
// (NT = 512: the walk and two-row transforms in one loop fit its registers.)
// kKb: convolutional gridding (kb_walk).
template <int N, int NT, bool kMulti, bool kKb>
__global__ void __launch_bounds__(NT)
k_grid_rows_direct_blocks(SetDesc first, const SetDesc* __restrict__ sets,
                          const int* __restrict__ img_set, int64_t gi0, int n_img,
                          float2* __restrict__ inter, const float2* __restrict__ tw, int cap) {
    constexpr int H = N / 2;
    extern __shared__ float4 rows_blk_lds4[];
    float2* tile = (float2*)rows_blk_lds4;  // [cap]
    float2* twl = tile + cap;               // twiddle table copy [N]
    __shared__ int red[NT / 64];
    const int tid = threadIdx.x;
    const int rb = (cap / N) & ~1;  // rows per block (the launch checks rb >= 2)
    bool have_tw = false;
    for (int li = blockIdx.x; li < n_img; li += gridDim.x) {
        const int64_t gi = gi0 + li;
        const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
        const int img = (int)(gi - sd.img_base);
        const int ch = img % sd.g.n_ch;
        int* rsg = sd.row_start + (int64_t)ch * (N + 1);
        if (rsg[N] >= 0) continue;  // windowed (uniform)
        if (!have_tw) {
            for (int i = tid; i < N; i += NT) twl[i] = tw[i];
            have_tw = true;
        }
        int n_ok = 0;
        for (int r0 = 0; r0 <= H; r0 += rb) {
            const int nr = min(rb, H + 1 - r0);
            __syncthreads();  // the previous block's transforms are done with the tile
            for (int i = tid; i < ((nr + 1) & ~1) * N; i += NT) tile[i] = make_float2(0.f, 0.f);
            __syncthreads();
            auto at = [&](int r, int c) { return &tile[r * N + c]; };
            if constexpr (kKb)
                n_ok += kb_walk<false>(sd, img, ch, N, r0, nr, tid, NT, at, nullptr);
            else
                n_ok += direct_walk<2>(sd, img, ch, N, r0, nr, tid, NT, at);
            __syncthreads();
            win_transform_rows<N, NT, true>(tile, r0, nr, N, 0, li, inter, twl,
                                            sd.cached_inter != 0);
        }
        const int tot = block_sum_int(n_ok, red);
        if (tid == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
        for (int i = tid; i < N; i += NT) rsg[i] = min(i, H + 1);
    }
}

// Set-up of the direct row pass, per set of the group: the count sums start
// at zero, and the phase coefficients of the baselines are written once (the
// plan's count pass does both for the planned sets); for the windowed direct
// row pass, each block's largest |u| and |v| over the unflagged baselines it
// visits (slot b of uv_max; slots past the grid hold zeros).
__global__ void __launch_bounds__(256)
k_direct_prep(const SetDesc* __restrict__ sets) {
    __shared__ float redm[2 * (256 / 64)];
    const SetDesc& sd = sets[blockIdx.y];
    const int n = max(sd.n_img, sd.g.n_bl);
    float mu = 0.f, mv = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < sd.n_img) sd.isums[i] = 0u;
        if (i < sd.g.n_bl && sd.g.corrections)
            sd.coef_out[i] = phase_coefficient(sd.g.corrections, sd.w_m, sd.cable_m, (unsigned)i);
        if (sd.uv_max && i < sd.g.n_bl && !sd.flags[i]) {
            mu = fmaxf(mu, fabsf(sd.u_m[i]));
            mv = fmaxf(mv, fabsf(sd.v_m[i]));
        }
    }
    if (!sd.uv_max) return;  // uniform per block
    block_max2(mu, mv, redm);
    if (threadIdx.x == 0)
        for (int j = blockIdx.x; j < kUvMaxParts; j += gridDim.x) {
            sd.uv_max[2 * j] = j == (int)blockIdx.x ? mu : 0.f;
            sd.uv_max[2 * j + 1] = j == (int)blockIdx.x ? mv : 0.f;
        }
}

// Second pass.  The gridded plane is Hermitian by construction (every
// visibility lands at (u, v) and its conjugate at (-u, -v)), so every image
// is real and each sequence c_y[p] = R_p(y) over the plan rows obeys
// c_y[N-p] = conj(c_y[p]).  One complex FFT therefore transforms two at once:
//   z[p] = c_y[p] + i c_{y+1}[p]  ->  Z[x] = img[y][x] + i img[y+1][x],
// with both image rows real; the imaginary parts of the output are 0.  The
// plan is transposed (p = u), so the buffer's column index is the image row y.
// CB image rows per workgroup, CB / 2 teams of L lanes.
template <int N, int CB_, bool kMulti>
__global__ void __launch_bounds__(CB_ / 2 * BatchLanes<N>::L)
k_cols_reg(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
           int64_t gi0, const float2* __restrict__ inter, const float2* __restrict__ tw) {
    constexpr int L = BatchLanes<N>::L, CB = CB_, TEAMS = CB_ / 2;
    constexpr int TS = tile_stride(N, TEAMS);
    constexpr int NT = TEAMS * L;
    using F = RegFft<N, L>;
    extern __shared__ float4 cols_reg_lds4[];
    float2* tile = (float2*)cols_reg_lds4;  // [TEAMS][TS]
    constexpr int H = N / 2;
    __shared__ unsigned char nonempty[H + 1];

    constexpr int n_cb = N / CB;
    const int cb = blockIdx.x % n_cb;
    const int li = blockIdx.x / n_cb;
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    if (!sd.images) return;  // a set imaged for its average only (uniform per workgroup)
    const int img = (int)(gi - sd.img_base);  // image within its set
    const int tid = threadIdx.x;
    const uint32_t count = sd.isums[img];
    if (cb == 0 && tid == 0 && sd.count_sums) sd.count_sums[img] = (double)count;
    const int* rsg = sd.row_start + (int64_t)(img % sd.g.n_ch) * (N + 1);
    // plan rows without entries hold zeros, or under uniform weighting empty
    // cells, 0 / 0 = NaN
    const float empty = sd.uniform ? __builtin_nanf("") : 0.f;
    float2* twl = tile + TEAMS * TS;  // twiddle table copy [N]
    for (int i = tid; i < N; i += NT) twl[i] = tw[i];
    for (int i = tid; i <= H; i += NT) nonempty[i] = sd.all_rows || rsg[i + 1] != rsg[i];
    __syncthreads();

    // stage z_t[p] and z_t[N - p] from buffer columns c0 + 2t, c0 + 2t + 1 of
    // plan row p: consecutive threads read consecutive 16-byte column pairs,
    // and with the column-blocked buffer consecutive rows of a block follow
    // each other, so the block's rows are one contiguous run; rows without plan
    // entries are zero (all of a thread's loads are issued before its LDS writes)
    const int c0 = cb * CB;
    constexpr int STAGE = (H + 1) * TEAMS, ITER = (STAGE + NT - 1) / NT;
    float4 p[ITER];
#pragma unroll
    for (int k = 0; k < ITER; k++) {
        const int i = tid + k * NT;
        const int vr = i / TEAMS, t = i % TEAMS;
        p[k] = make_float4(empty, empty, empty, empty);
        if (i < STAGE && nonempty[vr])
            p[k] = *(const float4*)&inter[inter_index(N, li, vr, c0 + 2 * t)];
    }
#pragma unroll
    for (int k = 0; k < ITER; k++) {
        const int i = tid + k * NT;
        if (i >= STAGE) break;
        const int vr = i / TEAMS, t = i % TEAMS;
        // a = c_y[p] = (p.x, p.y), b = c_{y+1}[p] = (p.z, p.w); row N - p holds conjugates
        float2* zr = tile + t * TS;
        zr[pad_index(vr)] = make_float2(p[k].x - p[k].w, p[k].y + p[k].z);
        if (vr > 0 && vr < H)
            zr[pad_index(N - vr)] = make_float2(p[k].x + p[k].w, p[k].z - p[k].y);
    }
    __syncthreads();

    const int team = tid / L, lane = tid % L;
    float2* row = tile + team * TS;
    float2 v[F::V];
    F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
    team_sync<false>();
    F::template run<1>(v, row, lane, TwTable{twl});
    const float s = (float)(1.0 / (double)count);
#pragma unroll
    for (int k = 0; k < F::V; k++) v[k] = cscale(v[k], s);
    team_sync<false>();
    F::store(v, lane, [&](int i, float2 x) { row[pad_index(i)] = x; });
    __syncthreads();
    // shifted store: images[img][shift(c0 + 2 t) + {0, 1}][shift(x) + {0, 1}]
    // = (Re Z_t[x], 0), (Re Z_t[x + 1], 0) and the same of Im Z_t in the next
    // row -- consecutive threads on consecutive column pairs, so each wave
    // writes 1 KiB runs of one image row (even x: shift(x + 1) = shift(x) + 1)
    float2* out = sd.images + (int64_t)img * N * N;
    for (int i = tid; i < TEAMS * H; i += NT) {
        const int t = i / H, ox = 2 * (i % H);
        const int x = ox < H ? ox + H : ox - H;
        const int y = c0 + 2 * t;
        const int oy = y < H ? y + H : y - H;
        const float2 z0 = tile[t * TS + pad_index(x)], z1 = tile[t * TS + pad_index(x + 1)];
        float4* o = (float4*)&out[(int64_t)oy * N + ox];
        store_stream(o, make_float4(z0.x, 0.f, z1.x, 0.f));
        store_stream(o + N / 2, make_float4(z0.y, 0.f, z1.y, 0.f));
    }
}

// Visibilities of a set's images compacted for the row tiles: one float2
// per (image, baseline), after the pol combination and the phase rotation, as
// GatherBatch::consume forms them, or NaN where consume would skip the value
// (NaN or |re|, |im| >= MAX_VIS).  A streaming pass over the records, so that
// the row tiles' random gathers read 8 bytes per baseline instead of 32-byte
// records, from a quarter of the cache lines, and do no FP64 work.
__device__ __forceinline__ void vis_compact_one(const SetDesc& sd, int64_t img, int bl) {
    const GatherArgs& g = sd.g;
    const int64_t i = img * g.n_bl + bl;
    const float2* rec = (const float2*)g.vis + 4 * i;  // record of (img, bl)
    const int p0 = g.pol == 1 ? 3 : (g.pol == 3 ? 1 : 0);
    const int p1 = g.pol == 3 ? 2 : 3;
    const float2 x0 = rec[p0];
    const float2 x1 = g.pol >= 2 ? rec[p1] : make_float2(0.f, 0.f);
    float2 v = select_pol2(x0, x1, g.pol);
    if (g.corrections) {
        // the coefficient the plan writes for the row tiles, formed here, so
        // that the compaction need not wait for the plan
        const double cf = phase_coefficient(g.corrections, sd.w_m, sd.cable_m, (unsigned)bl);
        const double nu = g.freq_hz[img % g.n_ch];
        v = cmul(v, unit_phasor(((cf * nu) / kVelLight)));
    }
    if (!vis_ok(v)) v = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    sd.compact[i] = v;
}

// One set: item i = (image, baseline) of its n_img x n_bl.
__global__ void __launch_bounds__(256)
k_vis_compact(SetDesc sd, int64_t n_items) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_items) return;
    vis_compact_one(sd, i / sd.g.n_bl, (int)(i % sd.g.n_bl));
}

// Several sets of one side: bpi workgroups per group image gi, each 256
// baselines of it; images of sets without a compact buffer are skipped.
__global__ void __launch_bounds__(256)
k_vis_compact_sets(const SetDesc* __restrict__ sets, const int* __restrict__ img_set, int bpi) {
    const int64_t gi = blockIdx.x / bpi;
    const SetDesc sd = sets[img_set[gi]];
    const int bl = (int)(blockIdx.x % bpi) * 256 + (int)threadIdx.x;
    if (!sd.compact || bl >= sd.g.n_bl) return;
    vis_compact_one(sd, gi - sd.img_base, bl);
}

hipError_t launch_vis_compact(const SetDesc& sd, int64_t n_img, hipStream_t stream) {
    const int64_t n_items = n_img * sd.g.n_bl;
    if (n_items == 0 || !sd.compact) return hipSuccess;
    const int64_t blocks = (n_items + 255) / 256;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_vis_compact, dim3((unsigned)blocks), dim3(256), 0, stream, sd, n_items);
    return hipGetLastError();
}

hipError_t launch_vis_compact_sets(const SetDesc* sets, const int* img_set, int64_t n_img,
                                   int max_bl, hipStream_t stream) {
    if (n_img == 0 || max_bl == 0) return hipSuccess;
    const int bpi = (max_bl + 255) / 256;
    const int64_t blocks = n_img * bpi;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_vis_compact_sets, dim3((unsigned)blocks), dim3(256), 0, stream, sets,
                       img_set, bpi);
    return hipGetLastError();
}

// Mean of a set's images without forming them.  The second pass is linear in
// the row-pass buffer, so the mean of the normalised images is the second
// pass of A = sum_i s_i R_i with the scale 1 / n_img, where R_i is image i's
// row-pass buffer (zero in plan rows without entries, NaN there under uniform
// weighting) and s_i = (float)(1 / sum counts of image i), the scale the
// second pass would apply (so an image with no gridded visibility makes the
// mean NaN, as 0 * (1/0) makes its image).  k_average_rows adds the set's
// images of one chunk into A, in image order (deterministic); the mean is then
// one second-pass launch over A with the set's descriptor pointed at the
// average (meta[0] = n_img as its count, meta[1 ..] an all-non-empty row start).
// kBlocked: the column-blocked buffer of the register kernels, else row-major.
template <bool kBlocked>
__global__ void __launch_bounds__(256)
k_average_rows(SetDesc sd, int64_t img0, int n_img, int64_t li0, int n,
               const float4* __restrict__ inter, float4* __restrict__ acc, int accumulate,
               unsigned* __restrict__ meta, int n_img_total) {
    const int h1 = n / 2 + 1;
    const int64_t per = (int64_t)h1 * n / 2;  // float4 per image
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        if (!accumulate) {
            for (int i = tid; i <= n; i += 256) meta[1 + i] = (unsigned)i;
            if (tid == 0) meta[0] = (unsigned)n_img_total;
        }
        // the count sums of images the second pass does not form
        if (!sd.images && sd.count_sums)
            for (int k = tid; k < n_img; k += 256)
                sd.count_sums[img0 + k] = (double)sd.isums[img0 + k];
    }
    const int64_t j = (int64_t)blockIdx.x * 256 + tid;
    if (j >= per) return;
    const int e = (int)(2 * j);  // first of the two float2 (one plan row)
    const int p = kBlocked ? (e % (h1 * kInterBlock)) / kInterBlock : e / n;
    const float z = sd.uniform ? __builtin_nanf("") : 0.f;
    float4 sum = accumulate ? acc[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < n_img; k++) {
        const int64_t img = img0 + k;
        const int* rsg = sd.row_start + (img % sd.g.n_ch) * (int64_t)(n + 1);
        const float sc = (float)(1.0 / (double)sd.isums[img]);
        const float4 v = (sd.all_rows || rsg[p + 1] != rsg[p]) ? inter[(li0 + k) * per + j]
                                                                : make_float4(z, z, z, z);
        sum.x += sc * v.x;
        sum.y += sc * v.y;
        sum.z += sc * v.z;
        sum.w += sc * v.w;
    }
    acc[j] = sum;
}

hipError_t launch_average_rows(const SetDesc& sd, int64_t img0, int n_img, int64_t li0,
                               const float* inter, float* acc, bool accumulate, unsigned* meta,
                               hipStream_t stream) {
    if (n_img <= 0) return hipSuccess;
    const int n = sd.n_pixels;
    const int64_t per = (int64_t)(n / 2 + 1) * n / 2;
    const int64_t blocks = (per + 255) / 256;
    auto k = batch_reg_supported(n) && n <= 2048 ? k_average_rows<true> : k_average_rows<false>;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, stream, sd, img0, n_img, li0, n,
                       (const float4*)inter, (float4*)acc, accumulate ? 1 : 0, meta, sd.n_img);
    return hipGetLastError();
}

// ------------------------------------------------------------- N = 4096
//
// At N = 4096 one row needs 4 wavefronts (256 lanes x 16 values) and 34 KiB
// of LDS, so a workgroup holds one to four rows or column pairs.  The plan is
// transposed as for the smaller sides, and the row-pass buffer is row-major:
// inter[img][p][y], plan rows p = 0 .. N/2.  The row kernel stores each
// transformed row with unit-stride lanes; the second-pass kernel stages its
// CB columns y of plan rows 0 .. N/2 through LDS (16-byte loads, two columns
// per thread), forming z = c_y + i c_{y+1} for rows p and N - p on the way,
// and writes image rows y and y + 1 whole.
// Twiddles come from two 64-entry tables (w(m) = lo[m % 64] hi[m / 64]) kept
// in LDS, instead of a 32 KiB table.

constexpr int kWideLanes = 256;

__host__ __device__ constexpr int wide_tile_stride(int n, int rows) { return tile_stride(n, rows); }

template <int N, int RB, bool kUniform = false>
constexpr size_t wide_rows_lds_bytes() {
    return sizeof(float2) * ((size_t)RB * padded_len(N) + 128) +
           (kUniform ? sizeof(unsigned) * (size_t)RB * N : 0);
}
template <int N, int CB>
constexpr size_t wide_cols_lds_bytes() {
    return sizeof(float2) * ((size_t)(CB / 2) * wide_tile_stride(N, CB / 2) + 128) + N;
}

template <int N, int RB, bool kMulti, bool kUniform>
__global__ void __launch_bounds__(RB * kWideLanes)
k_grid_rows_wide(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
                 int64_t gi0, float2* __restrict__ inter, const float2* __restrict__ tws) {
    constexpr int L = kWideLanes, TS = padded_len(N), NT = RB * L;
    using F = RegFft<N, L>;
    extern __shared__ float4 rows_wide_lds4[];
    float2* tile = (float2*)rows_wide_lds4;  // [RB][TS]
    float2* twl = tile + RB * TS;             // lo[64], hi[64]
    unsigned* cnt = (unsigned*)(twl + 128);   // [RB][N] cell counts (kUniform)
    __shared__ int rs[RB + 1];
    __shared__ int red[NT / 64 + 1];

    constexpr int H = N / 2;
    constexpr int n_rb = (H + 1 + RB - 1) / RB;
    const unsigned nwg = gridDim.x;
    unsigned bid = blockIdx.x;
    if ((nwg & 7u) == 0) bid = (bid & 7u) * (nwg >> 3) + (bid >> 3);
    const int tile_id = (int)(bid % n_rb);
    const int li = (int)(bid / n_rb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const GatherArgs& g = sd.g;
    const int img = (int)(gi - sd.img_base);
    const int ch = img % g.n_ch;
    const int r0 = tile_id * RB;
    const int nr = min(RB, H + 1 - r0);
    const int tid = threadIdx.x;
    const int* rsg = sd.row_start + (int64_t)ch * (N + 1);
    if (tid <= nr) rs[tid] = rsg[r0 + tid];
    if (tid > nr && tid <= RB) rs[tid] = rsg[r0 + nr];
    __syncthreads();
    if (rs[RB] == rs[0]) return;  // uniform: nothing lands in this tile

    if (tid < 128) twl[tid] = tws[tid];
    for (int i = tid; i < RB * TS; i += NT) tile[i] = make_float2(0.f, 0.f);
    if constexpr (kUniform)
        for (int i = tid; i < RB * N; i += NT) cnt[i] = 0u;
    __syncthreads();
    const uint32_t* ent = sd.entries + (int64_t)ch * 2 * g.n_bl;
    const int n_ok = gather_range<4>(g, img, ch, ent, rs[0], rs[RB], tid, NT,
                                     [&](int e, int u_idx, float2 v) {
        int r = 0;
        while (r + 1 < RB && rs[r + 1] <= e) r++;
        float2* c = &tile[r * TS + pad_index(u_idx)];
        atomicAdd(&c->x, v.x);
        atomicAdd(&c->y, v.y);
        if constexpr (kUniform) atomicAdd(&cnt[r * N + u_idx], 1u);
        const int row = r0 + r;
        return (row == 0 || row == H) ? 1 : 2;
    });
    const int tot = block_sum_int(n_ok, red);
    if (tid == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    if constexpr (kUniform) {  // grid / count per cell, 0 / 0 = NaN in empty cells
        for (int i = tid; i < RB * N; i += NT) {
            float2* c = &tile[(i / N) * TS + pad_index(i % N)];
            const float k = (float)cnt[i];
            *c = make_float2(c->x / k, c->y / k);
        }
        __syncthreads();
    }

    // every team transforms its row (the exchanges synchronise the workgroup);
    // rows with entries are stored
    const int team = tid / L, lane = tid % L;
    float2* row = tile + team * TS;
    float2 v[F::V];
    F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
    __syncthreads();
    F::template run<1>(v, row, lane, TwSplit64{twl, twl + 64});
    if (team < nr && rs[team + 1] != rs[team]) {
        float2* dst = inter + ((int64_t)li * (H + 1) + r0 + team) * N;
        if (sd.cached_inter)
            F::store(v, lane, [&](int i, float2 x) { dst[i] = x; });
        else
            F::store(v, lane, [&](int i, float2 x) { store_stream(&dst[i], x); });
    }
}

template <int N, int CB, bool kMulti>
__global__ void __launch_bounds__(CB / 2 * kWideLanes)
k_cols_wide(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
            int64_t gi0, const float2* __restrict__ inter, const float2* __restrict__ tws) {
    constexpr int L = kWideLanes, TEAMS = CB / 2, NT = TEAMS * L;
    constexpr int TS = wide_tile_stride(N, TEAMS);
    constexpr int H = N / 2;
    using F = RegFft<N, L>;
    extern __shared__ float4 cols_wide_lds4[];
    float2* tile = (float2*)cols_wide_lds4;   // [TEAMS][TS]
    float2* twl = tile + TEAMS * TS;           // lo[64], hi[64]
    unsigned char* nonempty = (unsigned char*)(twl + 128);  // [N / 2 + 1]

    constexpr int n_cb = N / CB;
    // XCD-contiguous order: neighbouring column blocks share the cache lines
    // of every output row
    const unsigned nwg = gridDim.x;
    unsigned bid = blockIdx.x;
    if ((nwg & 7u) == 0) bid = (bid & 7u) * (nwg >> 3) + (bid >> 3);
    const int cb = (int)(bid % n_cb);
    const int li = (int)(bid / n_cb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    if (!sd.images) return;  // a set imaged for its average only (uniform per workgroup)
    const int img = (int)(gi - sd.img_base);
    const int tid = threadIdx.x;
    const uint32_t count = sd.isums[img];
    if (cb == 0 && tid == 0 && sd.count_sums) sd.count_sums[img] = (double)count;
    const int* rsg = sd.row_start + (int64_t)(img % sd.g.n_ch) * (N + 1);
    const float empty = sd.uniform ? __builtin_nanf("") : 0.f;  // rows without entries
    for (int i = tid; i <= H; i += NT) nonempty[i] = sd.all_rows || rsg[i + 1] != rsg[i];
    if (tid < 128) twl[tid] = tws[tid];
    __syncthreads();

    // stage: z_t[v] and z_t[N - v] from inter[li][v][c0 + 2t .. c0 + 2t + 1]
    const int c0 = cb * CB;
    const float2* src = inter + (int64_t)li * (H + 1) * N + c0;
    for (int i = tid; i < (H + 1) * TEAMS; i += NT) {
        const int vr = i / TEAMS, t = i % TEAMS;
        float4 p = make_float4(empty, empty, empty, empty);
        if (nonempty[vr]) p = *(const float4*)&src[(int64_t)vr * N + 2 * t];
        // a = c_x[v] = (p.x, p.y), b = c_{x+1}[v] = (p.z, p.w); row N - v holds conjugates
        float2* zr = tile + t * TS;
        zr[pad_index(vr)] = make_float2(p.x - p.w, p.y + p.z);
        if (vr > 0 && vr < H) zr[pad_index(N - vr)] = make_float2(p.x + p.w, p.z - p.y);
    }
    __syncthreads();

    const int team = tid / L, lane = tid % L;
    float2* row = tile + team * TS;
    float2 v[F::V];
    F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
    __syncthreads();
    F::template run<1>(v, row, lane, TwSplit64{twl, twl + 64});
    const float s = (float)(1.0 / (double)count);
#pragma unroll
    for (int k = 0; k < F::V; k++) v[k] = cscale(v[k], s);
    __syncthreads();
    F::store(v, lane, [&](int i, float2 x) { row[pad_index(i)] = x; });
    __syncthreads();
    // shifted store of whole image rows, as in k_cols_reg (the plan is
    // transposed, so pair t holds image rows c0 + 2t and c0 + 2t + 1)
    float2* out = sd.images + (int64_t)img * N * N;
    for (int i = tid; i < TEAMS * H; i += NT) {
        const int t = i / H, ox = 2 * (i % H);
        const int x = ox < H ? ox + H : ox - H;
        const int y = c0 + 2 * t;
        const int oy = y < H ? y + H : y - H;
        const float2 z0 = tile[t * TS + pad_index(x)], z1 = tile[t * TS + pad_index(x + 1)];
        float4* o = (float4*)&out[(int64_t)oy * N + ox];
        store_stream(o, make_float4(z0.x, 0.f, z1.x, 0.f));
        store_stream(o + N / 2, make_float4(z0.y, 0.f, z1.y, 0.f));
    }
}

template <int N, int RB, bool kUniform>
static hipError_t launch_rows_wide_t(const SetDesc& first, const SetDesc* sets,
                                     const int* img_set, int n_sets, int64_t gi0, int n_img,
                                     float* inter, const float* tws, hipStream_t stream) {
    constexpr int NT = RB * kWideLanes;
    constexpr size_t lds = wide_rows_lds_bytes<N, RB, kUniform>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const int64_t blocks = (int64_t)((N / 2 + 1 + RB - 1) / RB) * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    auto k = n_sets == 1 ? k_grid_rows_wide<N, RB, false, kUniform>
                         : k_grid_rows_wide<N, RB, true, kUniform>;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(NT), lds, stream, first, sets, img_set,
                       gi0, (float2*)inter, (const float2*)tws);
    return hipGetLastError();
}

template <int N, int CB>
static hipError_t launch_cols_wide_t(const SetDesc& first, const SetDesc* sets,
                                     const int* img_set, int n_sets, int64_t gi0, int n_img,
                                     const float* inter, const float* tws, hipStream_t stream) {
    constexpr int NT = CB / 2 * kWideLanes;
    constexpr size_t lds = wide_cols_lds_bytes<N, CB>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const int64_t blocks = (int64_t)(N / CB) * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    auto k = n_sets == 1 ? k_cols_wide<N, CB, false> : k_cols_wide<N, CB, true>;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(NT), lds, stream, first, sets, img_set,
                       gi0, (const float2*)inter, (const float2*)tws);
    return hipGetLastError();
}

template <int N, int RB, bool kUniform>
static hipError_t launch_rows_reg_t(const SetDesc& first, const SetDesc* sets,
                                    const int* img_set, int n_sets, int64_t gi0, int n_img,
                                    float* inter, const float* tw, hipStream_t stream) {
    constexpr int NT = rows_reg_threads<N, RB>();
    constexpr size_t lds = batch_lds_bytes<N, RB, kUniform>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const int64_t blocks = (int64_t)((N / 2 + 1 + RB - 1) / RB) * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    if (n_sets == 1)
        hipLaunchKernelGGL((k_grid_rows_reg<N, RB, false, kUniform>), dim3((unsigned)blocks),
                           dim3(NT), lds, stream, first, sets, img_set, gi0, (float2*)inter,
                           (const float2*)tw);
    else
        hipLaunchKernelGGL((k_grid_rows_reg<N, RB, true, kUniform>), dim3((unsigned)blocks),
                           dim3(NT), lds, stream, first, sets, img_set, gi0, (float2*)inter,
                           (const float2*)tw);
    return hipGetLastError();
}

template <int N>
constexpr size_t direct_lds_bytes() {
    return sizeof(float2) * ((size_t)(N / 2 + 1) * padded_len(N) + N);
}

template <int N, int NT>
static hipError_t launch_rows_direct_t(const SetDesc& first, const SetDesc* sets,
                                       const int* img_set, int n_sets, int64_t gi0, int n_img,
                                       float* inter, const float* tw, hipStream_t stream) {
    constexpr size_t lds = direct_lds_bytes<N>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    auto k = n_sets == 1 ? k_grid_rows_direct<N, NT, false> : k_grid_rows_direct<N, NT, true>;
    hipLaunchKernelGGL(k, dim3((unsigned)n_img), dim3(NT), lds, stream, first, sets, img_set,
                       gi0, (float2*)inter, (const float2*)tw);
    return hipGetLastError();
}

template <int N, int CB>
static hipError_t launch_cols_reg_t(const SetDesc& first, const SetDesc* sets,
                                    const int* img_set, int n_sets, int64_t gi0, int n_img,
                                    const float* inter, const float* tw, hipStream_t stream) {
    constexpr int NT = CB / 2 * BatchLanes<N>::L;
    constexpr size_t lds = batch_lds_bytes<N, CB / 2>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const int64_t blocks = (int64_t)(N / CB) * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    if (n_sets == 1)
        hipLaunchKernelGGL((k_cols_reg<N, CB, false>), dim3((unsigned)blocks), dim3(NT), lds,
                           stream, first, sets, img_set, gi0, (const float2*)inter,
                           (const float2*)tw);
    else
        hipLaunchKernelGGL((k_cols_reg<N, CB, true>), dim3((unsigned)blocks), dim3(NT), lds,
                           stream, first, sets, img_set, gi0, (const float2*)inter,
                           (const float2*)tw);
    return hipGetLastError();
}

// Built variants: plan rows per row tile, and output lines per second-pass
// block (image rows up to N = 2048, image columns at 4096; two per team).
// Small tiles keep few images' visibility records in flight per XCD: dense
// grids gather each image's records at random, and an XCD's 4 MiB L2 holds
// only a few images' records (768 x 256^2, 32896 baselines: rows per tile
// 32, 16, 8, 4 ran 509, 477, 447, 412 us, and 4 rows fetched 2.4x the
// record bytes).
#define BLINK_ROW_VARIANTS(X) \
    X(128, 4) X(128, 8) X(128, 16) X(128, 32) \
    X(256, 1) X(256, 2) X(256, 4) X(256, 8) X(256, 16) X(256, 32) \
    X(512, 2) X(512, 4) X(512, 8) X(512, 16) \
    X(1024, 4) X(1024, 8) X(1024, 16) \
    X(2048, 4) X(2048, 8)
#define BLINK_ROW_WIDE_VARIANTS(X) X(4096, 1) X(4096, 2) X(4096, 4)
// Row tiles of uniform weighting (the smallest built tile per side: the cell
// counts take LDS beside the tile)
#define BLINK_ROW_UNIFORM_VARIANTS(X) X(128, 8) X(256, 4) X(512, 2) X(1024, 4) X(2048, 4)
#define BLINK_ROW_WIDE_UNIFORM_VARIANTS(X) X(4096, 1)
#define BLINK_COL_WIDE_VARIANTS(X) X(4096, 2) X(4096, 4) X(4096, 8)
#define BLINK_COL_VARIANTS(X) \
    X(128, 16) X(128, 32) X(128, 64) \
    X(256, 16) X(256, 32) X(256, 64) \
    X(512, 16) X(512, 32) X(512, 64) \
    X(1024, 8) X(1024, 16) X(1024, 32) \
    X(2048, 4) X(2048, 8) X(2048, 16)

// Register sides of the direct row pass: (side, threads), and the side of
// the windowed direct row pass (k_grid_rows_direct_win) with its threads.
#define BLINK_DIRECT_VARIANTS(X) X(128, 512)
constexpr int kDirectWinSide = 256, kDirectWinThreads = 1024, kDirectBlockThreads = 512;

bool batch_reg_supported(int n) {
    switch (n) {
        case 128: case 256: case 512: case 1024: case 2048: case 4096:
            return true;
        default:
            return false;
    }
}

int64_t batch_inter_bytes(int n) {
    return (int64_t)sizeof(float2) * (n / 2 + 1) * n;  // row-major, rows 0 .. N/2
}

int batch_row_variants(int n, int* tiles, int max_tiles) {
    int k = 0;
#define X(NN, T) \
    if (n == NN && k < max_tiles) tiles[k++] = T;
    BLINK_ROW_VARIANTS(X)
    BLINK_ROW_WIDE_VARIANTS(X)
#undef X
    return k;
}

int batch_col_variants(int n, int* tiles, int max_tiles) {
    int k = 0;
#define X(NN, T) \
    if (n == NN && k < max_tiles) tiles[k++] = T;
    BLINK_COL_VARIANTS(X)
    BLINK_COL_WIDE_VARIANTS(X)
#undef X
    return k;
}

int batch_uniform_row_tile(int n) {
#define X(NN, T) \
    if (n == NN) return T;
    BLINK_ROW_UNIFORM_VARIANTS(X)
    BLINK_ROW_WIDE_UNIFORM_VARIANTS(X)
#undef X
    return 0;
}

hipError_t launch_grid_rows_reg(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                int n_sets, int n, int tile, int64_t gi0, int n_img, float* inter,
                                const float* tw, hipStream_t stream, bool uniform) {
    if (n_img == 0) return hipSuccess;
    if (uniform) {
#define X(NN, T) \
    if (n == NN) \
        return launch_rows_reg_t<NN, T, true>(first, sets, img_set, n_sets, gi0, n_img, inter, \
                                              tw, stream);
        BLINK_ROW_UNIFORM_VARIANTS(X)
#undef X
#define X(NN, T) \
    if (n == NN) \
        return launch_rows_wide_t<NN, T, true>(first, sets, img_set, n_sets, gi0, n_img, inter, \
                                               tw + 2 * NN, stream);
        BLINK_ROW_WIDE_UNIFORM_VARIANTS(X)
#undef X
        return hipErrorNotSupported;
    }
#define X(NN, T) \
    if (n == NN && tile == T) \
        return launch_rows_reg_t<NN, T, false>(first, sets, img_set, n_sets, gi0, n_img, inter, \
                                               tw, stream);
    BLINK_ROW_VARIANTS(X)
#undef X
#define X(NN, T) \
    if (n == NN && tile == T) \
        return launch_rows_wide_t<NN, T, false>(first, sets, img_set, n_sets, gi0, n_img, inter, \
                                                tw + 2 * NN, stream);
    BLINK_ROW_WIDE_VARIANTS(X)
#undef X
    return hipErrorNotSupported;
}

hipError_t launch_cols_reg(const SetDesc& first, const SetDesc* sets, const int* img_set,
                           int n_sets, int n, int tile, int64_t gi0, int n_img,
                           const float* inter, const float* tw, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
#define X(NN, T) \
    if (n == NN && tile == T) \
        return launch_cols_reg_t<NN, T>(first, sets, img_set, n_sets, gi0, n_img, inter, tw, \
                                        stream);
    BLINK_COL_VARIANTS(X)
#undef X
#define X(NN, T) \
    if (n == NN && tile == T) \
        return launch_cols_wide_t<NN, T>(first, sets, img_set, n_sets, gi0, n_img, inter, \
                                         tw + 2 * NN, stream);
    BLINK_COL_WIDE_VARIANTS(X)
#undef X
    return hipErrorNotSupported;
}

bool batch_direct_supported(int n) {
#define X(NN, T) \
    if (n == NN) return true;
    BLINK_DIRECT_VARIANTS(X)
#undef X
    return n == kDirectWinSide;
}

bool direct_uses_window(int n) { return n == kDirectWinSide; }

template <int N, int NT, bool kKb>
static hipError_t launch_rows_direct_win_t(const SetDesc& first, const SetDesc* sets,
                                           const int* img_set, int n_sets, int64_t gi0,
                                           int n_img, float* inter, const float* tw,
                                           hipStream_t stream) {
    // the tile takes the LDS the twiddle copy leaves; blocks of whole rows
    // need at least two rows
    const int64_t cap = g_lds_limit / (int64_t)sizeof(float2) - N;
    if (((cap / N) & ~1) < 2) return hipErrorInvalidConfiguration;
    const size_t lds = sizeof(float2) * (size_t)(cap + N);
    auto k = n_sets == 1 ? k_grid_rows_direct_win<N, NT, false, kKb>
                         : k_grid_rows_direct_win<N, NT, true, kKb>;
    hipLaunchKernelGGL(k, dim3((unsigned)n_img), dim3(NT), lds, stream, first, sets, img_set, gi0,
                       (float2*)inter, (const float2*)tw, (int)cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the images the window does not hold: a workgroup per CU takes them in turn
    auto kb = n_sets == 1 ? k_grid_rows_direct_blocks<N, kDirectBlockThreads, false, kKb>
                          : k_grid_rows_direct_blocks<N, kDirectBlockThreads, true, kKb>;
    const int blocks = (int)std::min<int64_t>(n_img, std::max(1, device_cu_count()));
    hipLaunchKernelGGL(kb, dim3((unsigned)blocks), dim3(kDirectBlockThreads), lds, stream, first,
                       sets, img_set,
                       gi0, n_img, (float2*)inter, (const float2*)tw, (int)cap);
    return hipGetLastError();
}

hipError_t launch_grid_rows_direct(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                   int n_sets, int n, int64_t gi0, int n_img, float* inter,
                                   const float* tw, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    if (n == kDirectWinSide)
        return first.kb_size > 0
                   ? launch_rows_direct_win_t<kDirectWinSide, kDirectWinThreads, true>(
                         first, sets, img_set, n_sets, gi0, n_img, inter, tw, stream)
                   : launch_rows_direct_win_t<kDirectWinSide, kDirectWinThreads, false>(
                         first, sets, img_set, n_sets, gi0, n_img, inter, tw, stream);
    if (first.kb_size > 0) return hipErrorNotSupported;  // convolutional: the window side only
#define X(NN, T) \
    if (n == NN) \
        return launch_rows_direct_t<NN, T>(first, sets, img_set, n_sets, gi0, n_img, inter, tw, \
                                           stream);
    BLINK_DIRECT_VARIANTS(X)
#undef X
    return hipErrorNotSupported;
}

hipError_t launch_direct_prep(const SetDesc* sets, int n_sets, int max_work,
                              hipStream_t stream) {
    if (n_sets == 0 || max_work <= 0) return hipSuccess;
    // (at most kUvMaxParts blocks: one partial maximum each)
    const unsigned blocks = (unsigned)std::min(kUvMaxParts, (max_work + 255) / 256);
    hipLaunchKernelGGL(k_direct_prep, dim3(blocks, (unsigned)n_sets), dim3(256), 0, stream, sets);
    return hipGetLastError();
}

hipError_t set_batch_kernel_attributes() {
    const int max_lds = (int)g_lds_limit;
    hipError_t e = hipSuccess;
    auto set = [&](const void* f) {
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
#define X(NN, T)                                              \
    set((const void*)k_grid_rows_reg<NN, T, false, false>);   \
    set((const void*)k_grid_rows_reg<NN, T, true, false>);
    BLINK_ROW_VARIANTS(X)
#undef X
#define X(NN, T)                                             \
    set((const void*)k_grid_rows_reg<NN, T, false, true>);   \
    set((const void*)k_grid_rows_reg<NN, T, true, true>);
    BLINK_ROW_UNIFORM_VARIANTS(X)
#undef X
#define X(NN, T)                                             \
    set((const void*)k_grid_rows_direct<NN, T, false>);   \
    set((const void*)k_grid_rows_direct<NN, T, true>);
    BLINK_DIRECT_VARIANTS(X)
#undef X
    for (const void* f :
         {(const void*)k_grid_rows_direct_win<kDirectWinSide, kDirectWinThreads, false, false>,
          (const void*)k_grid_rows_direct_win<kDirectWinSide, kDirectWinThreads, true, false>,
          (const void*)k_grid_rows_direct_win<kDirectWinSide, kDirectWinThreads, false, true>,
          (const void*)k_grid_rows_direct_win<kDirectWinSide, kDirectWinThreads, true, true>,
          (const void*)k_grid_rows_direct_blocks<kDirectWinSide, kDirectBlockThreads, false, false>,
          (const void*)k_grid_rows_direct_blocks<kDirectWinSide, kDirectBlockThreads, true, false>,
          (const void*)k_grid_rows_direct_blocks<kDirectWinSide, kDirectBlockThreads, false, true>,
          (const void*)k_grid_rows_direct_blocks<kDirectWinSide, kDirectBlockThreads, true, true>})
        set(f);
#define X(NN, T)                                  \
    set((const void*)k_cols_reg<NN, T, false>);   \
    set((const void*)k_cols_reg<NN, T, true>);
    BLINK_COL_VARIANTS(X)
#undef X
#define X(NN, T)                                               \
    set((const void*)k_grid_rows_wide<NN, T, false, false>);   \
    set((const void*)k_grid_rows_wide<NN, T, true, false>);
    BLINK_ROW_WIDE_VARIANTS(X)
#undef X
#define X(NN, T)                                              \
    set((const void*)k_grid_rows_wide<NN, T, false, true>);   \
    set((const void*)k_grid_rows_wide<NN, T, true, true>);
    BLINK_ROW_WIDE_UNIFORM_VARIANTS(X)
#undef X
#define X(NN, T)                                   \
    set((const void*)k_cols_wide<NN, T, false>);   \
    set((const void*)k_cols_wide<NN, T, true>);
    BLINK_COL_WIDE_VARIANTS(X)
#undef X
    return e;
}

}  // namespace blink
