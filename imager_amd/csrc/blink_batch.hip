// Batched visibilities -> images kernels for power-of-two grid sides, built on
// the register-resident FFT (blink_fft_reg.hpp).
//
//   k_grid_rows_reg : (row tile, image) -> plan entries gathered (pol, phase
//                     correction, NaN/MAX_VIS) and accumulated in LDS, each
//                     non-empty row transformed, the tile written TRANSPOSED
//                     to the row-pass buffer inter[img][u][v];
//   k_cols_reg      : (column block, image) -> each column read contiguously
//                     from inter (rows with no plan entries read as zero),
//                     transformed, scaled by (float)(1/sum counts) and stored
//                     fft-shifted into images[img][y][x].
#include "blink_fft_reg.hpp"
#include "blink_gather.hpp"
#include "blink_kernels.hpp"

namespace blink {

// Row stride of the LDS tiles: the padded row plus one element, an odd number
// of float2, so that walking down a column of the tile hits distinct banks.
__host__ __device__ constexpr int tile_stride(int n) { return padded_len(n) + 1; }

template <int N>
struct BatchShape;
// lanes per row (L), rows per row tile / columns per column block (RB)
template <> struct BatchShape<128> { static constexpr int L = 8, RB = 16; };
template <> struct BatchShape<256> { static constexpr int L = 16, RB = 16; };
template <> struct BatchShape<512> { static constexpr int L = 32, RB = 16; };
template <> struct BatchShape<1024> { static constexpr int L = 64, RB = 16; };
template <> struct BatchShape<2048> { static constexpr int L = 64, RB = 8; };

template <int N>
constexpr size_t batch_lds_bytes() {
    return sizeof(float2) * (size_t)BatchShape<N>::RB * tile_stride(N);
}

template <int N>
__global__ void __launch_bounds__(BatchShape<N>::RB * BatchShape<N>::L)
k_grid_rows_reg(GatherArgs g, const int* __restrict__ row_start,
                const uint32_t* __restrict__ entries, int64_t img0, int n_img,
                float2* __restrict__ inter, uint32_t* __restrict__ isums,
                const float2* __restrict__ tw) {
    constexpr int L = BatchShape<N>::L, RB = BatchShape<N>::RB, TS = tile_stride(N);
    constexpr int NT = RB * L;
    using F = RegFft<N, L>;
    extern __shared__ float4 rows_reg_lds4[];
    float2* tile = (float2*)rows_reg_lds4;  // [RB][TS]
    __shared__ int rs[RB + 1];
    __shared__ int red[NT / 64 + 1];

    constexpr int n_rb = N / RB;
    const unsigned nwg = gridDim.x;
    unsigned bid = blockIdx.x;
    if ((nwg & 7u) == 0) bid = (bid & 7u) * (nwg >> 3) + (bid >> 3);  // XCD-contiguous images
    const int tile_id = (int)(bid % n_rb);
    const int li = (int)(bid / n_rb);
    const int64_t img = img0 + li;
    const int ch = (int)(img % g.n_ch);
    const int r0 = tile_id * RB;
    const int tid = threadIdx.x;
    const int* rsg = row_start + (int64_t)ch * (N + 1);
    if (tid <= RB) rs[tid] = rsg[r0 + tid];
    __syncthreads();
    if (rs[RB] == rs[0]) return;  // uniform: nothing lands in this tile

    for (int i = tid; i < RB * TS; i += NT) tile[i] = make_float2(0.f, 0.f);
    __syncthreads();
    const uint32_t* ent = entries + (int64_t)ch * 2 * g.n_bl;
    int n_ok = 0;
    for (int e = rs[0] + tid; e < rs[RB]; e += NT) {
        int u_idx;
        float2 v;
        if (!gather_entry(g, img, ch, ent[e], u_idx, v)) continue;
        int r = 0;
        while (r + 1 < RB && rs[r + 1] <= e) r++;
        float2* c = &tile[r * TS + pad_index(u_idx)];
        atomicAdd(&c->x, v.x);
        atomicAdd(&c->y, v.y);
        n_ok++;
    }
    const int tot = block_sum_int(n_ok, red);
    if (tid == 0 && tot) atomicAdd(&isums[img], (uint32_t)tot);
    __syncthreads();

    const int team = tid / L, lane = tid % L;
    float2* row = tile + team * TS;
    if (rs[team + 1] != rs[team]) {
        float2 v[F::V];
        F::load(v, lane, [&](int i) { return row[pad_index(i)]; });
        team_sync<false>();
        F::template run<1>(v, row, lane, tw);
        F::store(v, lane, [&](int i, float2 x) { row[pad_index(i)] = x; });
    }
    __syncthreads();
    // transposed store: inter[li][c][r0 + rr], RB consecutive rows per column
    float2* dst = inter + (int64_t)li * N * N + r0;
    for (int i = tid; i < RB * N; i += NT) {
        const int c = i / RB, rr = i % RB;
        if (rs[rr + 1] != rs[rr]) dst[(int64_t)c * N + rr] = tile[rr * TS + pad_index(c)];
    }
}

template <int N>
__global__ void __launch_bounds__(BatchShape<N>::RB * BatchShape<N>::L)
k_cols_reg(const float2* __restrict__ inter, float2* __restrict__ images, int64_t img0,
           int n_ch, const int* __restrict__ row_start, const uint32_t* __restrict__ isums,
           const float2* __restrict__ tw) {
    constexpr int L = BatchShape<N>::L, CB = BatchShape<N>::RB, TS = tile_stride(N);
    constexpr int NT = CB * L;
    using F = RegFft<N, L>;
    extern __shared__ float4 cols_reg_lds4[];
    float2* tile = (float2*)cols_reg_lds4;  // [CB][TS]
    __shared__ unsigned char nonempty[N];

    constexpr int n_cb = N / CB;
    const int cb = blockIdx.x % n_cb;
    const int li = blockIdx.x / n_cb;
    const int64_t img = img0 + li;
    const int tid = threadIdx.x;
    const int* rsg = row_start + (img % n_ch) * (int64_t)(N + 1);
    for (int i = tid; i < N; i += NT) nonempty[i] = rsg[i + 1] != rsg[i];
    __syncthreads();

    const int team = tid / L, lane = tid % L;
    const int c0 = cb * CB;
    const float2* src = inter + ((int64_t)li * N + c0 + team) * N;
    float2 v[F::V];
    F::load(v, lane, [&](int i) { return nonempty[i] ? src[i] : make_float2(0.f, 0.f); });
    float2* row = tile + team * TS;
    F::template run<1>(v, row, lane, tw);
    const float s = (float)(1.0 / (double)isums[img]);
#pragma unroll
    for (int k = 0; k < F::V; k++) v[k] = cscale(v[k], s);
    team_sync<false>();
    F::store(v, lane, [&](int i, float2 x) { row[pad_index(i)] = x; });
    __syncthreads();
    // shifted store: images[img][shift(r)][shift(c0 + w)], CB columns per row
    float2* out = images + img * (int64_t)N * N;
    constexpr int H = N / 2;
    const int oc0 = c0 < H ? c0 + H : c0 - H;  // shift of a block that does not straddle N/2
    for (int i = tid; i < CB * N; i += NT) {
        const int r = i / CB, w = i % CB;
        const int orow = r < H ? r + H : r - H;
        out[(int64_t)orow * N + oc0 + w] = tile[w * TS + pad_index(r)];
    }
}

template <int N>
static hipError_t launch_rows_reg_t(const GatherArgs& g, const int* row_start,
                                    const uint32_t* entries, int64_t img0, int n_img, float* inter,
                                    uint32_t* isums, const float* tw, hipStream_t stream) {
    constexpr int NT = BatchShape<N>::RB * BatchShape<N>::L;
    constexpr size_t lds = batch_lds_bytes<N>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)((int64_t)(N / BatchShape<N>::RB) * n_img));
    hipLaunchKernelGGL(k_grid_rows_reg<N>, grid, dim3(NT), lds, stream, g, row_start, entries,
                       img0, n_img, (float2*)inter, isums, (const float2*)tw);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_cols_reg_t(const float* inter, float* images, int64_t img0, int n_img,
                                    int n_ch, const int* row_start, const uint32_t* isums,
                                    const float* tw, hipStream_t stream) {
    constexpr int NT = BatchShape<N>::RB * BatchShape<N>::L;
    constexpr size_t lds = batch_lds_bytes<N>();
    if ((int64_t)lds > g_lds_limit) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)((int64_t)(N / BatchShape<N>::RB) * n_img));
    hipLaunchKernelGGL(k_cols_reg<N>, grid, dim3(NT), lds, stream, (const float2*)inter,
                       (float2*)images, img0, n_ch, row_start, isums, (const float2*)tw);
    return hipGetLastError();
}

bool batch_reg_supported(int n) {
    switch (n) {
        case 128: case 256: case 512: case 1024: case 2048:
            return true;
        default:
            return false;
    }
}

hipError_t launch_grid_rows_reg(const GatherArgs& g, const int* row_start, const uint32_t* entries,
                                int n, int64_t img0, int n_img, float* inter, uint32_t* isums,
                                const float* tw, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    switch (n) {
        case 128: return launch_rows_reg_t<128>(g, row_start, entries, img0, n_img, inter, isums, tw, stream);
        case 256: return launch_rows_reg_t<256>(g, row_start, entries, img0, n_img, inter, isums, tw, stream);
        case 512: return launch_rows_reg_t<512>(g, row_start, entries, img0, n_img, inter, isums, tw, stream);
        case 1024: return launch_rows_reg_t<1024>(g, row_start, entries, img0, n_img, inter, isums, tw, stream);
        case 2048: return launch_rows_reg_t<2048>(g, row_start, entries, img0, n_img, inter, isums, tw, stream);
        default: return hipErrorNotSupported;
    }
}

hipError_t launch_cols_reg(const float* inter, float* images, int n, int64_t img0, int n_img,
                           int n_ch, const int* row_start, const uint32_t* isums, const float* tw,
                           hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    switch (n) {
        case 128: return launch_cols_reg_t<128>(inter, images, img0, n_img, n_ch, row_start, isums, tw, stream);
        case 256: return launch_cols_reg_t<256>(inter, images, img0, n_img, n_ch, row_start, isums, tw, stream);
        case 512: return launch_cols_reg_t<512>(inter, images, img0, n_img, n_ch, row_start, isums, tw, stream);
        case 1024: return launch_cols_reg_t<1024>(inter, images, img0, n_img, n_ch, row_start, isums, tw, stream);
        case 2048: return launch_cols_reg_t<2048>(inter, images, img0, n_img, n_ch, row_start, isums, tw, stream);
        default: return hipErrorNotSupported;
    }
}

hipError_t set_batch_kernel_attributes() {
    const int max_lds = (int)g_lds_limit;
    hipError_t e = hipSuccess;
    auto set = [&](const void* f) {
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    set((const void*)k_grid_rows_reg<128>);
    set((const void*)k_grid_rows_reg<256>);
    set((const void*)k_grid_rows_reg<512>);
    set((const void*)k_grid_rows_reg<1024>);
    set((const void*)k_grid_rows_reg<2048>);
    set((const void*)k_cols_reg<128>);
    set((const void*)k_cols_reg<256>);
    set((const void*)k_cols_reg<512>);
    set((const void*)k_cols_reg<1024>);
    set((const void*)k_cols_reg<2048>);
    return e;
}

}  // namespace blink
