// gfx950 kernels of the visibility -> dirty-image path, with their host-side
// launchers.  Every launcher validates the shape it is given against what the
// kernel and its grid assume before it enqueues anything.
#include "blink_kernels.hpp"
#include "blink_gather.hpp"

#include <algorithm>

namespace blink {

int64_t g_lds_limit = 64 * 1024;
int64_t g_lds_budget = 48 * 1024;

int device_cu_count() {
    static const int n_cu = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        return n;
    }();
    return n_cu;
}

// Launch-time guard: a kernel whose LDS request exceeds the device limit is
// not launched.
static bool lds_fits(size_t bytes) { return (int64_t)bytes <= g_lds_limit; }

// Workgroup of the generic direct row passes (one per image) and its teams.
constexpr int kDirectThreads = 1024, kDirectTeam = 64;

// Sides up to kTwLdsMax copy the twiddle table into LDS in the generic
// second-pass kernels, so that the Stockham passes do not wait on global
// loads between team syncs (768 x 180^2: 121 -> 115 us per launch; the
// plan-path row kernel, whose workgroups are short, measured 444 -> 469 us
// with it and keeps the global table).
constexpr int kTwLdsMax = 1024;
__host__ __device__ constexpr size_t tw_lds_bytes(int n) {
    return n <= kTwLdsMax ? sizeof(float2) * (size_t)n : 0;
}


// ------------------------------------------------------------------ helpers

__global__ void k_zero(float4* __restrict__ p, int64_t n_vec) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_vec; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_zero_tail(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

static unsigned grid_for(int64_t work, int block, int max_blocks) {
    int64_t b = (work + block - 1) / block;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (unsigned)b;
}

hipError_t launch_zero(void* ptr, int64_t bytes, hipStream_t stream) {
    if (bytes <= 0) return hipSuccess;
    float* f = (float*)ptr;
    const int64_t n_float = bytes / 4;
    const bool aligned = ((uintptr_t)ptr % 16) == 0;
    const int64_t n_vec = aligned ? n_float / 4 : 0;
    if (n_vec > 0) {
        hipLaunchKernelGGL(k_zero, dim3(grid_for(n_vec, 256, 8192)), dim3(256), 0, stream,
                           (float4*)f, n_vec);
    }
    const int64_t tail = n_float - n_vec * 4;
    if (tail > 0) {
        hipLaunchKernelGGL(k_zero_tail, dim3(grid_for(tail, 256, 1 << 20)), dim3(256), 0, stream,
                           f + n_vec * 4, tail);
    }
    return hipGetLastError();
}

// ----------------------------------------------------- (a) phase corrections

// One thread per (t, ch, bl) record: the 4 pol products (32 bytes) are
// rotated by exp(i 2 pi turns) in place.
__global__ void k_phase_correct(float4* __restrict__ vis, int n_bl, int n_ch, int64_t n_rec,
                                const float* __restrict__ w_m, const double* __restrict__ cable_m,
                                const double* __restrict__ freq_hz, unsigned which) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rec; i += stride) {
        const unsigned bl = (unsigned)(i % n_bl);
        const int ch = (int)((i / n_bl) % n_ch);
        const double nu = freq_hz[ch];
        double turns = 0.0;
        if (which & 1u) turns += reduce_turns(geometric_turns(w_m[bl], nu));
        if (which & 2u) {
            unsigned a1, a2;
            baseline_antennas(bl, a1, a2);
            turns += reduce_turns(cable_turns(cable_m[a2] - cable_m[a1], nu));
        }
        const float2 ph = unit_phasor(turns);
        float4 a = vis[2 * i], b = vis[2 * i + 1];
        float4 ra, rb;
        ra.x = a.x * ph.x - a.y * ph.y;  ra.y = a.y * ph.x + a.x * ph.y;
        ra.z = a.z * ph.x - a.w * ph.y;  ra.w = a.w * ph.x + a.z * ph.y;
        rb.x = b.x * ph.x - b.y * ph.y;  rb.y = b.y * ph.x + b.x * ph.y;
        rb.z = b.z * ph.x - b.w * ph.y;  rb.w = b.w * ph.x + b.z * ph.y;
        vis[2 * i] = ra;
        vis[2 * i + 1] = rb;
    }
}

hipError_t launch_phase_correct(float* vis, int n_ant, int n_ch, int n_t, const float* w_m,
                                const double* cable_m, const double* freq_hz, unsigned which,
                                hipStream_t stream) {
    const int n_bl = n_ant * (n_ant + 1) / 2;
    const int64_t n_rec = (int64_t)n_bl * n_ch * n_t;
    if (n_rec == 0) return hipSuccess;
    hipLaunchKernelGGL(k_phase_correct, dim3(grid_for(n_rec, 256, 65536)), dim3(256), 0, stream,
                       (float4*)vis, n_bl, n_ch, n_rec, w_m, cable_m, freq_hz, which);
    return hipGetLastError();
}

// ----------------------------------------------------------- cell indices

__global__ void k_uv_cells(const float* __restrict__ u_m, const float* __restrict__ v_m,
                           const int* __restrict__ flags, int n_bl,
                           const double* __restrict__ freq_hz, int n_ch, double du, double dv,
                           double min_uv, int n_pixels, int4* __restrict__ cells) {
    const int64_t n = (int64_t)n_bl * n_ch;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int bl = (int)(i % n_bl);
        const int ch = (int)(i / n_bl);
        int4 c = make_int4(-1, -1, -1, -1);
        int up, vp;
        if (!flags[bl] &&
            cell_of(u_m[bl], v_m[bl], wavelength_of(freq_hz[ch]), du, dv, min_uv, n_pixels, up, vp)) {
            c = make_int4(wrap_index(up, n_pixels), wrap_index(vp, n_pixels),
                          wrap_index(-up, n_pixels), wrap_index(-vp, n_pixels));
        }
        cells[i] = c;
    }
}

hipError_t launch_uv_cells(const float* u_m, const float* v_m, const int* flags, int n_bl,
                           const double* freq_hz, int n_ch, double du, double dv, double min_uv,
                           int n_pixels, int* cells, hipStream_t stream) {
    const int64_t n = (int64_t)n_bl * n_ch;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_uv_cells, dim3(grid_for(n, 256, 65536)), dim3(256), 0, stream, u_m, v_m,
                       flags, n_bl, freq_hz, n_ch, du, dv, min_uv, n_pixels, (int4*)cells);
    return hipGetLastError();
}

// ------------------------------------------------------------ array geometry

__global__ void k_zenith_uvw(const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int n_bl, float* __restrict__ u_m,
                             float* __restrict__ v_m, float* __restrict__ w_m) {
    const int bl = blockIdx.x * blockDim.x + threadIdx.x;
    if (bl >= n_bl) return;
    unsigned a1, a2;
    baseline_antennas((unsigned)bl, a1, a2);
    u_m[bl] = (float)(-(x[a1] - x[a2]));
    v_m[bl] = (float)(-(y[a1] - y[a2]));
    w_m[bl] = (float)(-(z[a1] - z[a2]));
}

__global__ void k_baseline_flags(const int* __restrict__ ant_flags, int n_bl,
                                 int* __restrict__ bl_flags) {
    const int bl = blockIdx.x * blockDim.x + threadIdx.x;
    if (bl >= n_bl) return;
    unsigned a1, a2;
    baseline_antennas((unsigned)bl, a1, a2);
    bl_flags[bl] = (a1 == a2 || ant_flags[a1] || ant_flags[a2]) ? 1 : 0;
}

// One workgroup: max |u|, |v| over unflagged baselines (exact in double).
__global__ void __launch_bounds__(1024)
k_uv_max(const float* __restrict__ u_m, const float* __restrict__ v_m,
         const int* __restrict__ flags, int n_bl, double* __restrict__ uv_max) {
    __shared__ double red[2][16];
    double um = -INFINITY, vm = -INFINITY;
    for (int i = threadIdx.x; i < n_bl; i += blockDim.x) {
        if (flags[i]) continue;
        um = fmax(um, fabs((double)u_m[i]));
        vm = fmax(vm, fabs((double)v_m[i]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        um = fmax(um, __shfl_down(um, o, 64));
        vm = fmax(vm, __shfl_down(vm, o, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = um;
        red[1][wid] = vm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            um = fmax(um, red[0][w]);
            vm = fmax(vm, red[1][w]);
        }
        uv_max[0] = um;
        uv_max[1] = vm;
    }
}

hipError_t launch_zenith_uvw(const double* x, const double* y, const double* z, int n_ant,
                             float* u_m, float* v_m, float* w_m, hipStream_t stream) {
    const int n_bl = n_ant * (n_ant + 1) / 2;
    if (n_bl == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zenith_uvw, dim3((n_bl + 255) / 256), dim3(256), 0, stream, x, y, z,
                       n_bl, u_m, v_m, w_m);
    return hipGetLastError();
}

hipError_t launch_baseline_flags(const int* ant_flags, int n_ant, int* bl_flags,
                                 hipStream_t stream) {
    const int n_bl = n_ant * (n_ant + 1) / 2;
    if (n_bl == 0) return hipSuccess;
    hipLaunchKernelGGL(k_baseline_flags, dim3((n_bl + 255) / 256), dim3(256), 0, stream,
                       ant_flags, n_bl, bl_flags);
    return hipGetLastError();
}

hipError_t launch_uv_max(const float* u_m, const float* v_m, const int* flags, int n_bl,
                         double* uv_max, hipStream_t stream) {
    hipLaunchKernelGGL(k_uv_max, dim3(1), dim3(1024), 0, stream, u_m, v_m, flags, n_bl, uv_max);
    return hipGetLastError();
}

// ------------------------------------------------ (b) gridding, plain atomic

// One thread per (t, ch, bl), FP32 global atomics (the plain path, in the
// shape of gridding_kernel, src/gpu/gridding_gpu.cpp:58-124, with per-image
// counts as the CPU path keeps them, src/pacer_imager.cpp:387,408).
__global__ void k_grid_atomic(const float4* __restrict__ vis, int n_bl, int n_ch, int64_t n_rec,
                              const float* __restrict__ u_m, const float* __restrict__ v_m,
                              const int* __restrict__ flags, const double* __restrict__ freq_hz,
                              double du, double dv, double min_uv, int n_pixels, int pol,
                              float* __restrict__ grids, float* __restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gs = (int64_t)n_pixels * n_pixels;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rec; i += stride) {
        const int bl = (int)(i % n_bl);
        const int64_t img = i / n_bl;
        const int ch = (int)(img % n_ch);
        if (flags[bl]) continue;
        const float2 val = select_pol(vis[2 * i], vis[2 * i + 1], pol);
        if (!vis_ok(val)) continue;
        int up, vp;
        if (!cell_of(u_m[bl], v_m[bl], wavelength_of(freq_hz[ch]), du, dv, min_uv, n_pixels, up, vp))
            continue;
        const int64_t p = img * gs + (int64_t)wrap_index(vp, n_pixels) * n_pixels + wrap_index(up, n_pixels);
        const int64_t q = img * gs + (int64_t)wrap_index(-vp, n_pixels) * n_pixels + wrap_index(-up, n_pixels);
        unsafeAtomicAdd(&grids[2 * p], val.x);
        unsafeAtomicAdd(&grids[2 * p + 1], val.y);
        unsafeAtomicAdd(&counts[p], 1.0f);
        unsafeAtomicAdd(&grids[2 * q], val.x);
        unsafeAtomicAdd(&grids[2 * q + 1], -val.y);
        unsafeAtomicAdd(&counts[q], 1.0f);
    }
}

hipError_t launch_grid_atomic(const float* vis, int n_ant, int n_ch, int n_t, const float* u_m,
                              const float* v_m, const int* flags, const double* freq_hz, double du,
                              double dv, double min_uv, int n_pixels, int pol, float* grids,
                              float* counts, hipStream_t stream) {
    const int n_bl = n_ant * (n_ant + 1) / 2;
    const int64_t n_rec = (int64_t)n_bl * n_ch * n_t;
    if (n_rec == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_atomic, dim3(grid_for(n_rec, 256, 65536)), dim3(256), 0, stream,
                       (const float4*)vis, n_bl, n_ch, n_rec, u_m, v_m, flags, freq_hz, du, dv,
                       min_uv, n_pixels, pol, grids, counts);
    return hipGetLastError();
}

// --------------------------------------- (b) convolutional (Kaiser-Bessel) gridding

// One thread per (t, ch, bl).  The visibility and its conjugate are spread
// over K x K cells with weights kt[ky][j] * kt[kx][i] (double), the kernel
// row chosen from the 1023-times oversampled fractional cell position as
// convolve_with_kernel does (src/pacer_imager.cpp:654-714); counts go to the
// nearest cells (:387,408).  FP32 global atomics.
__global__ void k_grid_kb(const float4* __restrict__ vis, int n_bl, int n_ch, int64_t n_rec,
                          const float* __restrict__ u_m, const float* __restrict__ v_m,
                          const int* __restrict__ flags, const double* __restrict__ freq_hz,
                          double du, double dv, double min_uv, int n_pixels, int pol, int K,
                          const double* __restrict__ kt, float* __restrict__ grids,
                          float* __restrict__ counts) {
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gs = (int64_t)n_pixels * n_pixels;
    const int mid = K / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rec; i += stride) {
        const int bl = (int)(i % n_bl);
        const int64_t img = i / n_bl;
        const int ch = (int)(img % n_ch);
        if (flags[bl]) continue;
        const float2 val = select_pol(vis[2 * i], vis[2 * i + 1], pol);
        if (!vis_ok(val)) continue;
        const double wl = wavelength_of(freq_hz[ch]);
        int up, vp;
        if (!cell_of(u_m[bl], v_m[bl], wl, du, dv, min_uv, n_pixels, up, vp)) continue;
        const double ue = (double)u_m[bl] / wl / du;
        const double ve = (double)v_m[bl] / wl / dv;
        int kx = (int)round((ue - (double)up) * 1023.0);
        int ky = (int)round((ve - (double)vp) * 1023.0);
        kx = (kx + 1534) % 1023;
        ky = (ky + 1534) % 1023;
        const double* xk = kt + (int64_t)kx * K;
        const double* yk = kt + (int64_t)ky * K;
        float* g = grids + 2 * img * gs;
        for (int s = 1; s >= -1; s -= 2) {
            for (int j = 0; j < K; j++) {
                int y;
                if (!kb_cell(vp - mid + j, s, n_pixels, y)) continue;
                const double wy = yk[j];
                for (int c = 0; c < K; c++) {
                    int x;
                    if (!kb_cell(up - mid + c, s, n_pixels, x)) continue;
                    const double w = wy * xk[c];
                    const double re = (double)val.x * w;
                    const double im = (double)s * (double)val.y * w;
                    const int64_t p = (int64_t)y * n_pixels + x;
                    unsafeAtomicAdd(&g[2 * p], (float)re);
                    unsafeAtomicAdd(&g[2 * p + 1], (float)im);
                }
            }
        }
        unsafeAtomicAdd(&counts[img * gs + (int64_t)wrap_index(vp, n_pixels) * n_pixels +
                                wrap_index(up, n_pixels)], 1.0f);
        unsafeAtomicAdd(&counts[img * gs + (int64_t)wrap_index(-vp, n_pixels) * n_pixels +
                                wrap_index(-up, n_pixels)], 1.0f);
    }
}

hipError_t launch_grid_kb(const float* vis, int n_ant, int n_ch, int n_t, const float* u_m,
                          const float* v_m, const int* flags, const double* freq_hz, double du,
                          double dv, double min_uv, int n_pixels, int pol, int kernel_size,
                          const double* table, float* grids, float* counts, hipStream_t stream) {
    const int n_bl = n_ant * (n_ant + 1) / 2;
    const int64_t n_rec = (int64_t)n_bl * n_ch * n_t;
    if (n_rec == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_kb, dim3(grid_for(n_rec, 256, 65536)), dim3(256), 0, stream,
                       (const float4*)vis, n_bl, n_ch, n_rec, u_m, v_m, flags, freq_hz, du, dv,
                       min_uv, n_pixels, pol, kernel_size, table, grids, counts);
    return hipGetLastError();
}

// --------------------------------------------- (b) gridding plan (row CSR)

// Per channel: every gridded (baseline, conjugate) pair becomes two entries,
// counting-sorted by grid row.  row_start[ch][0..N] (exclusive prefix sum of
// the row histogram) and entries[ch][0..2*n_bl) packed as
// u | conj << 13 | bl << 14.  One workgroup per channel.
__global__ void __launch_bounds__(kPlanThreads)
k_plan_rows(const float* __restrict__ u_m, const float* __restrict__ v_m,
            const int* __restrict__ flags, int n_bl, const double* __restrict__ freq_hz,
            double du, double dv, double min_uv, int n_pixels, int* __restrict__ row_start,
            uint32_t* __restrict__ entries) {
    extern __shared__ int plan_lds[];
    int* hist = plan_lds;                 // [n_pixels]
    int* cursor = plan_lds + n_pixels;    // [n_pixels]
    __shared__ int warp_tot[kPlanThreads / 64];
    const int ch = blockIdx.x;
    const int tid = threadIdx.x;
    const double wl = wavelength_of(freq_hz[ch]);
    int* rs = row_start + (int64_t)ch * (n_pixels + 1);
    uint32_t* ent = entries + (int64_t)ch * 2 * n_bl;

    for (int i = tid; i < n_pixels; i += kPlanThreads) hist[i] = 0;
    __syncthreads();
    for (int bl = tid; bl < n_bl; bl += kPlanThreads) {
        int up, vp;
        if (flags[bl] || !cell_of(u_m[bl], v_m[bl], wl, du, dv, min_uv, n_pixels, up, vp)) continue;
        atomicAdd(&hist[wrap_index(vp, n_pixels)], 1);
        atomicAdd(&hist[wrap_index(-vp, n_pixels)], 1);
    }
    __syncthreads();
    // Exclusive scan: each thread sums a contiguous chunk, the chunk totals
    // are scanned across the block, then each chunk is rewritten.
    const int chunk = (n_pixels + kPlanThreads - 1) / kPlanThreads;
    const int c0 = min(tid * chunk, n_pixels), c1 = min(c0 + chunk, n_pixels);
    int local = 0;
    for (int i = c0; i < c1; i++) local += hist[i];
    // inclusive scan of `local` across the block
    const int lane = tid & 63, wid = tid >> 6;
    int x = local;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) warp_tot[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int t = (lane < kPlanThreads / 64) ? warp_tot[lane] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < kPlanThreads / 64) warp_tot[lane] = t;
    }
    __syncthreads();
    int run = x - local + (wid > 0 ? warp_tot[wid - 1] : 0);
    for (int i = c0; i < c1; i++) {
        const int h = hist[i];
        rs[i] = run;
        cursor[i] = run;
        run += h;
    }
    if (tid == kPlanThreads - 1) rs[n_pixels] = warp_tot[kPlanThreads / 64 - 1];
    __syncthreads();
    for (int bl = tid; bl < n_bl; bl += kPlanThreads) {
        int up, vp;
        if (flags[bl] || !cell_of(u_m[bl], v_m[bl], wl, du, dv, min_uv, n_pixels, up, vp)) continue;
        const int p = atomicAdd(&cursor[wrap_index(vp, n_pixels)], 1);
        ent[p] = (uint32_t)wrap_index(up, n_pixels) | ((uint32_t)bl << 14);
        const int q = atomicAdd(&cursor[wrap_index(-vp, n_pixels)], 1);
        ent[q] = (uint32_t)wrap_index(-up, n_pixels) | (1u << 13) | ((uint32_t)bl << 14);
    }
}

hipError_t launch_plan_rows(const float* u_m, const float* v_m, const int* flags, int n_bl,
                            const double* freq_hz, int n_ch, double du, double dv, double min_uv,
                            int n_pixels, int* row_start, uint32_t* entries, hipStream_t stream) {
    if (n_ch == 0) return hipSuccess;
    const size_t lds = sizeof(int) * 2 * (size_t)n_pixels;
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_plan_rows, dim3(n_ch), dim3(kPlanThreads), lds, stream, u_m, v_m, flags,
                       n_bl, freq_hz, du, dv, min_uv, n_pixels, row_start, entries);
    return hipGetLastError();
}

// ------------------------------------------ (b) gridding plans of a batched call
//
// One plan unit per (set, channel): the counting sort of the channel's 2 n_bl
// cells by grid row, the same result k_plan_rows gives.  `parts` workgroups
// per unit share its baselines.
//   count    the cell of every baseline of the part (FP64, bit-exact), kept
//            packed for the scatter, and the part's histogram by row;
//   scan     per unit: row starts, and each part's first slot in every row;
//   scatter  each part sorts its entries by row in LDS and writes them out
//            as contiguous runs (no FP64 work, no scattered 4-byte stores).
// Entry order inside a row is not fixed; the sums it feeds are FP32 sums in
// any order either way.

constexpr int kUnitThreads = 512;
// baselines per thread in a plan part (parts hold at most kPlanPartMax)
constexpr int kPlanPerThread = kPlanPartMax / kUnitThreads;
static_assert(kPlanPartMax % kUnitThreads == 0, "plan parts split evenly over the threads");

__device__ __forceinline__ void part_range(int n_bl, int parts, int part, int& b0, int& b1) {
    const int per = (n_bl + parts - 1) / parts;
    b0 = min(n_bl, part * per);
    b1 = min(n_bl, b0 + per);
}

// Exclusive scan of x over the workgroup (kUnitThreads threads); returns the
// thread's prefix and the total in *total.
__device__ __forceinline__ int unit_scan(int x, int* warp_tot, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int y = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(y, o, 64);
        if (lane >= o) y += t;
    }
    if (lane == 63) warp_tot[wid] = y;
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < kUnitThreads / 64; w++) {
        if (w < wid) before += warp_tot[w];
        tot += warp_tot[w];
    }
    *total = tot;
    return before + y - x;
}

__global__ void __launch_bounds__(kUnitThreads)
k_plan_count(const SetDesc* __restrict__ sets, const PlanUnit* __restrict__ units, int parts,
             int n_max, int* __restrict__ partial, uint32_t* __restrict__ cells) {
    extern __shared__ int unit_lds[];
    int* hist = unit_lds;  // [n]
    const int unit = blockIdx.x / parts, part = blockIdx.x % parts;
    const PlanUnit pu = units[unit];
    const SetDesc sd = sets[pu.set];
    const int n = sd.n_pixels, n_bl = sd.g.n_bl;
    const CellUnit cu(sd.g.freq_hz[pu.ch], sd);
    for (int i = threadIdx.x; i < n; i += kUnitThreads) hist[i] = 0;
    __syncthreads();
    int b0, b1;
    part_range(n_bl, parts, part, b0, b1);
    uint32_t* cl = cells + pu.cell_off;
    if (part == 0)  // the count sums of this channel's images start at zero
        for (int t = threadIdx.x; t < sd.n_img / sd.g.n_ch; t += kUnitThreads)
            sd.isums[(int64_t)t * sd.g.n_ch + pu.ch] = 0u;
    if (pu.ch == 0 && sd.g.corrections)  // one unit per set writes the phase coefficients
        for (int bl = b0 + (int)threadIdx.x; bl < b1; bl += kUnitThreads)
            sd.coef_out[bl] = phase_coefficient(sd.g.corrections, sd.w_m, sd.cable_m, (unsigned)bl);
    // at most kPlanPerThread baselines per thread: every load is issued first
    int fl[kPlanPerThread];
    float uu[kPlanPerThread], vv[kPlanPerThread];
#pragma unroll
    for (int k = 0; k < kPlanPerThread; k++) {
        const int bl = b0 + (int)threadIdx.x + k * kUnitThreads;
        fl[k] = 1;
        if (bl < b1) {
            fl[k] = sd.flags[bl];
            uu[k] = sd.u_m[bl];
            vv[k] = sd.v_m[bl];
        }
    }
#pragma unroll
    for (int k = 0; k < kPlanPerThread; k++) {
        const int bl = b0 + (int)threadIdx.x + k * kUnitThreads;
        if (bl >= b1) break;
        int up, vp;
        uint32_t packed = kNoCell;
        if (!fl[k] && plan_cell(uu[k], vv[k], cu, sd, up, vp)) {
            const int r = wrap_index(sd.uv_swap ? up : vp, n);
            packed = ((uint32_t)r << 16) | (uint32_t)wrap_index(sd.uv_swap ? vp : up, n);
            // only rows 0 .. N/2 are planned (the consumers rebuild the rest
            // from the Hermitian symmetry): the cell or its conjugate, or both
            // in rows 0 and N/2
            const int r2 = r == 0 ? 0 : n - r;
            if (r <= n / 2) atomicAdd(&hist[r], 1);
            if (r2 <= n / 2) atomicAdd(&hist[r2], 1);
        }
        cl[bl] = packed;
    }
    __syncthreads();
    int* out = partial + ((int64_t)unit * parts + part) * n_max;
    for (int i = threadIdx.x; i < n; i += kUnitThreads) out[i] = hist[i];
}

__global__ void __launch_bounds__(kUnitThreads)
k_plan_scan(const SetDesc* __restrict__ sets, const PlanUnit* __restrict__ units, int parts,
            int n_max, int* __restrict__ partial) {
    extern __shared__ int unit_lds[];
    int* row_tot = unit_lds;  // [n] entries per row over all parts, then row starts
    __shared__ int warp_tot[kUnitThreads / 64];
    const int unit = blockIdx.x;
    const PlanUnit pu = units[unit];
    const SetDesc sd = sets[pu.set];
    const int n = sd.n_pixels;
    int* rs = sd.row_start + (int64_t)pu.ch * (n + 1);
    int* base = partial + (int64_t)unit * parts * n_max;
    const int tid = threadIdx.x;
    // row totals: one thread per row, parts summed in order (coalesced reads)
    for (int i = tid; i < n; i += kUnitThreads) {
        int t = 0;
#pragma unroll 8
        for (int p = 0; p < parts; p++) t += base[(int64_t)p * n_max + i];
        row_tot[i] = t;
    }
    __syncthreads();
    // exclusive scan of the row totals: contiguous chunks per thread
    const int chunk = (n + kUnitThreads - 1) / kUnitThreads;
    const int c0 = min(tid * chunk, n), c1 = min(c0 + chunk, n);
    int local = 0;
    for (int i = c0; i < c1; i++) local += row_tot[i];
    int total;
    int run = unit_scan(local, warp_tot, &total);
    for (int i = c0; i < c1; i++) {
        const int t = row_tot[i];
        row_tot[i] = run;
        run += t;
    }
    __syncthreads();
    // row starts, and each part's first slot in every row (one thread per row)
    for (int i = tid; i < n; i += kUnitThreads) {
        int r = row_tot[i];
        rs[i] = r;
        for (int p0 = 0; p0 < parts; p0 += 8) {
            int h[8];  // eight independent loads in flight, then the running sums
#pragma unroll
            for (int k = 0; k < 8; k++)
                h[k] = p0 + k < parts ? base[(int64_t)(p0 + k) * n_max + i] : 0;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (p0 + k < parts) {
                    base[(int64_t)(p0 + k) * n_max + i] = r;
                    r += h[k];
                }
        }
    }
    if (tid == 0) rs[n] = total;
    // the non-empty rows 0 .. N/2 and their first entries (row_list_len)
    const int h = n / 2, h1 = h + 1;
    int* rl = sd.row_list + (int64_t)pu.ch * row_list_len(n);
    auto start = [&](int i) { return i < n ? row_tot[i] : total; };
    const int hc = (h1 + kUnitThreads - 1) / kUnitThreads;
    const int q0 = min(tid * hc, h1), q1 = min(q0 + hc, h1);
    int cnt = 0;
    for (int i = q0; i < q1; i++) cnt += start(i + 1) != start(i);
    __syncthreads();  // warp_tot is reused
    int n_rows;
    int at = unit_scan(cnt, warp_tot, &n_rows);
    for (int i = q0; i < q1; i++)
        if (start(i + 1) != start(i)) {
            rl[at] = i;
            rl[h1 + at] = start(i);
            at++;
        }
    for (int p = n_rows + tid; p <= h1; p += kUnitThreads) rl[h1 + p] = start(h1);
}

__global__ void __launch_bounds__(kUnitThreads)
k_plan_scatter(const SetDesc* __restrict__ sets, const PlanUnit* __restrict__ units, int parts,
               int n_max, const int* __restrict__ partial, const uint32_t* __restrict__ cells) {
    extern __shared__ int unit_lds[];
    __shared__ int warp_tot[kUnitThreads / 64];
    const int unit = blockIdx.x / parts, part = blockIdx.x % parts;
    const PlanUnit pu = units[unit];
    const SetDesc sd = sets[pu.set];
    const int n = sd.n_pixels, n_bl = sd.g.n_bl;
    int* first = unit_lds;            // [n] the part's first global slot per row
    int* lstart = first + n_max;      // [n] the row's first local slot
    int* lfill = lstart + n_max;      // [n]
    uint32_t* lent = (uint32_t*)(lfill + n_max);              // [2 kPlanPartMax]
    unsigned short* lrow = (unsigned short*)(lent + 2 * kPlanPartMax);  // [2 kPlanPartMax]
    const int* off = partial + ((int64_t)unit * parts + part) * n_max;
    const int* rs = sd.row_start + (int64_t)pu.ch * (n + 1);
    const int tid = threadIdx.x;
    // the part's count per row is the next part's first slot (or the row
    // end) minus its own; local starts are their exclusive scan
    const int chunk = (n + kUnitThreads - 1) / kUnitThreads;
    const int c0 = min(tid * chunk, n), c1 = min(c0 + chunk, n);
    int local = 0;
    for (int i = c0; i < c1; i++) {
        const int o = off[i];
        const int e = part + 1 < parts ? off[n_max + i] : rs[i + 1];
        first[i] = o;
        lfill[i] = e - o;  // count, for now
        local += e - o;
    }
    int n_local;
    int run = unit_scan(local, warp_tot, &n_local);
    for (int i = c0; i < c1; i++) {
        const int c = lfill[i];
        lstart[i] = run;
        lfill[i] = 0;
        run += c;
    }
    __syncthreads();
    int b0, b1;
    part_range(n_bl, parts, part, b0, b1);
    const uint32_t* cl = cells + pu.cell_off;
    uint32_t pk[kPlanPerThread];
#pragma unroll
    for (int k = 0; k < kPlanPerThread; k++) {
        const int bl = b0 + tid + k * kUnitThreads;
        pk[k] = bl < b1 ? cl[bl] : kNoCell;
    }
#pragma unroll
    for (int k = 0; k < kPlanPerThread; k++) {
        const int bl = b0 + tid + k * kUnitThreads;
        const uint32_t packed = pk[k];
        if (packed == kNoCell) continue;
        const int r = (int)(packed >> 16), u = (int)(packed & 0xffffu);
        if (r <= n / 2) {
            const int p = lstart[r] + atomicAdd(&lfill[r], 1);
            lent[p] = (uint32_t)u | ((uint32_t)bl << 14);
            lrow[p] = (unsigned short)r;
        }
        const int r2 = r == 0 ? 0 : n - r, u2 = u == 0 ? 0 : n - u;
        if (r2 <= n / 2) {
            const int q = lstart[r2] + atomicAdd(&lfill[r2], 1);
            lent[q] = (uint32_t)u2 | (1u << 13) | ((uint32_t)bl << 14);
            lrow[q] = (unsigned short)r2;
        }
    }
    __syncthreads();
    uint32_t* ent = sd.entries + (int64_t)pu.ch * 2 * n_bl;
    for (int i = tid; i < n_local; i += kUnitThreads) {
        const int r = lrow[i];
        ent[first[r] + (i - lstart[r])] = lent[i];
    }
}

hipError_t launch_plan_units(const SetDesc* sets, const PlanUnit* units, int n_units, int parts,
                             int n_max, int* partial, uint32_t* cells, hipStream_t stream) {
    if (n_units == 0) return hipSuccess;
    if (parts < 1 || n_max < 1 || n_max > kMaxPixels) return hipErrorInvalidValue;
    const int64_t blocks = (int64_t)n_units * parts;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    const size_t lds_count = sizeof(int) * (size_t)n_max;
    const size_t lds_scatter = sizeof(int) * 3 * (size_t)n_max +
                               (sizeof(uint32_t) + sizeof(unsigned short)) * 2 * kPlanPartMax;
    if (!lds_fits(lds_scatter)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_plan_count, dim3((unsigned)blocks), dim3(kUnitThreads), lds_count,
                       stream, sets, units, parts, n_max, partial, cells);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_plan_scan, dim3((unsigned)n_units), dim3(kUnitThreads), lds_count, stream,
                       sets, units, parts, n_max, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_plan_scatter, dim3((unsigned)blocks), dim3(kUnitThreads), lds_scatter,
                       stream, sets, units, parts, n_max, (const int*)partial,
                       (const uint32_t*)cells);
    return hipGetLastError();
}

// Locates the tile row of entry e given the tile's row starts rs[0..nr].
__device__ __forceinline__ int row_of_entry(const int* rs, int nr, int e) {
    int r = 0;
    while (r + 1 < nr && rs[r + 1] <= e) r++;
    return r;
}

// ------------------------------- (b) gridding, LDS-privatised row tiles

// Workgroup = (row tile of RB rows, image).  The tile's plan entries are
// accumulated with LDS atomics (complex grid and counts), then the tile is
// added to the global grids/counts (each cell is owned by one workgroup, so
// the flush needs no atomics).
__global__ void __launch_bounds__(kTileThreads)
k_grid_tiles(GatherArgs g, const int* __restrict__ row_start, const uint32_t* __restrict__ entries,
             int n_pixels, int rb, float2* __restrict__ grids, float* __restrict__ counts,
             int overwrite, uint32_t* __restrict__ isums, int64_t img0) {
    extern __shared__ float4 tiles_lds4[];
    float2* tile = (float2*)tiles_lds4;                       // [rb][N]
    float* cnt = (float*)(tile + (size_t)rb * n_pixels);      // [rb][N]
    __shared__ int rs[kMaxTileRows + 1];
    __shared__ int red[kTileThreads / 64];
    const int n_rb = (n_pixels + rb - 1) / rb;
    const int tile_id = blockIdx.x % n_rb;
    const int64_t li = blockIdx.x / n_rb;       // image within the launch
    const int64_t img = img0 + li;              // image within the set
    const int ch = (int)(img % g.n_ch);
    const int r0 = tile_id * rb;
    const int nr = min(rb, n_pixels - r0);
    const int* rsg = row_start + (int64_t)ch * (n_pixels + 1);
    const uint32_t* ent = entries + (int64_t)ch * 2 * g.n_bl;
    if (threadIdx.x <= nr) rs[threadIdx.x] = rsg[r0 + threadIdx.x];
    __syncthreads();
    const int64_t base = li * (int64_t)n_pixels * n_pixels + (int64_t)r0 * n_pixels;
    const int cells = nr * n_pixels;
    if (rs[nr] == rs[0]) {  // nothing lands in this tile
        if (overwrite)
            for (int i = threadIdx.x; i < cells; i += blockDim.x) grids[base + i] = make_float2(0.f, 0.f);
        return;
    }
    for (int i = threadIdx.x; i < cells; i += blockDim.x) {
        tile[i] = make_float2(0.f, 0.f);
        cnt[i] = 0.f;
    }
    __syncthreads();
    const int n_ok = gather_range<4>(g, img, ch, ent, rs[0], rs[nr], (int)threadIdx.x,
                                     (int)blockDim.x, [&](int e, int u_idx, float2 v) {
        const int r = row_of_entry(rs, nr, e);
        const int c = r * n_pixels + u_idx;
        atomicAdd(&tile[c].x, v.x);
        atomicAdd(&tile[c].y, v.y);
        atomicAdd(&cnt[c], 1.0f);
        return 1;
    });
    if (isums) {
        const int tot = block_sum_int(n_ok, red);
        if (threadIdx.x == 0 && tot) atomicAdd(&isums[img], (uint32_t)tot);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cells; i += blockDim.x) {
        const int r = i / n_pixels;
        if (overwrite) {
            grids[base + i] = tile[i];
            if (counts) counts[base + i] = cnt[i];
            continue;
        }
        if (rs[r + 1] == rs[r]) continue;  // empty row: nothing to add
        const float2 t = tile[i];
        float2 o = grids[base + i];
        o.x += t.x;
        o.y += t.y;
        grids[base + i] = o;
        counts[base + i] += cnt[i];
    }
}

hipError_t launch_grid_tiles(const GatherArgs& g, const int* row_start, const uint32_t* entries,
                             int n_images, int n_pixels, float* grids, float* counts,
                             hipStream_t stream, int overwrite, uint32_t* isums, int64_t img0) {
    if (n_images == 0) return hipSuccess;
    // 12 bytes per cell (complex + count) within the LDS budget.
    int rb = (int)(g_lds_budget / (12 * (int64_t)n_pixels));
    rb = std::max(1, std::min(rb, kMaxTileRows));
    const int n_rb = (n_pixels + rb - 1) / rb;
    const size_t lds = (size_t)rb * n_pixels * 12;
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_grid_tiles, dim3((unsigned)((int64_t)n_rb * n_images)), dim3(kTileThreads),
                       lds, stream, g, row_start, entries, n_pixels, rb, (float2*)grids, counts,
                       overwrite, isums, img0);
    return hipGetLastError();
}

// Hermitian half plane for the DFT transform: rows v = 0 .. N/2 of each image
// of a chunk, gridded in LDS tiles of rb rows and written row-major
// to half[li][v][u] (rows of a chunk contiguous); rows without entries are
// written as zeros.  An entry in a row 0 < v < N/2 stands for its conjugate
// partner in row N - v too, so it counts twice in the count sum.
__global__ void __launch_bounds__(kTileThreads)
k_grid_half_rows(GatherArgs g, const int* __restrict__ row_start,
                 const uint32_t* __restrict__ entries, int n, int rb, int64_t img0,
                 float2* __restrict__ half, uint32_t* __restrict__ isums, int uniform) {
    extern __shared__ float4 ghr_lds4[];
    float4* tile4 = ghr_lds4;  // [rb][n] complex, as n/2 float4 per row
    float2* tile = (float2*)ghr_lds4;
    unsigned* cnt = (unsigned*)(tile + (size_t)rb * n);  // [rb][n] counts ("U" weighting)
    __shared__ int rs[kMaxTileRows + 1];
    __shared__ int red[kTileThreads / 64];
    const int h = n / 2;
    const int n_rb = (h + 1 + rb - 1) / rb;
    // XCD-aware order: each XCD gets a contiguous run of (image, tile) pairs,
    // so the tiles of one image share that XCD's L2 copy of its records.
    const unsigned bid = xcd_order(blockIdx.x, gridDim.x);
    const int tile_id = (int)(bid % n_rb);
    const int64_t li = bid / n_rb;
    const int64_t img = img0 + li;
    const int ch = (int)(img % g.n_ch);
    const int r0 = tile_id * rb;
    const int nr = min(rb, h + 1 - r0);
    const int* rsg = row_start + (int64_t)ch * (n + 1);
    if ((int)threadIdx.x <= nr) rs[threadIdx.x] = rsg[r0 + threadIdx.x];
    __syncthreads();
    float4* dst = (float4*)(half + (li * (h + 1) + r0) * (int64_t)n);
    const int q = nr * n / 2;
    if (rs[nr] == rs[0]) {  // no entries: zero rows (empty cells, NaN, under "U")
        const float z = uniform ? __builtin_nanf("") : 0.f;
        for (int i = threadIdx.x; i < q; i += blockDim.x) dst[i] = make_float4(z, z, z, z);
        return;
    }
    for (int i = threadIdx.x; i < q; i += blockDim.x) tile4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (uniform)
        for (int i = threadIdx.x; i < nr * n; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    const uint32_t* ent = entries + (int64_t)ch * 2 * g.n_bl;
    const int n_ok = gather_range<4>(g, img, ch, ent, rs[0], rs[nr], (int)threadIdx.x,
                                     (int)blockDim.x, [&](int e, int u_idx, float2 v) {
        const int r = row_of_entry(rs, nr, e);
        atomicAdd(&tile[r * n + u_idx].x, v.x);
        atomicAdd(&tile[r * n + u_idx].y, v.y);
        if (uniform) atomicAdd(&cnt[r * n + u_idx], 1u);
        const int row = r0 + r;
        return (row == 0 || row == h) ? 1 : 2;
    });
    const int tot = block_sum_int(n_ok, red);
    if (threadIdx.x == 0 && tot) atomicAdd(&isums[img], (uint32_t)tot);
    __syncthreads();
    if (uniform) {  // grid / count per cell, 0 / 0 = NaN in empty cells
        for (int i = threadIdx.x; i < nr * n; i += blockDim.x) {
            const float k = (float)cnt[i];
            tile[i] = make_float2(tile[i].x / k, tile[i].y / k);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < q; i += blockDim.x) dst[i] = tile4[i];
}

hipError_t launch_grid_half_rows(const GatherArgs& g, const int* row_start, const uint32_t* entries,
                                 int n, int64_t n_img, int64_t img0, float* half,
                                 uint32_t* isums, hipStream_t stream, bool uniform) {
    if (n_img == 0) return hipSuccess;
    if (n % 2) return hipErrorInvalidValue;
    // Two rows per workgroup: the gather of plan entries is latency bound, and
    // small tiles keep the most workgroups (and loads) in flight per CU; this
    // measured fastest against 4 to 16 rows on 768-channel 256-antenna sets.
    const int rb = 2;
    const int64_t n_rb = (n / 2 + 1 + rb - 1) / rb;
    const int64_t blocks = n_rb * n_img;
    const size_t lds = (sizeof(float2) + (uniform ? sizeof(unsigned) : 0)) * (size_t)rb * n;
    if (!lds_fits(lds) || blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_grid_half_rows, dim3((unsigned)blocks), dim3(kTileThreads), lds, stream, g,
                       row_start, entries, n, rb, img0, (float2*)half, isums, uniform ? 1 : 0);
    return hipGetLastError();
}

// Direct variant of k_grid_half_rows for dense sets with one timestep per
// channel (no plan; see k_grid_rows_direct_fft): one workgroup per image of
// the chunk holds rows v = 0 .. N/2 in LDS, walks every baseline in order and
// writes the half plane row-major to half[li][v][u], rows without entries as
// zeros.  The count sums start at zero (the direct set-up launch).
__global__ void __launch_bounds__(kDirectThreads)
k_grid_half_direct(SetDesc sd, int64_t img0, float2* __restrict__ half) {
    extern __shared__ float4 ghd_lds4[];
    const int n = sd.n_pixels, h = n / 2;
    float2* tile = (float2*)ghd_lds4;  // [h + 1][n]
    __shared__ int red[kDirectThreads / 64];
    const int64_t li = xcd_order(blockIdx.x, gridDim.x);
    const int64_t img = img0 + li;  // image within the set
    const int ch = (int)(img % sd.g.n_ch);
    const int q = (h + 1) * n / 2;  // float4 per image (N even)
    for (int i = threadIdx.x; i < q; i += kDirectThreads) ghd_lds4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int n_ok = direct_walk<2>(sd, img, ch, n, 0, h + 1, (int)threadIdx.x, kDirectThreads,
                                    [&](int r, int c) { return &tile[r * n + c]; });
    const int tot = block_sum_int(n_ok, red);
    if (threadIdx.x == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    float4* dst = (float4*)(half + li * (h + 1) * (int64_t)n);
    for (int i = threadIdx.x; i < q; i += kDirectThreads) dst[i] = ghd_lds4[i];
}

hipError_t launch_grid_half_direct(const SetDesc& sd, int64_t n_img, int64_t img0, float* half,
                                   hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    const int n = sd.n_pixels;
    if (n % 2 || sd.uv_swap) return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)(n / 2 + 1) * n;
    if (!lds_fits(lds) || n_img > 0x7fffffff) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_grid_half_direct, dim3((unsigned)n_img), dim3(kDirectThreads), lds,
                       stream, sd, img0, (float2*)half);
    return hipGetLastError();
}

// grid /= count per cell, the "U" weighting of src/pacer_imager.cpp:421-428.
__global__ void k_uniform_weight(float2* __restrict__ grids, const float* __restrict__ counts,
                                 int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float2 g = grids[i];
        const float c = counts[i];
        g.x = g.x / c;
        g.y = g.y / c;
        grids[i] = g;
    }
}

hipError_t launch_uniform_weight(float* grids, const float* counts, int64_t n_cells,
                                 hipStream_t stream) {
    if (n_cells == 0) return hipSuccess;
    hipLaunchKernelGGL(k_uniform_weight, dim3(grid_for(n_cells, 256, 65536)), dim3(256), 0, stream,
                       (float2*)grids, counts, n_cells);
    return hipGetLastError();
}

// --------------------------------------------------------------- (c) FFT

// Row pass: each team loads one row from global memory, transforms it in LDS
// and writes it back (in == out allowed: a row is fully read before written).
template <int S, bool kBlockTeam>
__global__ void __launch_bounds__(kFftThreads)
k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out, FftPlan plan, int team,
           int64_t n_rows, const float2* __restrict__ tw) {
    extern __shared__ float4 rows_lds4[];
    float2* lds = (float2*)rows_lds4;
    const int n = plan.n;
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* buf = lds + (size_t)t_id * 2 * n;
    float2* tmp = buf + n;
    const int64_t row = (int64_t)blockIdx.x * teams + t_id;
    const bool active = row < n_rows;
    if (active) {
        const float2* src = in + row * n;
        for (int i = lane; i < n; i += team) buf[i] = src[i];
    }
    team_sync<kBlockTeam>();
    if (kBlockTeam || active) {
        const float2* res = fft_row<S, kBlockTeam>(buf, tmp, plan, tw, lane, team);
        if (active) {
            float2* dst = out + row * n;
            for (int i = lane; i < n; i += team) dst[i] = res[i];
        }
    }
}

// Column pass.  A workgroup owns a set of columns: W contiguous columns
// starting at c0 and, when `paired` is set, the W columns starting at
// c0 + N/2 as well (then the fft_shift maps the set onto itself, which makes
// an in-place shifted store safe).  The columns are loaded into an LDS tile
// (transposed, one padded LDS row per column), transformed by the teams, and
// stored row by row, optionally shifted and scaled.
template <int S, bool kBlockTeam>
__global__ void __launch_bounds__(kFftThreads)
k_fft_cols(ColArgs a, FftPlan plan, int team, const float2* __restrict__ tw) {
    extern __shared__ float4 cols_lds4[];
    const int n = plan.n;
    const int ld = n + 1;  // padded LDS row
    const int wt = a.paired ? 2 * a.width : a.width;
    float2* tile = (float2*)cols_lds4;                 // [wt][ld]
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* tmp = tile + (size_t)wt * ld + (size_t)t_id * n;
    const int span = a.paired ? n / 2 : n;
    const int n_cb = (span + a.width - 1) / a.width;
    const int cb = blockIdx.x % n_cb;
    const int64_t img = blockIdx.x / n_cb;
    const int c0 = cb * a.width;
    const int w_here = min(a.width, span - c0);
    const float2* src = a.in + img * a.in_img_stride;
    const int64_t oimg = a.out_img_offset + img;
    float2* dst = a.out + img * a.out_img_stride;
    const int* rs = a.row_start ? a.row_start + (oimg % a.n_ch) * (int64_t)(n + 1) : nullptr;
    const uint32_t* isums = a.isums;

    // load: element (r, w) of the set; columns c0+w and c0+N/2+w
    const int total = n * wt;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int r = i / wt;
        const int wl = i % wt;
        const int half = wl / a.width;  // 0 or 1
        const int w = wl % a.width;
        float2 v = make_float2(0.f, 0.f);
        if (w < w_here && (!rs || rs[r + 1] != rs[r]))
            v = src[(int64_t)r * n + c0 + half * (n / 2) + w];
        tile[(size_t)wl * ld + r] = v;
    }
    __syncthreads();
    const int ncols = a.paired ? 2 * w_here : w_here;
    const int rounds = kBlockTeam ? ncols : (ncols + teams - 1) / teams;
    for (int k = 0; k < rounds; k++) {
        const int cidx = kBlockTeam ? k : k * teams + t_id;
        if (!kBlockTeam && cidx >= ncols) break;
        const int wl = a.paired ? (cidx < w_here ? cidx : a.width + (cidx - w_here)) : cidx;
        float2* col = tile + (size_t)wl * ld;
        const float2* res = fft_row<S, kBlockTeam>(col, tmp, plan, tw, lane, team);
        if (res != col) {
            for (int i = lane; i < n; i += team) col[i] = res[i];
            team_sync<kBlockTeam>();
        }
    }
    __syncthreads();
    float scale = 1.0f;
    if (a.sums) scale = (float)(1.0 / a.sums[oimg]);
    else if (isums) scale = (float)(1.0 / (double)isums[oimg]);
    const bool do_scale = a.sums || isums;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int r = i / wt;
        const int wl = i % wt;
        const int half = wl / a.width;
        const int w = wl % a.width;
        if (w >= w_here) continue;
        const int c = c0 + half * (n / 2) + w;
        float2 v = tile[(size_t)wl * ld + r];
        if (do_scale) v = cscale(v, scale);
        const int orow = a.shift ? shift_index(r, n) : r;
        const int ocol = a.shift ? shift_index(c, n) : c;
        dst[(int64_t)orow * n + ocol] = v;
    }
}

// ---------------------------------------------- (d) sums, shift+norm, averaging

// sums[img] = sum of values of one image, in double.  One block per image.
__global__ void __launch_bounds__(1024)
k_vector_sum(const float* __restrict__ values, int64_t image_size, double* __restrict__ sums) {
    __shared__ double part[16];
    const float* v = values + (int64_t)blockIdx.x * image_size;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < image_size; i += blockDim.x) s += (double)v[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        sums[blockIdx.x] = t;
    }
}

hipError_t launch_vector_sum(const float* values, int64_t image_size, int n_images, double* sums,
                             hipStream_t stream) {
    if (n_images == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vector_sum, dim3(n_images), dim3(1024), 0, stream, values, image_size, sums);
    return hipGetLastError();
}

// In-place fft_shift + scale.  Pixel p = (y, x) and its image q =
// (shift(y), shift(x)) form a 2-cycle (or p == q); the smaller index swaps.
__global__ void k_shift_norm(float2* __restrict__ img, int nx, int ny, int64_t n_total,
                             const double* __restrict__ sums, int n_sums) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t per = (int64_t)nx * ny;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_total; i += stride) {
        const int64_t im = i / per;
        const int64_t p = i % per;
        const int y = (int)(p / nx), x = (int)(p % nx);
        const int64_t q = (int64_t)shift_index(y, ny) * nx + shift_index(x, nx);
        if (q < p) continue;
        const float s = sums ? (float)(1.0 / sums[im % n_sums]) : 1.0f;
        float2* base = img + im * per;
        const float2 a = base[p];
        const float2 b = base[q];
        if (q == p) {
            base[p] = sums ? cscale(a, s) : a;
        } else {
            base[p] = sums ? cscale(b, s) : b;
            base[q] = sums ? cscale(a, s) : a;
        }
    }
}

hipError_t launch_shift_norm(float* images, int nx, int ny, int n_images, const double* sums,
                             int n_sums, hipStream_t stream) {
    const int64_t n = (int64_t)nx * ny * n_images;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shift_norm, dim3(grid_for(n, 256, 65536)), dim3(256), 0, stream,
                       (float2*)images, nx, ny, n, sums, n_sums);
    return hipGetLastError();
}

// Mean over images per pixel, summed in image order in FP32
// (src/gpu/gpu_utils.cpp:143-150; real part only for src/utils.cpp:4-22).
__global__ void k_average(const float2* __restrict__ images, int64_t image_size, int n_images,
                          int real_only, float2* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < image_size; i += stride) {
        float2 t = images[i];
        for (int k = 1; k < n_images; k++) {
            const float2 v = images[i + (int64_t)k * image_size];
            t.x += v.x;
            t.y += v.y;
        }
        const float n = (float)n_images;
        out[i] = real_only ? make_float2(t.x / n, 0.f) : make_float2(t.x / n, t.y / n);
    }
}

hipError_t launch_average(const float* images, int64_t image_size, int n_images, int real_only,
                          float* avg, hipStream_t stream) {
    if (image_size == 0) return hipSuccess;
    hipLaunchKernelGGL(k_average, dim3(grid_for(image_size, 256, 65536)), dim3(256), 0, stream,
                       (const float2*)images, image_size, n_images, real_only, (float2*)avg);
    return hipGetLastError();
}

// ------------------------------------- batched: gridding fused with row FFT

// Workgroup = (tile of RB rows, image).  Plan entries of the tile are
// gathered (pol, phase correction, NaN/MAX_VIS), accumulated with LDS atomics
// into the tile, each non-empty row is transformed in LDS and written to the
// row-pass buffer `inter` (image index relative to the chunk).  Empty rows are
// not written: the column pass reads the plan and treats them as zero.  The
// number of gridded cells per image goes to isums (integer atomics).
template <int S, bool kBlockTeam, bool kMulti>
__global__ void __launch_bounds__(kFftThreads)
k_grid_rows_fft(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
                FftPlan plan, int team, int rb, int64_t gi0, float2* __restrict__ inter,
                const float2* __restrict__ tw) {
    extern __shared__ float4 grf_lds4[];
    const int n = plan.n;
    float2* tile = (float2*)grf_lds4;                 // [rb][n]
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* tmp = tile + (size_t)rb * n + (size_t)t_id * n;
    __shared__ int rs[kMaxTileRows + 1];
    __shared__ int red[kFftThreads / 64];

    // Rows 0 .. N/2 only: the gridded plane is Hermitian, so row N - v is the
    // conjugate mirror of row v and the column pass rebuilds it.
    const int h = n / 2;
    const int n_rb = (h + 1 + rb - 1) / rb;
    // XCD-aware order: consecutive workgroup ids land on different XCDs, so
    // the ids are remapped to give each XCD a contiguous run of (image, tile)
    // pairs and keep an image's visibilities in one L2.
    const unsigned nwg = gridDim.x;
    unsigned bid = blockIdx.x;
    if ((nwg & 7u) == 0) bid = (bid & 7u) * (nwg >> 3) + (bid >> 3);
    const int tile_id = (int)(bid % n_rb);
    const int li = (int)(bid / n_rb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const GatherArgs& g = sd.g;
    const int img = (int)(gi - sd.img_base);  // image within its set
    const int ch = img % g.n_ch;
    const int r0 = tile_id * rb;
    const int nr = min(rb, h + 1 - r0);
    const int* rsg = sd.row_start + (int64_t)ch * (n + 1);
    const uint32_t* ent = sd.entries + (int64_t)ch * 2 * g.n_bl;
    if ((int)threadIdx.x <= nr) rs[threadIdx.x] = rsg[r0 + threadIdx.x];
    __syncthreads();
    if (rs[nr] == rs[0]) return;  // uniform: no entries in this tile
    // "U" weighting: a count per tile cell after the team scratch
    unsigned* cnt = (unsigned*)(tile + (size_t)rb * n + (size_t)teams * n);
    const bool uniform = sd.uniform != 0;
    for (int i = threadIdx.x; i < nr * n; i += blockDim.x) tile[i] = make_float2(0.f, 0.f);
    if (uniform)
        for (int i = threadIdx.x; i < nr * n; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    const int n_ok = gather_range<4>(g, img, ch, ent, rs[0], rs[nr], (int)threadIdx.x,
                                     (int)blockDim.x, [&](int e, int u_idx, float2 v) {
        const int r = row_of_entry(rs, nr, e);
        atomicAdd(&tile[r * n + u_idx].x, v.x);
        atomicAdd(&tile[r * n + u_idx].y, v.y);
        if (uniform) atomicAdd(&cnt[r * n + u_idx], 1u);
        // an entry stands for its mirror too, except in the self-conjugate rows
        // (0, and N/2 for even N), which list both
        const int row = r0 + r;
        return (row == 0 || 2 * row == n) ? 1 : 2;
    });
    const int tot = block_sum_int(n_ok, red);
    if (threadIdx.x == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    if (uniform) {  // grid / count per cell (src/pacer_imager.cpp:421-428), 0 / 0 = NaN
        for (int i = threadIdx.x; i < nr * n; i += blockDim.x) {
            const float k = (float)cnt[i];
            tile[i] = make_float2(tile[i].x / k, tile[i].y / k);
        }
        __syncthreads();
    }
    const int rounds = kBlockTeam ? nr : (nr + teams - 1) / teams;
    for (int k = 0; k < rounds; k++) {
        const int r = kBlockTeam ? k : k * teams + t_id;
        if (!kBlockTeam && r >= nr) break;
        if (rs[r + 1] == rs[r]) continue;  // uniform per team
        float2* row = tile + (size_t)r * n;
        const float2* res = fft_row<S, kBlockTeam>(row, tmp, plan, tw, lane, team);
        float2* dst = inter + ((int64_t)li * (h + 1) + r0 + r) * n;
        for (int i = lane; i < n; i += team) store_stream(&dst[i], res[i]);
        team_sync<kBlockTeam>();
    }
}

// Second pass of the generic batched path, from the half-height row-major
// row-pass buffer inter[img][p][y], plan rows p = 0 .. N/2 (the plan is
// transposed: plan rows are grid columns u, transformed along v by the row
// pass).  As in the register path, the images are real, so one complex FFT
// over p of z = c_y + i c_{y+1} gives image rows y and y+1; plan rows N - p
// are the conjugates of rows p.  A workgroup owns `cb` (even) buffer columns
// = cb/2 pairs, staged through LDS with 16-byte loads, transformed by the
// teams, scaled by 1/sum and stored fftshifted with imaginary parts 0.
template <int S, bool kBlockTeam, bool kMulti>
__global__ void __launch_bounds__(kFftThreads)
k_cols_half(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
            FftPlan plan, int team, int cb, int64_t gi0, const float2* __restrict__ inter,
            const float2* __restrict__ tw) {
    extern __shared__ float4 ch_lds4[];
    const int n = plan.n, h = n / 2, pairs = cb / 2;
    const int ld = n + 1;                                   // padded LDS row
    float2* rows = (float2*)ch_lds4;                        // [pairs][ld]
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* tmp = rows + (size_t)pairs * ld + (size_t)t_id * n;
    float2* twl = rows + (size_t)pairs * ld + (size_t)teams * n;  // [n] when n <= kTwLdsMax
    unsigned char* nonempty = (unsigned char*)twl + tw_lds_bytes(n);
    const float2* twp = n <= kTwLdsMax ? twl : tw;
    if (n <= kTwLdsMax)
        for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = tw[i];

    const int n_cb = n / cb;
    const int cbk = (int)(blockIdx.x % n_cb);
    const int li = (int)(blockIdx.x / n_cb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    if (!sd.images) return;  // a set imaged for its average only (uniform per workgroup)
    const int img = (int)(gi - sd.img_base);
    const uint32_t count = sd.isums[img];
    if (cbk == 0 && threadIdx.x == 0 && sd.count_sums) sd.count_sums[img] = (double)count;
    const int* rsg = sd.row_start + (int64_t)(img % sd.g.n_ch) * (n + 1);
    const float empty = sd.uniform ? __builtin_nanf("") : 0.f;  // rows without entries
    for (int i = threadIdx.x; i <= h; i += blockDim.x)
        nonempty[i] = sd.all_rows || rsg[i + 1] != rsg[i];
    __syncthreads();

    const int c0 = cbk * cb;
    const float2* src = inter + (int64_t)li * (h + 1) * n + c0;
    // four loads in flight per thread, then their LDS writes
    const int stage = (h + 1) * pairs, nt = (int)blockDim.x;
    for (int i0 = threadIdx.x; i0 < stage; i0 += 4 * nt) {
        float4 p[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int i = i0 + k * nt;
            const int v = i / pairs, t = i % pairs;
            p[k] = make_float4(empty, empty, empty, empty);
            if (i < stage && nonempty[v]) p[k] = *(const float4*)&src[(int64_t)v * n + 2 * t];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int i = i0 + k * nt;
            if (i >= stage) break;
            const int v = i / pairs, t = i % pairs;
            float2* zr = rows + (size_t)t * ld;
            zr[v] = make_float2(p[k].x - p[k].w, p[k].y + p[k].z);  // c_x[v] + i c_{x+1}[v]
            if (v > 0 && v < h) zr[n - v] = make_float2(p[k].x + p[k].w, p[k].z - p[k].y);  // row N - v
        }
    }
    __syncthreads();
    const int rounds = kBlockTeam ? pairs : (pairs + teams - 1) / teams;
    for (int k = 0; k < rounds; k++) {
        const int t = kBlockTeam ? k : k * teams + t_id;
        if (!kBlockTeam && t >= pairs) break;
        float2* zr = rows + (size_t)t * ld;
        const float2* res = fft_row<S, kBlockTeam>(zr, tmp, plan, twp, lane, team);
        if (res != zr) {
            for (int i = lane; i < n; i += team) zr[i] = res[i];
            team_sync<kBlockTeam>();
        }
    }
    __syncthreads();
    const float s = (float)(1.0 / (double)count);
    float2* out = sd.images + (int64_t)img * n * n;
    // the plan is transposed: pair t holds Z_t[x] = img[y][x] + i img[y + 1][x]
    // for y = c0 + 2t, so it writes image rows shift(y) and shift(y + 1) (not
    // adjacent when N/2 is odd and y + 1 = N/2) whole, consecutive threads on
    // consecutive output column pairs (even, so 16-byte aligned as N is even)
    for (int i = threadIdx.x; i < h * pairs; i += blockDim.x) {
        const int t = i / h, ox = 2 * (i % h);
        const int x0 = ox < h ? ox + h : ox - h, x1 = ox + 1 < h ? ox + 1 + h : ox + 1 - h;
        const float2 z0 = rows[(size_t)t * ld + x0], z1 = rows[(size_t)t * ld + x1];
        const int y = c0 + 2 * t;
        const int oy0 = y < h ? y + h : y - h, oy1 = y + 1 < h ? y + 1 + h : y + 1 - h;
        store_stream((float4*)&out[(int64_t)oy0 * n + ox], make_float4(z0.x * s, 0.f, z1.x * s, 0.f));
        store_stream((float4*)&out[(int64_t)oy1 * n + ox], make_float4(z0.y * s, 0.f, z1.y * s, 0.f));
    }
}

// ------------------------------------------------------------ launchers

static int fft_team_for(int n) {
    // lanes per row: about 8-16 values per lane, one wave or the whole block
    if (n <= 128) return 16;
    if (n <= 512) return 32;
    if (n <= 1024) return 64;
    return kFftThreads;
}

bool make_fft_plan(int n, FftPlan& plan) {
    if (n < 2 || n > kMaxPixels) return false;
    plan.n = n;
    plan.n_passes = 0;
    int rem = n, span = 1;
    while (rem > 1) {
        int r;
        if (rem % 16 == 0) r = 16;
        else if (rem % 8 == 0) r = 8;
        else if (rem % 4 == 0) r = 4;
        else if (rem % 2 == 0) r = 2;
        else if (rem % 3 == 0) r = 3;
        else if (rem % 5 == 0) r = 5;
        else if (rem % 7 == 0) r = 7;
        else if (rem % 11 == 0) r = 11;
        else if (rem % 13 == 0) r = 13;
        else return false;
        if (plan.n_passes >= kMaxFftPasses) return false;
        plan.radix[plan.n_passes] = r;
        plan.span[plan.n_passes] = span;
        plan.n_passes++;
        span *= r;
        rem /= r;
    }
    return true;
}

template <int S>
static hipError_t fft_rows_t(const float* in, float* out, const FftPlan& plan, int64_t n_rows,
                             const float* tw, hipStream_t stream) {
    const int n = plan.n;
    const int team = fft_team_for(n);
    if (team == kFftThreads) {
        const size_t lds = sizeof(float2) * 2 * (size_t)n;
        if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
        hipLaunchKernelGGL((k_fft_rows<S, true>), dim3((unsigned)n_rows), dim3(kFftThreads), lds,
                           stream, (const float2*)in, (float2*)out, plan, team, n_rows,
                           (const float2*)tw);
    } else {
        const int teams = kFftThreads / team;
        const size_t lds = sizeof(float2) * 2 * (size_t)n * teams;
        if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
        const unsigned blocks = (unsigned)((n_rows + teams - 1) / teams);
        hipLaunchKernelGGL((k_fft_rows<S, false>), dim3(blocks), dim3(kFftThreads), lds, stream,
                           (const float2*)in, (float2*)out, plan, team, n_rows, (const float2*)tw);
    }
    return hipGetLastError();
}

hipError_t launch_fft_rows(const float* in, float* out, const FftPlan& plan, int64_t n_rows,
                           int sign, const float* tw, hipStream_t stream) {
    if (n_rows == 0) return hipSuccess;
    return sign > 0 ? fft_rows_t<1>(in, out, plan, n_rows, tw, stream)
                    : fft_rows_t<-1>(in, out, plan, n_rows, tw, stream);
}

// Column width: the largest divisor of the column span that fits in the LDS
// budget next to the team scratch rows, at most kMaxColWidth.
static int col_width_for(int n, int teams, bool paired) {
    const int64_t budget = g_lds_budget - (int64_t)teams * n * 8;
    int w = (int)(budget / ((int64_t)(n + 1) * 8 * (paired ? 2 : 1)));
    w = std::min(w, kMaxColWidth);
    const int span = paired ? n / 2 : n;
    w = std::max(1, std::min(w, span));
    while (span % w) w--;
    return w;
}

bool cols_paired_fits(int n) {
    const int team = fft_team_for(n);
    const int teams = team == kFftThreads ? 1 : kFftThreads / team;
    const size_t lds = sizeof(float2) * (2 * (size_t)(n + 1) + (size_t)teams * n);
    return n % 2 == 0 && lds_fits(lds);
}

template <int S>
static hipError_t fft_cols_t(ColArgs a, const FftPlan& plan, int n_images, const float* tw,
                             hipStream_t stream) {
    const int n = plan.n;
    const int team = fft_team_for(n);
    const bool block_team = team == kFftThreads;
    const int teams = block_team ? 1 : kFftThreads / team;
    a.width = col_width_for(n, teams, a.paired != 0);
    const int span = a.paired ? n / 2 : n;
    const int n_cb = (span + a.width - 1) / a.width;
    const int wt = a.paired ? 2 * a.width : a.width;
    const size_t lds = sizeof(float2) * ((size_t)wt * (n + 1) + (size_t)teams * n);
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)((int64_t)n_cb * n_images));
    if (block_team)
        hipLaunchKernelGGL((k_fft_cols<S, true>), grid, dim3(kFftThreads), lds, stream, a, plan,
                           team, (const float2*)tw);
    else
        hipLaunchKernelGGL((k_fft_cols<S, false>), grid, dim3(kFftThreads), lds, stream, a, plan,
                           team, (const float2*)tw);
    return hipGetLastError();
}

hipError_t launch_fft_cols(const ColArgs& a, const FftPlan& plan, int n_images, int sign,
                           const float* tw, hipStream_t stream) {
    if (n_images == 0) return hipSuccess;
    return sign > 0 ? fft_cols_t<1>(a, plan, n_images, tw, stream)
                    : fft_cols_t<-1>(a, plan, n_images, tw, stream);
}

int grid_rows_tile_rows(int n) {
    const int team = fft_team_for(n);
    const int teams = team == kFftThreads ? 1 : kFftThreads / team;
    const int64_t budget = g_lds_budget - (int64_t)teams * n * 8;
    int rb = (int)(budget / ((int64_t)n * 8));
    rb = std::min(rb, kMaxTileRows);
    if (rb >= teams) rb = (rb / teams) * teams;
    return std::max(1, std::min(rb, n));
}

hipError_t launch_grid_rows_fft(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                                float* inter, const float* tw, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    const int n = plan.n;
    const int team = fft_team_for(n);
    const bool block_team = team == kFftThreads;
    const int teams = block_team ? 1 : kFftThreads / team;
    // rows per tile: as many as fit up to 4, and at least ~2048 workgroups in
    // the launch, so that a few images still spread over the whole chip.  The
    // gather of plan entries is latency bound: small tiles keep more
    // workgroups in flight (4 rows measured fastest against 2 to 16 at 180^2).
    const int64_t want = (int64_t)(n / 2 + 1) * n_img / 2048;
    int rb = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)grid_rows_tile_rows(n), want, 4}));
    // uniform weighting keeps a count per tile cell beside the tile
    const bool uniform = first.uniform != 0;
    auto lds_of = [&](int r) {
        return sizeof(float2) * ((size_t)r * n + (size_t)teams * n) +
               (uniform ? sizeof(unsigned) * (size_t)r * n : 0);
    };
    while (rb > 1 && !lds_fits(lds_of(rb))) rb--;
    const int n_rb = (n / 2 + 1 + rb - 1) / rb;
    const size_t lds = lds_of(rb);
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)((int64_t)n_rb * n_img));
    const bool multi = n_sets > 1;
    auto k = block_team
                 ? (multi ? k_grid_rows_fft<1, true, true> : k_grid_rows_fft<1, true, false>)
                 : (multi ? k_grid_rows_fft<1, false, true> : k_grid_rows_fft<1, false, false>);
    hipLaunchKernelGGL(k, grid, dim3(kFftThreads), lds, stream, first, sets, img_set, plan, team,
                       rb, gi0, (float2*)inter, (const float2*)tw);
    return hipGetLastError();
}

// Direct row pass of the generic kernels (no plan), for dense sets with one
// timestep per channel, as k_grid_rows_direct does it for the register
// sides: one workgroup per image with the whole half plane (plan rows
// 0 .. N/2 of an even or odd side) in LDS, every baseline walked in order
// (direct_walk), then rows 0 .. the highest holding a cell transformed by
// teams of one wave and stored row-major as the generic second pass reads
// them, and the image's row starts written (the rows above empty).  Sides
// whose half plane
// and team scratch fit in LDS (up to 186).  The count sums start at zero
// (the direct set-up launch).
static size_t direct_fft_lds_bytes(int n) {
    return sizeof(float2) *
           ((size_t)(n / 2 + 1) * n + (size_t)(kDirectThreads / kDirectTeam) * n + n);
}

template <bool kMulti>
__global__ void __launch_bounds__(kDirectThreads)
k_grid_rows_direct_fft(SetDesc first, const SetDesc* __restrict__ sets,
                       const int* __restrict__ img_set, FftPlan plan, int64_t gi0,
                       float2* __restrict__ inter, const float2* __restrict__ tw) {
    extern __shared__ float4 grd_lds4[];
    constexpr int teams = kDirectThreads / kDirectTeam;
    const int n = plan.n, h = n / 2;
    float2* tile = (float2*)grd_lds4;  // [h + 1][n]
    const int t_id = (int)threadIdx.x / kDirectTeam, lane = (int)threadIdx.x % kDirectTeam;
    float2* tmp = tile + (size_t)(h + 1) * n + (size_t)t_id * n;
    // twiddle table copy [n] after the teams' scratch: the passes read it
    // from LDS instead of waiting on global loads between team syncs
    float2* twl = tile + (size_t)(h + 1) * n + (size_t)teams * n;
    __shared__ int red[kDirectThreads / 64];

    const int li = (int)xcd_order(blockIdx.x, gridDim.x);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    const int img = (int)(gi - sd.img_base);
    const int ch = img % sd.g.n_ch;
    for (int i = threadIdx.x; i < (h + 1) * n; i += kDirectThreads) tile[i] = make_float2(0.f, 0.f);
    for (int i = threadIdx.x; i < n; i += kDirectThreads) twl[i] = tw[i];
    __syncthreads();
    int top = -1;
    const int n_ok = direct_walk<2>(
        sd, img, ch, n, 0, h + 1, (int)threadIdx.x, kDirectThreads,
        [&](int r, int c) { return &tile[r * n + c]; }, nullptr, &top);
    const int n_rows = block_max_int(top, red) + 1;
    const int tot = block_sum_int(n_ok, red);
    if (threadIdx.x == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    // rows 0 .. n_rows - 1 (the highest holding a cell); the row starts tell
    // the second pass that the others are empty
    for (int r = t_id; r < n_rows; r += teams) {
        float2* row = tile + (size_t)r * n;
        const float2* res = fft_row<1, false>(row, tmp, plan, twl, lane, kDirectTeam);
        float2* dst = inter + ((int64_t)li * (h + 1) + r) * n;
        for (int i = lane; i < n; i += kDirectTeam) store_stream(&dst[i], res[i]);
        team_sync<false>();
    }
    int* rsg = sd.row_start + (int64_t)ch * (n + 1);
    for (int i = threadIdx.x; i <= n; i += kDirectThreads) rsg[i] = min(i, n_rows);
}

bool direct_half_plane_fits(int n) {
    return n % 2 == 0 && n <= kMaxPixels && lds_fits(sizeof(float2) * (size_t)(n / 2 + 1) * n);
}

bool direct_row_pass_fits(int n) {
    if (batch_reg_supported(n)) return batch_direct_supported(n);
    return n >= 2 && n <= kMaxPixels && lds_fits(direct_fft_lds_bytes(n));
}

hipError_t launch_grid_rows_direct_fft(const SetDesc& first, const SetDesc* sets,
                                       const int* img_set, int n_sets, const FftPlan& plan,
                                       int64_t gi0, int n_img, float* inter, const float* tw,
                                       hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    const size_t lds = direct_fft_lds_bytes(plan.n);
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    auto k = n_sets > 1 ? k_grid_rows_direct_fft<true> : k_grid_rows_direct_fft<false>;
    hipLaunchKernelGGL(k, dim3((unsigned)n_img), dim3(kDirectThreads), lds, stream, first, sets,
                       img_set, plan, gi0, (float2*)inter, (const float2*)tw);
    return hipGetLastError();
}

// Convolutional gridding fused with the row FFT (generic kernels, transposed
// plan: plan rows are grid columns x, positions rows y).  Workgroup = (tile of
// rb consecutive plan rows r0 .. r1 of 0 .. N/2, image).  Every visibility
// that is gridded adds its K x K footprint around its nearest cell (u, v) with
// value V and, mirrored, around (-u, -v) with conj(V) (convolve_with_kernel,
// src/pacer_imager.cpp:654-714, placed as calc_xy_grid2 places it, :716-750).
// The plan lists, per visibility, the one of its two nearest cells that falls
// in rows 1 .. N/2-1 (its partner is in row N - p), or both when they fall in
// rows 0 and N/2.  A footprint around row p covers rows p - K/2 .. p + K/2,
// so a tile gathers the entries of rows r0 - K/2 .. r1 + K/2 and, when r1 +
// K/2 reaches N/2, of rows N - r1 - K/2 .. N/2 (their partners' footprints
// reach back below N/2); each entry adds its own footprint and, for rows 1 ..
// N/2-1, its partner's, keeping the cells in rows r0 .. r1.  Counts stay at
// the nearest cells, as the reference keeps them.  Every row of the tile is
// transformed and written, row-major or, for the register sides, in the
// column-blocked layout of the register second pass.
template <int S, bool kBlockTeam, bool kMulti>
__global__ void __launch_bounds__(kFftThreads)
k_grid_rows_kb(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
               FftPlan plan, int team, int rb, int64_t gi0, float2* __restrict__ inter,
               const float2* __restrict__ tw, int blocked) {
#pragma clang fp contract(off)
    extern __shared__ float4 grk_lds4[];
    const int n = plan.n;
    float2* tile = (float2*)grk_lds4;  // [rb][n]
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* tmp = tile + (size_t)rb * n + (size_t)t_id * n;
    __shared__ int red[kFftThreads / 64];

    const int h = n / 2;
    const int n_rb = (h + 1 + rb - 1) / rb;
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
    const int r0 = tile_id * rb;
    const int nr = min(rb, h + 1 - r0);
    const int r1 = r0 + nr - 1;
    const int K = sd.kb_size, mid = K / 2;
    int lo = max(0, r0 - mid);
    const int hi = min(h, r1 + mid);
    if (r1 + mid >= h) lo = max(0, min(lo, n - r1 - mid));
    const int* rsg = sd.row_start + (int64_t)ch * (n + 1);
    const int e0 = rsg[lo], e1 = rsg[hi + 1];
    for (int i = threadIdx.x; i < nr * n; i += blockDim.x) tile[i] = make_float2(0.f, 0.f);
    __syncthreads();

    const uint32_t* ent = sd.entries + (int64_t)ch * 2 * g.n_bl;
    const float2* rec = (const float2*)g.vis + 4 * ((int64_t)img * g.n_bl);
    const double nu = g.freq_hz[ch];
    const double wl = wavelength_of(nu);
    const int p0 = g.pol == 1 ? 3 : (g.pol == 3 ? 1 : 0);
    const int p1 = g.pol == 3 ? 2 : 3;
    int n_ok = 0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += (int)blockDim.x) {
        const uint32_t ev = ent[e];
        const unsigned bl = ev >> 14;
        const int own = (ev >> 13) & 1u ? -1 : 1;  // sign of the listed cell
        float2 v = select_pol2(rec[4 * (size_t)bl + p0],
                               g.pol >= 2 ? rec[4 * (size_t)bl + p1] : make_float2(0.f, 0.f), g.pol);
        if (g.corrections) v = cmul(v, unit_phasor(((g.phase_coef[bl] * nu) / kVelLight)));
        if (!vis_ok(v)) continue;
        int up, vp;
        if (!cell_of(sd.u_m[bl], sd.v_m[bl], wl, sd.du, sd.dv, sd.min_uv, n, up, vp)) continue;
        // the listed cell's plan row (its column x, as the plan wraps it), and
        // the kernel rows of the fractional position
        const int p = wrap_index(own * up, n);
        if (p >= r0 && p <= r1) n_ok += (p == 0 || p == h) ? 1 : 2;
        const double ue = (double)sd.u_m[bl] / wl / sd.du;
        const double ve = (double)sd.v_m[bl] / wl / sd.dv;
        int kx = (int)round((ue - (double)up) * 1023.0);
        int ky = (int)round((ve - (double)vp) * 1023.0);
        kx = (kx + 1534) % 1023;
        ky = (ky + 1534) % 1023;
        const double* xk = sd.kb_table + (int64_t)kx * K;
        const double* yk = sd.kb_table + (int64_t)ky * K;
        const int n_fp = (p == 0 || p == h) ? 1 : 2;  // own footprint, then the partner's
        for (int f = 0; f < n_fp; f++) {
            const int sgn = f == 0 ? own : -own;
            for (int c = 0; c < K; c++) {
                int x;
                if (!kb_cell(up - mid + c, sgn, n, x) || x < r0 || x > r1) continue;
                const double wx = xk[c];
                float2* row = tile + (size_t)(x - r0) * n;
                for (int j = 0; j < K; j++) {
                    int y;
                    if (!kb_cell(vp - mid + j, sgn, n, y)) continue;
                    const double w = yk[j] * wx;
                    atomicAdd(&row[y].x, (float)((double)v.x * w));
                    atomicAdd(&row[y].y, (float)((double)sgn * (double)v.y * w));
                }
            }
        }
    }
    const int tot = block_sum_int(n_ok, red);
    if (threadIdx.x == 0 && tot) atomicAdd(&sd.isums[img], (uint32_t)tot);
    __syncthreads();
    const int rounds = kBlockTeam ? nr : (nr + teams - 1) / teams;
    const int64_t h1 = h + 1;
    for (int k = 0; k < rounds; k++) {
        const int r = kBlockTeam ? k : k * teams + t_id;
        if (!kBlockTeam && r >= nr) break;
        float2* row = tile + (size_t)r * n;
        const float2* res = fft_row<S, kBlockTeam>(row, tmp, plan, tw, lane, team);
        const int v = r0 + r;
        if (blocked) {  // element (v, x) at ((x / 32) (N/2+1) + v) 32 + x % 32
            float2* dst = inter + (int64_t)li * h1 * n + (int64_t)v * kInterBlockCols;
            for (int i = lane; i < n; i += team)
                store_stream(&dst[(int64_t)(i / kInterBlockCols) * h1 * kInterBlockCols +
                                  i % kInterBlockCols],
                             res[i]);
        } else {
            float2* dst = inter + ((int64_t)li * h1 + v) * n;
            for (int i = lane; i < n; i += team) store_stream(&dst[i], res[i]);
        }
        team_sync<kBlockTeam>();
    }
}

hipError_t launch_grid_rows_kb(const SetDesc& first, const SetDesc* sets, const int* img_set,
                               int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                               float* inter, const float* tw, bool blocked, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    if (first.kb_size < 1) return hipErrorInvalidValue;
    const int n = plan.n;
    const int team = fft_team_for(n);
    const bool block_team = team == kFftThreads;
    const int teams = block_team ? 1 : kFftThreads / team;
    // rows per tile: a footprint covers K rows, so taller tiles re-read fewer
    // entries; as many rows as fit up to 8, and at least ~2048 workgroups
    const int64_t want = std::max<int64_t>(1, (int64_t)(n / 2 + 1) * n_img / 2048);
    int rb = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)grid_rows_tile_rows(n), want, 8}));
    auto lds_of = [&](int r) { return sizeof(float2) * ((size_t)r * n + (size_t)teams * n); };
    while (rb > 1 && !lds_fits(lds_of(rb))) rb--;
    const size_t lds = lds_of(rb);
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    const int n_rb = (n / 2 + 1 + rb - 1) / rb;
    const int64_t blocks = (int64_t)n_rb * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    const bool multi = n_sets > 1;
    auto k = block_team
                 ? (multi ? k_grid_rows_kb<1, true, true> : k_grid_rows_kb<1, true, false>)
                 : (multi ? k_grid_rows_kb<1, false, true> : k_grid_rows_kb<1, false, false>);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(kFftThreads), lds, stream, first, sets,
                       img_set, plan, team, rb, gi0, (float2*)inter, (const float2*)tw,
                       blocked ? 1 : 0);
    return hipGetLastError();
}

// Second pass of the generic batched path for odd N = 2m + 1, from the
// row-major row-pass buffer inter[img][p][y], plan rows p = 0 .. m.  As for
// even N the images are real, so one complex FFT over p of z = c_y + i
// c_{y+1} gives image rows y and y + 1, with rows N - p (p = 1 .. m: an odd
// side has no self-conjugate middle row) the conjugates of rows p; column
// pairs are slots s = 0 .. m (y = 2s), and the last column N - 1 pairs with
// zero columns.  A workgroup owns `pairs` slots; 8-byte accesses (odd rows
// are not 16-byte aligned); the store applies the odd-side fftshift
// (shift_index: p <-> p + m + 1 for p < m, the middle element fixed,
// src/pacer_imager.cpp:30-35,110-128) element by element.
template <int S, bool kBlockTeam, bool kMulti>
__global__ void __launch_bounds__(kFftThreads)
k_cols_half_odd(SetDesc first, const SetDesc* __restrict__ sets, const int* __restrict__ img_set,
                FftPlan plan, int team, int pairs, int64_t gi0, const float2* __restrict__ inter,
                const float2* __restrict__ tw) {
    extern __shared__ float4 cho_lds4[];
    const int n = plan.n, h = n / 2;
    const int ld = n + 1;
    float2* rows = (float2*)cho_lds4;  // [pairs][ld]
    const int teams = kBlockTeam ? 1 : (int)blockDim.x / team;
    const int t_id = kBlockTeam ? 0 : (int)threadIdx.x / team;
    const int lane = kBlockTeam ? (int)threadIdx.x : (int)threadIdx.x % team;
    float2* tmp = rows + (size_t)pairs * ld + (size_t)t_id * n;
    float2* twl = rows + (size_t)pairs * ld + (size_t)teams * n;  // [n] when n <= kTwLdsMax
    unsigned char* nonempty = (unsigned char*)twl + tw_lds_bytes(n);
    const float2* twp = n <= kTwLdsMax ? twl : tw;
    if (n <= kTwLdsMax)
        for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = tw[i];

    const int n_slots = h + 1;
    const int n_cb = (n_slots + pairs - 1) / pairs;
    const int cbk = (int)(blockIdx.x % n_cb);
    const int li = (int)(blockIdx.x / n_cb);
    const int64_t gi = gi0 + li;
    const SetDesc sd = set_of<kMulti>(first, sets, img_set, gi);
    if (!sd.images) return;  // uniform per workgroup
    const int img = (int)(gi - sd.img_base);
    const uint32_t count = sd.isums[img];
    if (cbk == 0 && threadIdx.x == 0 && sd.count_sums) sd.count_sums[img] = (double)count;
    const int* rsg = sd.row_start + (int64_t)(img % sd.g.n_ch) * (n + 1);
    const float empty = sd.uniform ? __builtin_nanf("") : 0.f;  // rows without entries
    for (int i = threadIdx.x; i <= h; i += blockDim.x)
        nonempty[i] = sd.all_rows || rsg[i + 1] != rsg[i];
    __syncthreads();

    const int s0 = cbk * pairs;
    const int np = min(pairs, n_slots - s0);  // slots of this workgroup (uniform)
    const float2* src = inter + (int64_t)li * (h + 1) * n;
    const int stage = (h + 1) * np;
    for (int i = threadIdx.x; i < stage; i += (int)blockDim.x) {
        const int v = i / np, t = i % np;
        const int y = 2 * (s0 + t);
        const bool has_b = y + 1 < n;  // the last column pairs with zeros
        float2 a = make_float2(empty, empty);
        float2 b = has_b ? make_float2(empty, empty) : make_float2(0.f, 0.f);
        if (nonempty[v]) {
            a = src[(int64_t)v * n + y];
            if (has_b) b = src[(int64_t)v * n + y + 1];
        }
        float2* zr = rows + (size_t)t * ld;
        zr[v] = make_float2(a.x - b.y, a.y + b.x);                   // c_y[v] + i c_{y+1}[v]
        if (v > 0) zr[n - v] = make_float2(a.x + b.y, b.x - a.y);    // row N - v
    }
    __syncthreads();
    const int rounds = kBlockTeam ? np : (np + teams - 1) / teams;
    for (int k = 0; k < rounds; k++) {
        const int t = kBlockTeam ? k : k * teams + t_id;
        if (!kBlockTeam && t >= np) break;
        float2* zr = rows + (size_t)t * ld;
        const float2* res = fft_row<S, kBlockTeam>(zr, tmp, plan, twp, lane, team);
        if (res != zr) {
            for (int i = lane; i < n; i += team) zr[i] = res[i];
            team_sync<kBlockTeam>();
        }
    }
    __syncthreads();
    const float s = (float)(1.0 / (double)count);
    float2* out = sd.images + (int64_t)img * n * n;
    for (int i = threadIdx.x; i < np * n; i += (int)blockDim.x) {
        const int t = i / n, x = i % n;
        const int y = 2 * (s0 + t);
        const float2 z = rows[(size_t)t * ld + x];
        const int ox = shift_index(x, n);
        store_stream(&out[(int64_t)shift_index(y, n) * n + ox], make_float2(z.x * s, 0.f));
        if (y + 1 < n)
            store_stream(&out[(int64_t)shift_index(y + 1, n) * n + ox], make_float2(z.y * s, 0.f));
    }
}

// Buffer columns (image rows) per workgroup of k_cols_half: the largest even divisor of
// n, at most 64, whose LDS rows fit next to the team scratch and that still
// gives at least 256 workgroups over the n_img images (one per CU), else the
// smallest, 2 (a single 180^2 image then runs 90 workgroups, not 3).
static int half_cols_block(int n, int teams, int64_t n_img, int max_cb) {
    int fallback = 0;
    for (int cb = std::min(max_cb, n); cb >= 2; cb--) {
        if (cb % 2 || n % cb) continue;
        const int64_t lds = sizeof(float2) * ((int64_t)(cb / 2) * (n + 1) + (int64_t)teams * n) +
                            (int64_t)tw_lds_bytes(n) + (n / 2 + 1);
        if (lds > g_lds_budget) continue;
        if ((int64_t)(n / cb) * n_img >= 256) return cb;
        fallback = cb;
    }
    return fallback;
}

hipError_t launch_cols_half(const SetDesc& first, const SetDesc* sets, const int* img_set,
                            int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                            const float* inter, const float* tw, hipStream_t stream) {
    if (n_img == 0) return hipSuccess;
    const int n = plan.n;
    // Sides up to 256: teams of 16 lanes and at most one column pair per team,
    // so that every pair of a workgroup is transformed in one round (768 x
    // 180^2: 30-column blocks with 16-lane teams 84 us per launch, against
    // 115 us with 60-column blocks and 32-lane teams in two rounds).
    const bool one_round = n <= 256;
    const int team = one_round ? 16 : fft_team_for(n);
    const bool block_team = team == kFftThreads;
    const int teams = block_team ? 1 : kFftThreads / team;
    if (n % 2) {
        // odd N: slots of column pairs per workgroup, as many as fit up to 32
        // (one per team up to 256), and at least ~256 workgroups over the
        // images
        const int n_slots = n / 2 + 1;
        auto lds_of = [&](int p) {
            return sizeof(float2) * ((size_t)p * (n + 1) + (size_t)teams * n) + tw_lds_bytes(n) +
                   (size_t)(n / 2 + 1);
        };
        int pairs = std::min(one_round ? teams : 32, n_slots);
        while (pairs > 1 && (!lds_fits(lds_of(pairs)) ||
                             (int64_t)((n_slots + pairs - 1) / pairs) * n_img < 256))
            pairs--;
        if (!lds_fits(lds_of(pairs))) return hipErrorInvalidConfiguration;
        const int64_t blocks = (int64_t)((n_slots + pairs - 1) / pairs) * n_img;
        if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
        const bool multi = n_sets > 1;
        auto k = block_team
                     ? (multi ? k_cols_half_odd<1, true, true> : k_cols_half_odd<1, true, false>)
                     : (multi ? k_cols_half_odd<1, false, true> : k_cols_half_odd<1, false, false>);
        hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(kFftThreads), lds_of(pairs), stream,
                           first, sets, img_set, plan, team, pairs, gi0, (const float2*)inter,
                           (const float2*)tw);
        return hipGetLastError();
    }
    const int cb = half_cols_block(n, teams, n_img, one_round ? 2 * teams : 64);
    if (!cb) return hipErrorInvalidConfiguration;
    const size_t lds = sizeof(float2) * ((size_t)(cb / 2) * (n + 1) + (size_t)teams * n) +
                       tw_lds_bytes(n) + (size_t)(n / 2 + 1);
    if (!lds_fits(lds)) return hipErrorInvalidConfiguration;
    const int64_t blocks = (int64_t)(n / cb) * n_img;
    if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
    const bool multi = n_sets > 1;
    auto k = block_team ? (multi ? k_cols_half<1, true, true> : k_cols_half<1, true, false>)
                        : (multi ? k_cols_half<1, false, true> : k_cols_half<1, false, false>);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(kFftThreads), lds, stream, first, sets,
                       img_set, plan, team, cb, gi0, (const float2*)inter, (const float2*)tw);
    return hipGetLastError();
}

__global__ void k_isums_to_double(const uint32_t* __restrict__ isums, double* __restrict__ sums,
                                  int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sums[i] = (double)isums[i];
}

hipError_t launch_isums_to_double(const uint32_t* isums, double* sums, int n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_isums_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       isums, sums, n);
    return hipGetLastError();
}

hipError_t set_kernel_attributes() {
    // Allow the dynamic LDS sizes used above.
    const int max_lds = (int)g_lds_limit;
    hipError_t e = hipSuccess;
    auto set = [&](const void* f) {
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    set((const void*)k_fft_rows<1, true>);
    set((const void*)k_fft_rows<1, false>);
    set((const void*)k_fft_rows<-1, true>);
    set((const void*)k_fft_rows<-1, false>);
    set((const void*)k_fft_cols<1, true>);
    set((const void*)k_fft_cols<1, false>);
    set((const void*)k_fft_cols<-1, true>);
    set((const void*)k_fft_cols<-1, false>);
    set((const void*)k_grid_rows_kb<1, true, false>);
    set((const void*)k_grid_rows_kb<1, false, false>);
    set((const void*)k_grid_rows_kb<1, true, true>);
    set((const void*)k_grid_rows_kb<1, false, true>);
    set((const void*)k_grid_half_direct);
    set((const void*)k_grid_rows_direct_fft<false>);
    set((const void*)k_grid_rows_direct_fft<true>);
    set((const void*)k_grid_rows_fft<1, true, false>);
    set((const void*)k_grid_rows_fft<1, false, false>);
    set((const void*)k_grid_rows_fft<1, true, true>);
    set((const void*)k_grid_rows_fft<1, false, true>);
    set((const void*)k_cols_half<1, true, false>);
    set((const void*)k_cols_half<1, false, false>);
    set((const void*)k_cols_half<1, true, true>);
    set((const void*)k_cols_half<1, false, true>);
    set((const void*)k_cols_half_odd<1, true, false>);
    set((const void*)k_cols_half_odd<1, false, false>);
    set((const void*)k_cols_half_odd<1, true, true>);
    set((const void*)k_cols_half_odd<1, false, true>);
    set((const void*)k_grid_tiles);
    set((const void*)k_grid_half_rows);
    set((const void*)k_plan_rows);
    set((const void*)k_plan_count);
    set((const void*)k_plan_scatter);
    return e;
}

}  // namespace blink
