// Host-side launch interface of the gfx950 kernels (blink_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "blink_divrn.hpp"
#include "blink_fft.hpp"

namespace blink {

constexpr int kMaxPixels = 8192;      // largest grid side (plan entries pack positions in 13 bits)
constexpr int kMaxBaselines = 1 << 18;  // plan packs the baseline in 18 bits
constexpr int kPlanThreads = 1024;
constexpr int kTileThreads = 256;
constexpr int kFftThreads = 256;
// Columns per block of the register kernels' column-blocked row-pass buffer.
constexpr int kInterBlockCols = 32;
constexpr int kMaxTileRows = 16;
constexpr int kMaxColWidth = 16;

// LDS bytes a workgroup may allocate dynamically (device limit), and the
// share the tile-sizing heuristics aim for.  Set once per process from the
// device properties (gfx950: 160 KiB per CU).
extern int64_t g_lds_limit;
extern int64_t g_lds_budget;

struct GatherArgs {
    const float4* vis;        // [t][ch][bl][2 x float4]
    // phase coefficient of every baseline in metres, w[bl] - (L[a2] - L[a1])
    // with each term present when its correction is on (read when
    // corrections != 0): the correction is exp(+2 pi i coef nu / c)
    const double* phase_coef;
    const double* freq_hz;
    unsigned corrections;
    int pol;
    int n_bl, n_ch;
};

// Row list of a plan, per channel: the non-empty rows 0 .. N/2 in ascending
// order (N/2 + 1 slots; the first k used when k rows are non-empty), then their
// first entries (N/2 + 2 slots: slot p < k holds row_start of the p-th
// non-empty row, every later slot the end of the entries of rows 0 .. N/2).
// Empty rows have empty entry ranges, so a tile of consecutive non-empty rows
// p .. p + R - 1 owns the contiguous entries [start[p], start[p + R]).
__host__ __device__ constexpr int row_list_len(int n) { return n + 3; }

// One image set of a batched call as the kernels see it (a table of these
// lives in device memory for the duration of the call or captured graph).
// Images of all sets that share a grid side and path form a group; a launch
// covers a contiguous range of group image indices gi, and a set owns
// gi = img_base .. img_base + n_img - 1.
struct SetDesc {
    GatherArgs g;             // visibilities, corrections, pol, n_bl, n_ch
    const float* w_m;         // inputs of the phase coefficients (g.phase_coef),
    const double* cable_m;    // which the plan launch writes to coef_out
    double* coef_out;
    const float* u_m;
    const float* v_m;
    const int* flags;
    double du, dv, min_uv;
    int n_pixels;
    int n_img;                // n_ch * n_t
    int64_t img_base;         // first group image index of the set
    int* row_start;           // plan: [n_ch][N + 1]
    int* row_list;            // plan: [n_ch][row_list_len(N)], see row_list_len
    uint32_t* entries;        // plan: [n_ch][2 n_bl]
    uint32_t* isums;          // [n_img] gridded visibilities per image; the plan's
                              // part 0 of each channel zeroes that channel's entries
    float2* images;           // [n_img][N][N]
    double* count_sums;       // optional [n_img]: isums as doubles, written by the
                              // column pass (first column block of each image)
    int uv_swap;              // transposed plan: its rows are grid columns u and an
                              // entry's position is v (rows are v otherwise)
    float2* compact;          // optional [n_img][n_bl]: each baseline's visibility
                              // after the pol combination and the phase rotation,
                              // NaN where it is not gridded (k_vis_compact); the row
                              // tiles then gather these instead of the records
    int uniform;              // "U" weighting: cells divided by their counts, so
                              // empty cells (and plan rows without entries) are NaN
    // convolutional gridding (kb_size > 0): every visibility and its conjugate
    // spread over kb_size x kb_size cells with the 1023-row kernel table
    // kb_table; the row pass then writes every plan row 0 .. N/2, and all_rows
    // tells the second pass to read them all (rows without plan entries are
    // not empty there)
    const double* kb_table;
    int kb_size;
    int all_rows;
    // the register row pass stores the row-pass buffer with plain (cached)
    // stores instead of streaming ones, so that a chunk small enough for the
    // Infinity Cache can be read back from it by the second pass
    int cached_inter;
    // windowed direct row pass (k_grid_rows_direct_win): [kUvMaxParts][2]
    // partial maxima of |u| and |v| over the unflagged baselines, written by
    // the direct set-up launch (in the set's otherwise unused row lists);
    // nullptr for other sets
    float* uv_max;
};

constexpr int kUvMaxParts = 64;

// A packed plan cell that is not gridded (flagged, NaN or out of range).
constexpr uint32_t kNoCell = 0xffffffffu;

// The set of group image gi, by value.  Kernels are built twice: for a
// group of one set (kMulti false) the descriptor is `first`, a kernel
// argument, so its fields arrive with the other arguments and add no
// dependent load at workgroup start; otherwise it is the table entry named
// by the image-to-set table.  Uniform per workgroup.
template <bool kMulti>
__device__ __forceinline__ SetDesc set_of(const SetDesc& first, const SetDesc* __restrict__ sets,
                                          const int* __restrict__ img_set, int64_t gi) {
    if constexpr (!kMulti) return first;
    else return sets[img_set[gi]];
}

struct ColArgs {
    const float2* in;          // [img][N][N], images of this launch
    float2* out;               // [img][N][N], images of this launch
    int64_t in_img_stride;     // elements
    int64_t out_img_stride;
    int64_t out_img_offset;    // global index of the launch's first image (sums, plan)
    int width;                 // W, set by the launcher
    int paired;                // also load columns c0 + N/2 .. (+W)
    int shift;                 // fft_shift on store
    const double* sums;        // scale (float)(1/sums[img]) on store, or null
    const uint32_t* isums;     // integer sums alternative, or null
    const int* row_start;      // plan rows: skip empty rows on load, or null
    int n_ch;
};

// Compute units of the current device (queried once per process: every device
// the library runs on is the same type).
int device_cu_count();

bool make_fft_plan(int n, FftPlan& plan);
// Whether the in-place (paired) shifted column pass fits in LDS for side n.
bool cols_paired_fits(int n);
int grid_rows_tile_rows(int n);
hipError_t set_kernel_attributes();

hipError_t launch_zero(void* ptr, int64_t bytes, hipStream_t stream);
hipError_t launch_phase_correct(float* vis, int n_ant, int n_ch, int n_t, const float* w_m,
                                const double* cable_m, const double* freq_hz, unsigned which,
                                hipStream_t stream);
hipError_t launch_uv_cells(const float* u_m, const float* v_m, const int* flags, int n_bl,
                           const double* freq_hz, int n_ch, double du, double dv, double min_uv,
                           int n_pixels, int* cells, hipStream_t stream);
hipError_t launch_grid_atomic(const float* vis, int n_ant, int n_ch, int n_t, const float* u_m,
                              const float* v_m, const int* flags, const double* freq_hz, double du,
                              double dv, double min_uv, int n_pixels, int pol, float* grids,
                              float* counts, hipStream_t stream);
hipError_t launch_zenith_uvw(const double* x, const double* y, const double* z, int n_ant,
                             float* u_m, float* v_m, float* w_m, hipStream_t stream);
hipError_t launch_baseline_flags(const int* ant_flags, int n_ant, int* bl_flags,
                                 hipStream_t stream);
hipError_t launch_uv_max(const float* u_m, const float* v_m, const int* flags, int n_bl,
                         double* uv_max, hipStream_t stream);
// Convolutional gridding with a kernel table of 1023 rows of kernel_size
// doubles (row = oversampled fractional cell position).
hipError_t launch_grid_kb(const float* vis, int n_ant, int n_ch, int n_t, const float* u_m,
                          const float* v_m, const int* flags, const double* freq_hz, double du,
                          double dv, double min_uv, int n_pixels, int pol, int kernel_size,
                          const double* table, float* grids, float* counts, hipStream_t stream);
hipError_t launch_plan_rows(const float* u_m, const float* v_m, const int* flags, int n_bl,
                            const double* freq_hz, int n_ch, double du, double dv, double min_uv,
                            int n_pixels, int* row_start, uint32_t* entries, hipStream_t stream);
// Row-tile gridding.  Accumulates into grids/counts, or with overwrite != 0
// writes every cell of the n_images grids (counts may then be null); valid
// visibilities are added to isums[img0 + image] when isums is given.
hipError_t launch_grid_tiles(const GatherArgs& g, const int* row_start, const uint32_t* entries,
                             int n_images, int n_pixels, float* grids, float* counts,
                             hipStream_t stream, int overwrite = 0, uint32_t* isums = nullptr,
                             int64_t img0 = 0);
hipError_t launch_uniform_weight(float* grids, const float* counts, int64_t n_cells,
                                 hipStream_t stream);
hipError_t launch_fft_rows(const float* in, float* out, const FftPlan& plan, int64_t n_rows,
                           int sign, const float* tw, hipStream_t stream);
hipError_t launch_fft_cols(const ColArgs& a, const FftPlan& plan, int n_images, int sign,
                           const float* tw, hipStream_t stream);
hipError_t launch_vector_sum(const float* values, int64_t image_size, int n_images, double* sums,
                             hipStream_t stream);
hipError_t launch_shift_norm(float* images, int nx, int ny, int n_images, const double* sums,
                             int n_sums, hipStream_t stream);
hipError_t launch_average(const float* images, int64_t image_size, int n_images, int real_only,
                          float* avg, hipStream_t stream);
// Generic batched path (any even side the FFT plan supports): row pass over
// plan rows 0 .. N/2 (grid columns u, transposed plan) into a row-major
// half-height buffer [img][N/2 + 1][N], and the second pass from it (two real
// image rows per complex FFT).
hipError_t launch_grid_rows_fft(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                                float* inter, const float* tw, hipStream_t stream);
hipError_t launch_cols_half(const SetDesc& first, const SetDesc* sets, const int* img_set,
                            int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                            const float* inter, const float* tw, hipStream_t stream);
// Row pass of convolutional (Kaiser-Bessel) gridding on the generic kernels:
// every plan row 0 .. N/2 of every image is gridded, transformed and written,
// row-major as launch_grid_rows_fft writes or, with `blocked`, in the
// column-blocked layout of the register second pass (N = 128 .. 2048); a
// tile gathers the plan entries of its rows +- K/2 and, near row N/2, of the
// rows whose Hermitian partners' footprints reach it.  The sets' kb_size must
// be > 0.
hipError_t launch_grid_rows_kb(const SetDesc& first, const SetDesc* sets, const int* img_set,
                               int n_sets, const FftPlan& plan, int64_t gi0, int n_img,
                               float* inter, const float* tw, bool blocked, hipStream_t stream);
// Gridding plans of a batched call: one unit per (set, channel); `parts`
// workgroups per unit split its baselines (at most kPlanPartMax each).  A plan
// lists the cells of rows 0 .. N/2 only (each baseline's cell and conjugate
// cell that fall there): the batched kernels grid the Hermitian half plane,
// and rows N/2+1 .. N-1 of row_start are left empty.
// Scratch: partial is [n_units][parts][n_max] ints, cells holds one packed
// cell per (unit, baseline) from cell_off on.
constexpr int kPlanPartMax = 4096;
struct PlanUnit {
    int set, ch;
    int64_t cell_off;
};
hipError_t launch_plan_units(const SetDesc* sets, const PlanUnit* units, int n_units, int parts,
                             int n_max, int* partial, uint32_t* cells, hipStream_t stream);
hipError_t launch_isums_to_double(const uint32_t* isums, double* sums, int n, hipStream_t stream);

// Split-precision DFT-as-GEMM (blink_dft_mfma.hip) over the Hermitian half
// plane: half[img][v][u] (v = 0 .. N/2, row-major) -> stage 1 (complex) ->
// inter[img][x][v] (row stride dft_inter_stride(n)) -> stage 2 (real, scaled
// by 1/sum counts, fftshifted, imaginary parts 0) -> images[img][y][x].
enum { kDftBf16 = 0, kDftFp16 = 1 };
int dft_inter_stride(int n);
hipError_t launch_grid_half_rows(const GatherArgs& g, const int* row_start, const uint32_t* entries,
                                 int n, int64_t n_img, int64_t img0, float* half,
                                 uint32_t* isums, hipStream_t stream, bool uniform = false);
hipError_t launch_dft_rows(const float* half, const uint16_t* planes1, float* inter, int n,
                           int64_t n_img, int prec, hipStream_t stream);
hipError_t launch_dft_cols(const float* inter, const uint16_t* planes2, float* images, int n,
                           int64_t n_img, int64_t img0, const uint32_t* isums,
                           double* count_sums, int prec, hipStream_t stream);
// Allows the DFT kernels' LDS-resident matrix slices (up to 160 KiB).
hipError_t set_dft_kernel_attributes();
// Split planes of D1 (stage 1) or D2 (stage 2) for side n and precision prec.
void make_dft_planes(int n, int stage, int prec, std::vector<uint16_t>& planes);
size_t dft_planes_bytes(int n, int stage);

// Register-FFT batched kernels (power-of-two sides 128..2048); the row-pass
// buffer is transposed: inter[img][u][v].
bool batch_reg_supported(int n);
// Bytes of the row-pass buffer per image of side n.
int64_t batch_inter_bytes(int n);
// Built tile sizes for side n: plan rows per row tile, output lines per second-pass block.
int batch_row_variants(int n, int* tiles, int max_tiles);
int batch_col_variants(int n, int* tiles, int max_tiles);
hipError_t set_batch_kernel_attributes();
// uniform: the "U" weighting (grid / count per cell before the transform),
// with the side's fixed uniform row tile (batch_uniform_row_tile) instead of
// `tile`.
hipError_t launch_grid_rows_reg(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                int n_sets, int n, int tile, int64_t gi0, int n_img, float* inter,
                                const float* tw, hipStream_t stream, bool uniform = false);
int batch_uniform_row_tile(int n);
// Direct row pass of dense single-timestep sets (no plan): one workgroup per
// image with its whole half plane in LDS, for the register sides
// batch_direct_supported (128, and 256 with a window of the plane: the
// sides direct_uses_window names) and the generic sides that fit (up to
// 186); direct_row_pass_fits tells which.  The kernels store the rows up to
// the highest that holds a cell and write the image's row starts (the rows
// above empty) for the second pass.  The count sums and phase coefficients
// (and, for the window, the |u|, |v| maxima) come from launch_direct_prep
// over the group's sets (max_work: the largest n_img or n_bl among them),
// launched before.
bool batch_direct_supported(int n);
bool direct_row_pass_fits(int n);
bool direct_uses_window(int n);
hipError_t launch_grid_rows_direct(const SetDesc& first, const SetDesc* sets, const int* img_set,
                                   int n_sets, int n, int64_t gi0, int n_img, float* inter,
                                   const float* tw, hipStream_t stream);
hipError_t launch_grid_rows_direct_fft(const SetDesc& first, const SetDesc* sets,
                                       const int* img_set, int n_sets, const FftPlan& plan,
                                       int64_t gi0, int n_img, float* inter, const float* tw,
                                       hipStream_t stream);
hipError_t launch_direct_prep(const SetDesc* sets, int n_sets, int max_work, hipStream_t stream);
// The DFT transform's half plane by the direct row pass (rows v, sd.uv_swap
// 0), for the images img0 .. img0 + n_img - 1 of one set, at the even sides
// whose half plane fits in LDS (direct_half_plane_fits).
bool direct_half_plane_fits(int n);
hipError_t launch_grid_half_direct(const SetDesc& sd, int64_t n_img, int64_t img0, float* half,
                                   hipStream_t stream);
hipError_t launch_cols_reg(const SetDesc& first, const SetDesc* sets, const int* img_set,
                           int n_sets, int n, int tile, int64_t gi0, int n_img,
                           const float* inter, const float* tw, hipStream_t stream);
// Mean of a set's images fused into the batched path: adds the set's images
// img0 .. img0 + n_img - 1 (row-pass buffer slots li0 ..) into acc, each scaled
// by (float)(1 / its count sum), in image order (accumulate = false starts
// from zero and writes meta: [0] = sd.n_img, [1 .. N+1] a row start with every
// plan row non-empty); then a second pass over acc with a descriptor whose
// isums / row_start point at meta and images at the average gives the mean
// image.  Also writes the count sums of the images when sd.images is null.
hipError_t launch_average_rows(const SetDesc& sd, int64_t img0, int n_img, int64_t li0,
                               const float* inter, float* acc, bool accumulate, unsigned* meta,
                               hipStream_t stream);
// Fills sd.compact for all n_img images of the set (from the visibilities and
// the set's own inputs only: it may run beside the plan).
hipError_t launch_vis_compact(const SetDesc& sd, int64_t n_img, hipStream_t stream);
// The same for the n_img images of a group of several sets (device set table
// and image -> set table; sets without a compact buffer are skipped).
hipError_t launch_vis_compact_sets(const SetDesc* sets, const int* img_set, int64_t n_img,
                                   int max_bl, hipStream_t stream);

}  // namespace blink
