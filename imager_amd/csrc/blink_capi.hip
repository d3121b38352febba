// C ABI of libblink_mi355x.so: argument checking, context state (twiddle
// tables, scratch buffers, captured graphs) and the orchestration of the
// kernels in blink_kernels.hip.
#include "../../include/blink_mi355x.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "blink_kernels.hpp"

using namespace blink;

// The fixed device region of blink_image_batch_noalloc.  In host code its
// name stands for the array, not for its device address: host code uses the
// address hipGetSymbolAddress gives each context at blink_create
// (blink_ctx::noalloc_base), never the name.
static __device__ __align__(256) unsigned char scratch[BLINK_NOALLOC_SCRATCH_BYTES];

namespace {

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct GraphEntry {
    std::string call;  // the call's key without its buffer addresses (graph_key)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* blob = nullptr;  // device copy of the call's descriptor tables
    hipEvent_t done = nullptr;  // recorded after every replay
    std::vector<const void*> ptrs;  // vis, images, count_sums, average of every set at capture
    std::vector<unsigned char> host_blob;  // source of the last descriptor refresh
    uint64_t last_use = 0;
};

void release_graph(GraphEntry& ge) {
    if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
    if (ge.graph) (void)hipGraphDestroy(ge.graph);
    if (ge.blob) (void)hipFree(ge.blob);
    if (ge.done) (void)hipEventDestroy(ge.done);
    ge = GraphEntry{};
}

// Captured graphs kept per context; the least recently used is dropped
// beyond this.
constexpr size_t kMaxGraphs = 16;
// Graphs kept per call (the same set fields) with different buffer addresses,
// so that a caller alternating between two output buffers replays both
// without updating either; a further address set re-points the least
// recently used of them.
constexpr int kBufferVariants = 2;

}  // namespace

static constexpr int kPassEvents = 16;

struct blink_ctx {
    int device = 0;
    hipDeviceProp_t props{};
    std::string error;
    std::map<int, float*> twiddles;               // N -> device table of N float2
    std::map<int, uint16_t*> dft_planes;          // (N, stage, precision) -> split DFT planes
    std::map<std::pair<int, int>, double*> kb_tables;  // (size, mode) -> gridding kernels
    std::map<std::string, DeviceBuffer> scratch;  // named grow-only buffers
    std::map<std::string, GraphEntry> graphs;     // descriptor bytes -> graph
    uint64_t graph_clock = 0;
    hipStream_t capture_stream = nullptr;
    // a call with several grid sides runs the largest side's group on the
    // call's stream and the others beside it on aux_stream
    hipStream_t aux_stream = nullptr;
    hipEvent_t fork_event = nullptr, join_event = nullptr;
    // the visibility compaction runs on aux_stream beside the plan
    hipEvent_t compact_fork = nullptr, compact_join = nullptr;
    // chunked groups: the second passes run on pass_stream, one chunk behind
    // the row passes; pass_ev[2k] marks chunk k's row pass done, pass_ev[2k+1]
    // its second pass (k modulo kPassEvents)
    hipStream_t pass_stream = nullptr;
    hipEvent_t pass_fork = nullptr, pass_join = nullptr;
    hipEvent_t pass_ev[2 * kPassEvents] = {};
    // blink_image_batch_noalloc: the device address of this GPU's fixed
    // region (null when it could not be found), the host source of the last
    // descriptor copy and an event recorded after that copy
    char* noalloc_base = nullptr;
    hipError_t noalloc_status = hipSuccess;
    std::vector<unsigned char> noalloc_blob;
    hipEvent_t noalloc_copied = nullptr;
    bool use_graphs = true;
    bool autotune = true;
    // tile choice of the batched kernels, {row tile, column block}: pinned
    // per grid side (blink_set_batch_tiles), or tuned per (grid side, batch
    // class: images per launch, in factors of 4), and the last tuned per side
    std::map<int, std::pair<int, int>> pinned_tiles;
    std::map<std::pair<int, int>, std::pair<int, int>> tuned_tiles;
    std::map<int, std::pair<int, int>> last_tuned;

    // measurement hooks: event pairs around every launch while enabled
    struct ProfRec {
        const char* name;
        hipEvent_t a, b;
    };
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<ProfRec> prof_pending;
    std::map<std::string, std::pair<int64_t, double>> prof_totals;

    int fail(int code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        error = buf;
        return code;
    }
};

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ctx->fail(BLINK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_ARG(cond, ...)                                                \
    do {                                                                    \
        if (!(cond)) return ctx->fail(BLINK_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

namespace {

hipEvent_t prof_take(blink_ctx* ctx) {
    hipEvent_t e = nullptr;
    if (!ctx->ev_pool.empty()) {
        e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

hipEvent_t prof_open(blink_ctx* ctx, hipStream_t s) {
    if (!ctx->prof_on) return nullptr;
    hipEvent_t a = prof_take(ctx);
    if (a && hipEventRecord(a, s) != hipSuccess) {
        ctx->ev_pool.push_back(a);
        return nullptr;
    }
    return a;
}

void prof_close(blink_ctx* ctx, hipStream_t s, const char* name, hipEvent_t a) {
    if (!a) return;
    hipEvent_t b = prof_take(ctx);
    if (!b || hipEventRecord(b, s) != hipSuccess) {
        ctx->ev_pool.push_back(a);
        if (b) ctx->ev_pool.push_back(b);
        return;
    }
    ctx->prof_pending.push_back({name, a, b});
}

int prof_collect(blink_ctx* ctx) {
    for (auto& r : ctx->prof_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& t = ctx->prof_totals[r.name];
            t.first += 1;
            t.second += ms;
        }
        ctx->ev_pool.push_back(r.a);
        ctx->ev_pool.push_back(r.b);
    }
    ctx->prof_pending.clear();
    return BLINK_OK;
}

}  // namespace

// Launch wrapper: checks the launch and, while profiling, brackets it with
// events on its stream.
#define LAUNCH(name, stream, expr)                                     \
    do {                                                               \
        hipEvent_t pa_ = prof_open(ctx, (stream));                     \
        HIP_TRY(expr);                                                 \
        prof_close(ctx, (stream), (name), pa_);                        \
    } while (0)

namespace {

int set_device(blink_ctx* ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return BLINK_OK;
}

// Streams and events of a context, each destroyed once (creation failures
// leave the later handles null).
void destroy_streams(blink_ctx* ctx) {
    for (hipStream_t* s : {&ctx->capture_stream, &ctx->aux_stream, &ctx->pass_stream}) {
        if (*s) (void)hipStreamDestroy(*s);
        *s = nullptr;
    }
    for (hipEvent_t* e : {&ctx->fork_event, &ctx->join_event, &ctx->compact_fork,
                          &ctx->compact_join, &ctx->noalloc_copied, &ctx->pass_fork,
                          &ctx->pass_join}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    for (hipEvent_t& e : ctx->pass_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

void drop_graphs(blink_ctx* ctx) {
    if (ctx->graphs.empty()) return;
    (void)hipDeviceSynchronize();  // a queued replay may still read a descriptor blob
    for (auto& kv : ctx->graphs) release_graph(kv.second);
    ctx->graphs.clear();
}

// Grow-only named scratch buffer.  Growing one invalidates every captured
// graph (they hold the old address), so those are dropped first.
int ctx_scratch(blink_ctx* ctx, const char* name, size_t bytes, void** out) {
    DeviceBuffer& b = ctx->scratch[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            HIP_TRY(hipDeviceSynchronize());
            drop_graphs(ctx);
            HIP_TRY(hipFree(b.ptr));
        }
        b.ptr = nullptr;
        b.bytes = 0;
        HIP_TRY(hipMalloc(&b.ptr, bytes));
        b.bytes = bytes;
    }
    *out = b.ptr;
    return BLINK_OK;
}

// tw[m] = exp(+2 pi i m / N), m < N, computed in double, rounded to FP32;
// followed by lo[j] = tw[j] and hi[j] = exp(+2 pi i 64 j / N), j < 64 (the
// two-level tables of the N = 4096 kernels).
int twiddles(blink_ctx* ctx, int n, const float** out) {
    auto it = ctx->twiddles.find(n);
    if (it != ctx->twiddles.end()) {
        *out = it->second;
        return BLINK_OK;
    }
    std::vector<float> h(2 * ((size_t)n + 128));
    auto put = [&](size_t at, double turns) {
        const double a = 2.0 * M_PI * turns;
        h[2 * at] = (float)std::cos(a);
        h[2 * at + 1] = (float)std::sin(a);
    };
    for (int m = 0; m < n; m++) put(m, (double)m / (double)n);
    for (int j = 0; j < 64; j++) {
        put((size_t)n + j, (double)j / (double)n);
        put((size_t)n + 64 + j, (double)(64 * j % n) / (double)n);
    }
    float* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(float) * h.size()));
    HIP_TRY(hipMemcpy(d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    ctx->twiddles[n] = d;
    *out = d;
    return BLINK_OK;
}

// Grid sides whose FFT tables blink_create builds: the sides the
// no-allocation entry point images.
constexpr int kPreparedSides[] = {128, 180, 256, 512, 1024, 2048, 4096};
static bool prepared_side(int n) {
    return std::find(std::begin(kPreparedSides), std::end(kPreparedSides), n) !=
           std::end(kPreparedSides);
}

static bool is_dft(int transform) {
    return transform == BLINK_TRANSFORM_DFT_BF16X3 || transform == BLINK_TRANSFORM_DFT_FP16X3;
}
static int dft_prec(int transform) {
    return transform == BLINK_TRANSFORM_DFT_FP16X3 ? kDftFp16 : kDftBf16;
}

// Split DFT matrix of side n for one stage and precision
// (blink_dft_mfma.hip), built once per context.
int dft_planes(blink_ctx* ctx, int n, int stage, int prec, const uint16_t** out) {
    const int key = n * 8 + stage * 2 + prec;
    auto it = ctx->dft_planes.find(key);
    if (it != ctx->dft_planes.end()) {
        *out = it->second;
        return BLINK_OK;
    }
    std::vector<uint16_t> h;
    make_dft_planes(n, stage, prec, h);
    uint16_t* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(uint16_t) * h.size()));
    HIP_TRY(hipMemcpy(d, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice));
    ctx->dft_planes[key] = d;
    *out = d;
    return BLINK_OK;
}

// Row-pass bytes per image of the DFT transform: the half plane and the
// stage-1 product.
static int64_t dft_inter_bytes(int64_t n) {
    return 8 * ((n / 2 + 1) * n + n * (int64_t)dft_inter_stride((int)n));
}

int check_fft_size(blink_ctx* ctx, int n, FftPlan& plan) {
    if (!make_fft_plan(n, plan))
        return ctx->fail(BLINK_ERR_UNSUPPORTED,
                         "grid side %d is not supported (need 2 <= N <= %d with prime factors "
                         "2, 3, 5, 7, 11, 13)",
                         n, kMaxPixels);
    return BLINK_OK;
}

// Rows then columns, optionally with the shift+scale epilogue on the column
// store.  in -> out (in == out allowed).
int fft2d(blink_ctx* ctx, const float* in, float* out, int n, int batch, int sign, bool shift,
          const double* sums, hipStream_t stream) {
    FftPlan plan;
    int rc = check_fft_size(ctx, n, plan);
    if (rc) return rc;
    const float* tw;
    if ((rc = twiddles(ctx, n, &tw))) return rc;
    LAUNCH("fft_rows", stream, launch_fft_rows(in, out, plan, (int64_t)batch * n, sign, tw, stream));
    // The shifted store is fused into the in-place column pass when a
    // workgroup can hold a column block and its shift partner; otherwise a
    // separate shift+norm pass follows.
    const bool fused = shift && cols_paired_fits(n);
    ColArgs a{};
    a.in = (const float2*)out;
    a.out = (float2*)out;
    a.in_img_stride = a.out_img_stride = (int64_t)n * n;
    a.out_img_offset = 0;
    a.paired = fused ? 1 : 0;
    a.shift = fused ? 1 : 0;
    a.sums = fused ? sums : nullptr;
    a.n_ch = 1;
    LAUNCH("fft_cols", stream, launch_fft_cols(a, plan, batch, sign, tw, stream));
    if (shift && !fused) LAUNCH("shift_norm", stream, launch_shift_norm(out, n, n, batch, sums, batch, stream));
    return BLINK_OK;
}

int check_vis_shape(blink_ctx* ctx, int n_ant, int n_ch, int n_t) {
    CHECK_ARG(n_ant >= 1 && n_ch >= 0 && n_t >= 0, "invalid shape n_ant=%d n_ch=%d n_t=%d", n_ant,
              n_ch, n_t);
    const int64_t n_bl = (int64_t)n_ant * (n_ant + 1) / 2;
    CHECK_ARG(n_bl < kMaxBaselines, "n_ant=%d gives %lld baselines, above the supported %d", n_ant,
              (long long)n_bl, kMaxBaselines);
    return BLINK_OK;
}

}  // namespace

// ==================================================================== C ABI

extern "C" {

BLINK_API const char* blink_version(void) { return "blink-mi355x 0.1 (gfx950)"; }

BLINK_API int blink_fft_supported(int n_pixels) {
    FftPlan p;
    return make_fft_plan(n_pixels, p) ? 1 : 0;
}

BLINK_API int blink_create(int device, blink_ctx** out) {
    if (!out) return BLINK_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return BLINK_ERR_HIP;
    if (device < 0 || device >= n_dev) return BLINK_ERR_INVALID_ARGUMENT;
    blink_ctx* ctx = new blink_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipGetDeviceProperties(&ctx->props, device) != hipSuccess) {
        delete ctx;
        return BLINK_ERR_HIP;
    }
    // Dynamic LDS per workgroup: gfx950 has 160 KiB per CU; keep 16 KiB for
    // the kernels' static LDS and headroom.
    int64_t lds = (int64_t)ctx->props.sharedMemPerBlock;
    if (lds <= 0) lds = 64 * 1024;
    g_lds_limit = std::min<int64_t>(lds, 160 * 1024) - 4 * 1024;
    g_lds_budget = g_lds_limit - 12 * 1024;
    if (set_kernel_attributes() != hipSuccess || set_batch_kernel_attributes() != hipSuccess ||
        set_dft_kernel_attributes() != hipSuccess) {
        (void)hipGetLastError();
    }
    if (hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork_event, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join_event, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->compact_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->compact_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->noalloc_copied, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->pass_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->pass_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->pass_join, hipEventDisableTiming) != hipSuccess ||
        !std::all_of(std::begin(ctx->pass_ev), std::end(ctx->pass_ev), [](hipEvent_t& e) {
            return hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        })) {
        destroy_streams(ctx);
        delete ctx;
        return BLINK_ERR_HIP;
    }
    // this GPU's copy of the no-allocation region (its own address per device;
    // the device is current here)
    void* region = nullptr;
    ctx->noalloc_status = hipGetSymbolAddress(&region, HIP_SYMBOL(scratch));
    if (ctx->noalloc_status == hipSuccess && region) {
        ctx->noalloc_base = (char*)region;
    } else {
        (void)hipGetLastError();
        ctx->noalloc_base = nullptr;
    }
    const char* ng = getenv("BLINK_NO_GRAPH");
    ctx->use_graphs = !(ng && ng[0] == '1');
    const char* at = getenv("BLINK_AUTOTUNE");
    ctx->autotune = !(at && at[0] == '0');
    // Twiddle tables of the common grid sides, so that later calls do not
    // allocate.
    for (int n : kPreparedSides) {
        const float* tw;
        if (twiddles(ctx, n, &tw) != BLINK_OK) {
            blink_destroy(ctx);
            return BLINK_ERR_HIP;
        }
    }
    *out = ctx;
    return BLINK_OK;
}

BLINK_API void blink_destroy(blink_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    blink_release_cache(ctx);
    for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
    for (auto& kv : ctx->dft_planes) (void)hipFree(kv.second);
    for (auto& kv : ctx->kb_tables) (void)hipFree(kv.second);
    prof_collect(ctx);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    destroy_streams(ctx);
    delete ctx;
}

BLINK_API const char* blink_last_error(const blink_ctx* ctx) {
    return ctx ? ctx->error.c_str() : "null context";
}

BLINK_API int blink_release_cache(blink_ctx* ctx) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    drop_graphs(ctx);
    for (auto& kv : ctx->scratch)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    ctx->scratch.clear();
    return BLINK_OK;
}

// ------------------------------------------------------------------ (a)

static int phase_correct(blink_ctx* ctx, float* vis, int n_ant, int n_ch, int n_t,
                         const float* w_m, const double* cable_m, const double* freq_hz,
                         unsigned which, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    int rc = check_vis_shape(ctx, n_ant, n_ch, n_t);
    if (rc) return rc;
    CHECK_ARG(vis && freq_hz, "vis and freq_hz must be non-null");
    CHECK_ARG(!(which & 1u) || w_m, "w_m must be non-null");
    CHECK_ARG(!(which & 2u) || cable_m, "cable_m must be non-null");
    if ((rc = set_device(ctx))) return rc;
    LAUNCH("phase_correct", (hipStream_t)stream,
           launch_phase_correct(vis, n_ant, n_ch, n_t, w_m, cable_m, freq_hz, which,
                                (hipStream_t)stream));
    return BLINK_OK;
}

BLINK_API int blink_apply_geometric_corrections(blink_ctx* ctx, float* vis, int n_ant, int n_ch,
                                                int n_t, const float* w_m, const double* freq_hz,
                                                void* stream) {
    return phase_correct(ctx, vis, n_ant, n_ch, n_t, w_m, nullptr, freq_hz, 1u, stream);
}

BLINK_API int blink_apply_cable_corrections(blink_ctx* ctx, float* vis, int n_ant, int n_ch,
                                            int n_t, const double* cable_m, const double* freq_hz,
                                            void* stream) {
    return phase_correct(ctx, vis, n_ant, n_ch, n_t, nullptr, cable_m, freq_hz, 2u, stream);
}

// ------------------------------------------------------------------ (b)

BLINK_API int blink_uv_cells(blink_ctx* ctx, const float* u_m, const float* v_m,
                             const int* bl_flags, int n_bl, const double* freq_hz, int n_ch,
                             double delta_u, double delta_v, double min_uv, int n_pixels,
                             int* cells, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(n_bl >= 0 && n_ch >= 0 && n_pixels >= 1, "invalid sizes");
    CHECK_ARG(u_m && v_m && bl_flags && freq_hz && cells, "null pointer argument");
    int rc = set_device(ctx);
    if (rc) return rc;
    LAUNCH("uv_cells", (hipStream_t)stream,
           launch_uv_cells(u_m, v_m, bl_flags, n_bl, freq_hz, n_ch, delta_u, delta_v, min_uv,
                           n_pixels, cells, (hipStream_t)stream));
    return BLINK_OK;
}

// Builds the per-channel row plan into context scratch.
static int build_plan(blink_ctx* ctx, const float* u_m, const float* v_m, const int* flags,
                      int n_bl, const double* freq_hz, int n_ch, double du, double dv,
                      double min_uv, int n_pixels, int** row_start, uint32_t** entries,
                      hipStream_t stream, const char* tag) {
    void *rs, *en;
    int rc;
    std::string a = std::string("plan_rows") + tag, b = std::string("plan_entries") + tag;
    if ((rc = ctx_scratch(ctx, a.c_str(), sizeof(int) * (size_t)n_ch * (n_pixels + 1), &rs))) return rc;
    if ((rc = ctx_scratch(ctx, b.c_str(), sizeof(uint32_t) * 2 * (size_t)n_ch * std::max(n_bl, 1), &en)))
        return rc;
    LAUNCH("plan_rows", stream,
           launch_plan_rows(u_m, v_m, flags, n_bl, freq_hz, n_ch, du, dv, min_uv, n_pixels,
                            (int*)rs, (uint32_t*)en, stream));
    *row_start = (int*)rs;
    *entries = (uint32_t*)en;
    return BLINK_OK;
}

BLINK_API int blink_gridding(blink_ctx* ctx, const float* vis, int n_ant, int n_ch, int n_t,
                             const float* u_m, const float* v_m, const int* bl_flags,
                             const double* freq_hz, double delta_u, double delta_v, double min_uv,
                             int n_pixels, int pol, int weighting, int algo, float* grids,
                             float* counts, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    int rc = check_vis_shape(ctx, n_ant, n_ch, n_t);
    if (rc) return rc;
    CHECK_ARG(vis && u_m && v_m && bl_flags && freq_hz && grids && counts, "null pointer argument");
    CHECK_ARG(pol >= 0 && pol <= 3, "pol must be 0..3 (XX, YY, I, V), got %d", pol);
    CHECK_ARG(weighting == BLINK_WEIGHTING_NATURAL || weighting == BLINK_WEIGHTING_UNIFORM,
              "weighting must be NATURAL or UNIFORM");
    CHECK_ARG(algo == BLINK_GRID_TILED || algo == BLINK_GRID_ATOMIC, "unknown gridding algorithm");
    CHECK_ARG(n_pixels >= 1 && n_pixels <= (algo == BLINK_GRID_TILED ? kMaxPixels : 65535),
              "n_pixels=%d out of range", n_pixels);
    if ((rc = set_device(ctx))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int n_bl = n_ant * (n_ant + 1) / 2;
    const int n_img = n_ch * n_t;
    if (algo == BLINK_GRID_ATOMIC) {
        LAUNCH("grid_atomic", s,
               launch_grid_atomic(vis, n_ant, n_ch, n_t, u_m, v_m, bl_flags, freq_hz, delta_u,
                                  delta_v, min_uv, n_pixels, pol, grids, counts, s));
    } else {
        int* rs;
        uint32_t* en;
        if ((rc = build_plan(ctx, u_m, v_m, bl_flags, n_bl, freq_hz, n_ch, delta_u, delta_v, min_uv,
                             n_pixels, &rs, &en, s, "")))
            return rc;
        GatherArgs g{};
        g.vis = (const float4*)vis;
        g.freq_hz = freq_hz;
        g.corrections = 0;
        g.pol = pol;
        g.n_bl = n_bl;
        g.n_ch = n_ch;
        LAUNCH("grid_tiles", s, launch_grid_tiles(g, rs, en, n_img, n_pixels, grids, counts, s));
    }
    if (weighting == BLINK_WEIGHTING_UNIFORM)
        LAUNCH("uniform_weight", s,
               launch_uniform_weight(grids, counts, (int64_t)n_img * n_pixels * n_pixels, s));
    return BLINK_OK;
}

// ------------------------------------------------------------ array geometry

BLINK_API int blink_zenith_uvw(blink_ctx* ctx, const double* x, const double* y,
                               const double* z, int n_ant, float* u_m, float* v_m, float* w_m,
                               void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(x && y && z && u_m && v_m && w_m, "null pointer argument");
    CHECK_ARG(n_ant >= 1 && (int64_t)n_ant * (n_ant + 1) / 2 <= kMaxBaselines,
              "n_ant=%d out of range", n_ant);
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("zenith_uvw", s, launch_zenith_uvw(x, y, z, n_ant, u_m, v_m, w_m, s));
    return BLINK_OK;
}

BLINK_API int blink_baseline_flags(blink_ctx* ctx, const int* ant_flags, int n_ant,
                                   int* bl_flags, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(ant_flags && bl_flags, "null pointer argument");
    CHECK_ARG(n_ant >= 1 && (int64_t)n_ant * (n_ant + 1) / 2 <= kMaxBaselines,
              "n_ant=%d out of range", n_ant);
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("baseline_flags", s, launch_baseline_flags(ant_flags, n_ant, bl_flags, s));
    return BLINK_OK;
}

BLINK_API int blink_uv_max(blink_ctx* ctx, const float* u_m, const float* v_m,
                           const int* bl_flags, int n_bl, double* uv_max, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(u_m && v_m && bl_flags && uv_max, "null pointer argument");
    CHECK_ARG(n_bl >= 0 && n_bl <= kMaxBaselines, "n_bl=%d out of range", n_bl);
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("uv_max", s, launch_uv_max(u_m, v_m, bl_flags, n_bl, uv_max, s));
    return BLINK_OK;
}

// Modified Bessel function I0 by its power series sum (x/2)^2m / (m!)^2,
// summed until a term falls to `rel` of the sum (the stopping rule of the
// reference's bessel0, src/pacer_imager.cpp:566-578).
static double bessel_i0(double x, double rel) {
    double sum = 1.0, term = 1.0, d = 0.0;
    do {
        d += 2.0;
        term *= x * x / (d * d);
        sum += term;
    } while (term > sum * rel);
    return sum;
}

// Gridding kernel table: row r (oversampled fractional position), column c
// (footprint cell) = the reference's _griddingKernels[r][c], built by
// makeKaiserBesselKernel (alpha 8.6, oversampling 1023, optional sinc) and
// makeKernels (src/pacer_imager.cpp:580-651).  The 1-D kernel of K * 1023
// samples is filled from its middle outwards and mirrored; for even K the
// reference also writes one value past its end, which is not kept here.
static int kb_table(blink_ctx* ctx, int K, int mode, const double** out) {
    auto key = std::make_pair(K, mode);
    auto it = ctx->kb_tables.find(key);
    if (it != ctx->kb_tables.end()) {
        *out = it->second;
        return BLINK_OK;
    }
    const int os = 1023;
    const double alpha = 8.6;
    const size_t n = (size_t)K * os, mid = n / 2;
    std::vector<double> k1(n, 0.0);
    const double ratio = 1.0 / (double)os;
    const double norm = (double)os / bessel_i0(alpha, 1e-8);
    for (size_t i = 0; i <= mid; i++) {
        const double x = (double)i;
        const double sinc = i == 0 ? ratio
                            : (mode == BLINK_KERNEL_KAISER_BESSEL ? std::sin(M_PI * ratio * x) / (M_PI * x)
                                                                  : ratio);
        const double t = (double)i / (double)mid;
        if (mid + i < n) k1[mid + i] = sinc * bessel_i0(alpha * std::sqrt(1.0 - t * t), 1e-8) * norm;
    }
    for (size_t i = 0; i < mid; i++) k1[i] = k1[n - 1 - i];
    std::vector<double> table((size_t)os * K);
    for (int i = 0; i < os; i++)
        for (int c = 0; c < K; c++) table[(size_t)(os - 1 - i) * K + c] = k1[(size_t)c * os + i];
    double* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(double) * table.size()));
    HIP_TRY(hipMemcpy(d, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice));
    ctx->kb_tables[key] = d;
    *out = d;
    return BLINK_OK;
}

BLINK_API int blink_gridding_kernel(blink_ctx* ctx, const float* vis, int n_ant, int n_ch,
                                    int n_t, const float* u_m, const float* v_m,
                                    const int* bl_flags, const double* freq_hz, double delta_u,
                                    double delta_v, double min_uv, int n_pixels, int pol,
                                    int weighting, int kernel_size, int kernel_mode,
                                    float* grids, float* counts, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    int rc = check_vis_shape(ctx, n_ant, n_ch, n_t);
    if (rc) return rc;
    CHECK_ARG(vis && u_m && v_m && bl_flags && freq_hz && grids && counts, "null pointer argument");
    CHECK_ARG(pol >= 0 && pol <= 3, "pol must be 0..3 (XX, YY, I, V), got %d", pol);
    CHECK_ARG(weighting == BLINK_WEIGHTING_NATURAL || weighting == BLINK_WEIGHTING_UNIFORM,
              "weighting must be NATURAL or UNIFORM");
    CHECK_ARG(kernel_size >= 1 && kernel_size <= BLINK_MAX_KERNEL_SIZE,
              "kernel_size must be 1..%d, got %d", BLINK_MAX_KERNEL_SIZE, kernel_size);
    CHECK_ARG(kernel_mode == BLINK_KERNEL_KAISER_BESSEL ||
                  kernel_mode == BLINK_KERNEL_KAISER_BESSEL_NO_SINC,
              "unknown kernel mode %d", kernel_mode);
    CHECK_ARG(n_pixels >= 1 && n_pixels <= 65535, "n_pixels=%d out of range", n_pixels);
    if ((rc = set_device(ctx))) return rc;
    const double* table;
    if ((rc = kb_table(ctx, kernel_size, kernel_mode, &table))) return rc;
    hipStream_t s = (hipStream_t)stream;
    LAUNCH("grid_kernel", s,
           launch_grid_kb(vis, n_ant, n_ch, n_t, u_m, v_m, bl_flags, freq_hz, delta_u, delta_v,
                          min_uv, n_pixels, pol, kernel_size, table, grids, counts, s));
    if (weighting == BLINK_WEIGHTING_UNIFORM)
        LAUNCH("uniform_weight", s,
               launch_uniform_weight(grids, counts, (int64_t)n_ch * n_t * n_pixels * n_pixels, s));
    return BLINK_OK;
}

// ------------------------------------------------------------------ (c)

BLINK_API int blink_fft2d_c2c(blink_ctx* ctx, const float* in, float* out, int n_pixels,
                              int batch, int direction, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(in && out, "null pointer argument");
    CHECK_ARG(batch >= 0, "batch must be >= 0");
    CHECK_ARG(direction == BLINK_FFT_BACKWARD || direction == BLINK_FFT_FORWARD,
              "direction must be BLINK_FFT_BACKWARD (+1) or BLINK_FFT_FORWARD (-1)");
    int rc = set_device(ctx);
    if (rc) return rc;
    if (batch == 0) return BLINK_OK;
    return fft2d(ctx, in, out, n_pixels, batch, direction, false, nullptr, (hipStream_t)stream);
}

// ------------------------------------------------------------------ (d)

BLINK_API int blink_vector_sum(blink_ctx* ctx, const float* values, int64_t image_size,
                               int n_images, double* sums, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(values && sums, "null pointer argument");
    CHECK_ARG(image_size >= 0 && n_images >= 0, "invalid sizes");
    int rc = set_device(ctx);
    if (rc) return rc;
    LAUNCH("vector_sum", (hipStream_t)stream,
           launch_vector_sum(values, image_size, n_images, sums, (hipStream_t)stream));
    return BLINK_OK;
}

BLINK_API int blink_fft_shift_and_norm(blink_ctx* ctx, float* images, int nx, int ny,
                                       int n_images, const double* sums, int n_sums,
                                       void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(images, "null pointer argument");
    CHECK_ARG(nx >= 1 && ny >= 1 && n_images >= 0, "invalid sizes nx=%d ny=%d", nx, ny);
    CHECK_ARG(!sums || n_sums >= 1, "n_sums must be >= 1 when sums are given");
    int rc = set_device(ctx);
    if (rc) return rc;
    LAUNCH("shift_norm", (hipStream_t)stream,
           launch_shift_norm(images, nx, ny, n_images, sums, n_sums, (hipStream_t)stream));
    return BLINK_OK;
}

BLINK_API int blink_image(blink_ctx* ctx, float* grids, float* counts, int n_pixels, int n_images,
                          float* images, int zero_grids, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(grids && counts && images, "grids are not initialised and cannot be imaged");
    CHECK_ARG(n_images >= 0, "n_images must be >= 0");
    CHECK_ARG(images != grids, "images must not alias grids");
    int rc = set_device(ctx);
    if (rc) return rc;
    FftPlan plan;
    if ((rc = check_fft_size(ctx, n_pixels, plan))) return rc;
    if (n_images == 0) return BLINK_OK;
    hipStream_t s = (hipStream_t)stream;
    void* sums;
    if ((rc = ctx_scratch(ctx, "image_sums", sizeof(double) * (size_t)n_images, &sums))) return rc;
    const int64_t gs = (int64_t)n_pixels * n_pixels;
    LAUNCH("vector_sum", s, launch_vector_sum(counts, gs, n_images, (double*)sums, s));
    if ((rc = fft2d(ctx, grids, images, n_pixels, n_images, BLINK_FFT_BACKWARD, true,
                    (const double*)sums, s)))
        return rc;
    if (zero_grids) {
        LAUNCH("zero", s, launch_zero(grids, (int64_t)sizeof(float) * 2 * gs * n_images, s));
        LAUNCH("zero", s, launch_zero(counts, (int64_t)sizeof(float) * gs * n_images, s));
    }
    return BLINK_OK;
}

BLINK_API int blink_image_averaging(blink_ctx* ctx, const float* images, int64_t image_size,
                                    int n_images, int real_only, float* avg, void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(images && avg, "null pointer argument");
    CHECK_ARG(image_size >= 0 && n_images >= 1, "invalid sizes");
    int rc = set_device(ctx);
    if (rc) return rc;
    LAUNCH("average", (hipStream_t)stream,
           launch_average(images, image_size, n_images, real_only, avg, (hipStream_t)stream));
    return BLINK_OK;
}

// ------------------------------------------------------------ batched
//
// A call is organised in groups: the sets that share a grid side and a
// transform.  The FFT groups run one row launch and one column launch per
// chunk of group images, whatever the number of sets in the group; the
// kernels find an image's set through a table of set descriptors (SetDesc)
// and an image-to-set table, uploaded once per captured graph.  Gridding
// plans are built for every (set, channel) of the call in three launches.

namespace {

struct SetScratch {
    int* row_start;
    int* row_list;
    uint32_t* entries;
    uint32_t* isums;
    double* coef;  // [n_bl] phase coefficients
    int chunk;  // DFT sets: images per launch triple
    float2* compact = nullptr;  // [n_img][n_bl] compacted visibilities, or none
    // FFT sets with an average: the weighted sum of the row-pass buffers and
    // the average pass's count and row start (launch_average_rows)
    float* avg_acc = nullptr;
    unsigned* avg_meta = nullptr;
    const double* kb_table = nullptr;  // convolutional gridding kernels
};

struct BatchGroup {
    int n = 0;
    int transform = BLINK_TRANSFORM_FFT;
    bool uniform = false;        // "U" weighting
    int kb_size = 0, kb_mode = 0;  // convolutional gridding (generic kernels)
    bool reg = false;            // register-FFT kernels (power-of-two sides 128..4096)
    bool kb_blocked = false;     // convolutional rows into the register second pass
    std::vector<int> sets;       // call indices of the sets, in descriptor order
    int first_desc = 0;          // index of the group's first SetDesc in the table
    int64_t n_img = 0;
    int chunk = 1;               // FFT groups: images per row/column launch pair
    size_t img_set_off = 0;      // blob offset of the image -> set table (2+ sets)
    int compact_max_bl = 0;      // largest n_bl of its sets with compacted visibilities
    // second passes on the context's pass stream, a chunk behind the row
    // passes, with two row-pass buffers (inter, inter2) used in turn
    bool overlap = false;
    // the register row pass stores its chunks with plain (cached) stores
    bool cached_inter = false;
    // direct row pass (no plan): dense sets with a plan per image
    bool direct = false;
    int direct_work = 0;  // the largest n_img or n_bl of its sets (launch_direct_prep)
};

struct BatchPlan {
    std::vector<BatchGroup> groups;
    std::vector<SetScratch> sc;  // per call index
    std::vector<unsigned char> blob;
    size_t units_off = 0;
    int n_units = 0, parts = 1, n_max = 0;
    int* partial = nullptr;
    uint32_t* cells = nullptr;
    float* inter = nullptr;
    // several groups: the group of index `main` (most image bytes) uses inter
    // on the call's stream, the others inter_aux on the aux stream
    float* inter_aux = nullptr;
    float* inter2 = nullptr;  // the main group's second buffer (overlap)
    int main = 0;
    // environment switches read for this call (part of the graph key)
    int64_t inter_budget = 0;
    bool compact_on = true, compact_dense = false;
    int cached_mode = -1;       // BLINK_INTER_CACHED: 0 never, 1 always, -1 by chunk size
    int overlap_chunks = -1;    // BLINK_OVERLAP_CHUNKS: chunks of the main group (-1: by size)
    int64_t overlap_bytes = 0;  // BLINK_OVERLAP_MIB: target chunk bytes of an overlapped group
    int direct_mode = -1;       // BLINK_DIRECT: 0 never, 1 every eligible group, -1 dense ones
    // device copy of the blob, when build_batch placed it (arena)
    void* blob_dev = nullptr;
    // host copy of a group's first set descriptor (a kernel argument)
    const SetDesc& host_desc(const BatchGroup& g) const {
        return ((const SetDesc*)blob.data())[g.first_desc];
    }
};

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

}  // namespace

// Where a batched call's scratch lives: the context's grow-only named buffers,
// or, for blink_image_batch_noalloc, consecutive 256-byte aligned pieces of
// one fixed device region.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static int batch_buffer(blink_ctx* ctx, Arena* arena, const char* name, size_t bytes,
                        void** out) {
    bytes = std::max<size_t>(bytes, 16);
    if (!arena) return ctx_scratch(ctx, name, bytes, out);
    const size_t off = align256(arena->used);
    if (off + bytes > arena->cap)
        return ctx->fail(BLINK_ERR_INVALID_ARGUMENT,
                         "the call needs more than the %zu MiB of the fixed scratch region "
                         "(%s: %zu bytes at offset %zu); split it into smaller calls",
                         arena->cap >> 20, name, bytes, off);
    *out = arena->base + off;
    arena->used = off + bytes;
    return BLINK_OK;
}

// Groups, scratch and descriptor tables of one call.  Every scratch buffer is
// reserved here, before any capture.  With an arena, the buffers and the
// descriptor blob (bp.blob_dev) are cut from it and the row-pass chunks are
// sized to what remains.
// Row-pass chunks of a register side (the main group of a call) that the
// Infinity Cache (256 MiB) can hold between the two passes: the row pass
// stores them with plain stores, the second pass reads them back from the
// cache, and each chunk's second pass runs beside the next chunk's row pass.
// kOverlapChunkBytes is the target chunk size (40 images at 1024^2, 3 at
// 4096^2).  Below 2048^2 only groups whose images share their plans (two or
// more timesteps per channel on average) are overlapped: where every image
// has a plan of its own, the row pass gathers from cold plan entries and the
// overlapped passes measured slower than one pass after the other (768
// channels x 1 timestep at 1024^2: 336k against 356k images/s; 64 x 10 at
// 1024^2: 416k against 376k).  From 2048^2 the rows are long enough to hide
// that (768 channels x 1 timestep at 4096^2: 137.6 against 159.0 ms per 3072
// images; single-image sets of six sides, main group 2048^2: 262k against
// 250k images/s).
static constexpr int64_t kOverlapChunkBytes = 168ll << 20;
// Direct row pass: a group whose half plane fits one workgroup's LDS (sides
// 128 and generic sides up to 186), whose sets have one timestep per channel
// and more than kDirectMinPerRow baselines per plan row, grids without a plan
// (k_grid_rows_direct, k_grid_rows_direct_fft): one workgroup per image walks
// all its baselines in order, computing their cells, instead of gathering
// plan entries (768 channels x 128^2, 32896 baselines: 0.32 against 0.60 ms
// per step).
static constexpr int64_t kDirectMinPerRow = 64;
static constexpr int64_t kCachedChunkBytes = 256ll << 20;

static int build_batch(blink_ctx* ctx, const blink_image_set* sets, int n_sets, BatchPlan& bp,
                       Arena* arena = nullptr) {
    int rc;
    // Row-pass buffer: up to 4 GiB of row-transformed grids per chunk
    // (BLINK_INTER_MIB overrides the budget, for tuning).  Chunks run one
    // after the other unless they are overlapped (below): small chunks alone
    // lose to one large launch (row-kernel ramp and tail per launch).
    int64_t inter_budget = 4096ll << 20;
    if (const char* e = getenv("BLINK_INTER_MIB")) {
        const long long mib = atoll(e);
        if (mib > 0) inter_budget = mib << 20;
    }
    // Overlapped chunks (see kOverlapChunkBytes); the switches are for tuning:
    // BLINK_OVERLAP_MIB (target chunk MiB, 0: no overlap), BLINK_OVERLAP_CHUNKS
    // (a fixed chunk count for the main FFT group, any side; 0 or 1: none),
    // BLINK_INTER_CACHED (0 / 1: streaming / plain row-pass stores always).
    bp.overlap_bytes = kOverlapChunkBytes;
    if (const char* e = getenv("BLINK_OVERLAP_MIB")) bp.overlap_bytes = (int64_t)atoll(e) << 20;
    if (const char* e = getenv("BLINK_OVERLAP_CHUNKS")) bp.overlap_chunks = std::max(0, atoi(e));
    if (const char* e = getenv("BLINK_INTER_CACHED")) bp.cached_mode = e[0] == '1' ? 1 : 0;
    if (arena) bp.overlap_chunks = 0, bp.overlap_bytes = 0;  // one row-pass buffer
    // Direct row pass (see kDirectMinPerRow); BLINK_DIRECT=0 / 1 turns it off /
    // on for every group of a side it supports, dense or not.
    if (const char* e = getenv("BLINK_DIRECT")) bp.direct_mode = e[0] == '1' ? 1 : 0;
    // groups (grid side, transform, weighting, gridding kernel) in order of
    // first appearance
    std::map<std::tuple<int, int, int, int, int>, int> gidx;
    for (int i = 0; i < n_sets; i++) {
        const int kb = sets[i].kernel_size;
        const std::tuple<int, int, int, int, int> key(sets[i].n_pixels, sets[i].transform,
                                                      sets[i].weighting, kb,
                                                      kb ? sets[i].kernel_mode : 0);
        auto it = gidx.find(key);
        if (it == gidx.end()) {
            it = gidx.emplace(key, (int)bp.groups.size()).first;
            BatchGroup g;
            g.n = sets[i].n_pixels;
            g.transform = sets[i].transform;
            g.uniform = sets[i].weighting == BLINK_WEIGHTING_UNIFORM;
            g.kb_size = kb;
            g.kb_mode = kb ? sets[i].kernel_mode : 0;
            // convolutional gridding runs its row pass on the generic kernels,
            // any side; at the register sides up to 2048 it writes their
            // column-blocked buffer, and the register second pass reads it
            g.reg = g.transform == BLINK_TRANSFORM_FFT && batch_reg_supported(g.n) && kb == 0;
            g.kb_blocked = g.transform == BLINK_TRANSFORM_FFT && kb > 0 &&
                           batch_reg_supported(g.n) && g.n <= 2048;
            bp.groups.push_back(g);
        }
        BatchGroup& g = bp.groups[it->second];
        g.sets.push_back(i);
        g.n_img += (int64_t)sets[i].n_ch * sets[i].n_t;
    }
    // the main group (most image bytes) keeps its own row-pass buffer (two
    // with overlap); the others share a second one on the aux stream
    {
        int64_t best = -1;
        for (size_t k = 0; k < bp.groups.size(); k++) {
            const int64_t bytes = bp.groups[k].n_img * (int64_t)bp.groups[k].n * bp.groups[k].n;
            if (bytes > best) best = bytes, bp.main = (int)k;
        }
    }
    // Groups on the direct row pass (no plan): sides whose half plane fits a
    // workgroup (direct_row_pass_fits), the FFT transform, natural weighting
    // and nearest-cell gridding, sets with one timestep per channel (every
    // image has a plan of its own) and dense grids (more than
    // kDirectMinPerRow baselines per plan row); with one workgroup per image,
    // a group needs images for at least half the CUs, unless it runs beside
    // the call's main group (one 180^2 image: 0.129 against 0.047 ms on the
    // plan path).
    std::vector<char> set_direct(n_sets, 0);
    const int64_t cus = device_cu_count();
    for (BatchGroup& g : bp.groups) {
        const bool beside_main = bp.groups.size() > 1 && &g - bp.groups.data() != bp.main;
        if (bp.direct_mode < 0 && 2 * g.n_img < cus && !beside_main) continue;
        // (the DFT transforms' half plane too: the same walk, no row FFT)
        if (g.uniform || bp.direct_mode == 0) continue;
        // convolutional gridding: the windowed direct row pass (256^2) only
        if (g.kb_size && !(g.kb_blocked && direct_uses_window(g.n))) continue;
        if (is_dft(g.transform) ? !direct_half_plane_fits(g.n) : !direct_row_pass_fits(g.n))
            continue;
        bool ok = true;
        int work = 0;
        for (int i : g.sets) {
            const int64_t n_bl = (int64_t)sets[i].n_ant * (sets[i].n_ant + 1) / 2;
            if (bp.direct_mode < 0 &&
                (sets[i].n_t != 1 || n_bl <= kDirectMinPerRow * (int64_t)(g.n / 2 + 1)))
                ok = false;
            work = std::max<int>({work, (int)n_bl, sets[i].n_ch * sets[i].n_t});
        }
        if (!ok) continue;
        g.direct = true;
        g.direct_work = work;
        for (int i : g.sets) set_direct[i] = 1;
    }
    // scratch sizes of the plans
    size_t rs_bytes = 0, en_bytes = 0;
    std::vector<size_t> rs_off(n_sets), en_off(n_sets), is_off(n_sets), co_off(n_sets);
    int64_t n_img_all = 0, n_units = 0, min_bl = INT64_MAX, max_bl = 0, n_cells = 0, n_bl_all = 0;
    int n_max = 0;
    for (int i = 0; i < n_sets; i++) {
        const blink_image_set& st = sets[i];
        const int64_t n = st.n_pixels;
        const int n_bl = st.n_ant * (st.n_ant + 1) / 2;
        rs_off[i] = rs_bytes;
        rs_bytes += align16(sizeof(int) * (size_t)st.n_ch * (n + 1)) +
                    align16(sizeof(int) * (size_t)st.n_ch * row_list_len((int)n));
        is_off[i] = (size_t)n_img_all;
        n_img_all += (int64_t)st.n_ch * st.n_t;
        co_off[i] = (size_t)n_bl_all;
        n_bl_all += n_bl;
        en_off[i] = en_bytes;
        if (set_direct[i]) continue;  // no plan
        en_bytes += align16(sizeof(uint32_t) * 2 * (size_t)st.n_ch * std::max(n_bl, 1));
        n_units += st.n_ch;
        n_cells += (int64_t)st.n_ch * n_bl;
        min_bl = std::min<int64_t>(min_bl, n_bl);
        max_bl = std::max<int64_t>(max_bl, n_bl);
        n_max = std::max(n_max, st.n_pixels);
    }
    // plan parts: about 1024 workgroups over the units with at least 256
    // baselines per part, and never more than kPlanPartMax baselines per part
    // (fewer parts shorten the scan, which walks every part of every row:
    // 64 channels x 8256 baselines plan in 30 us with 16 parts, 41 us with 32)
    int64_t parts = std::max<int64_t>(1, (1024 + n_units - 1) / std::max<int64_t>(n_units, 1));
    if (n_units == 0) min_bl = max_bl = 0, n_max = 1;
    parts = std::min<int64_t>({parts, 64, std::max<int64_t>(1, min_bl / 256)});
    parts = std::max<int64_t>(parts, (max_bl + kPlanPartMax - 1) / kPlanPartMax);
    bp.parts = (int)parts;
    bp.n_units = (int)n_units;
    bp.n_max = n_max;
    const size_t partial_bytes = sizeof(int) * (size_t)n_units * parts * n_max;

    // A set of a register-path side with at most 64 baselines per plan row
    // (a sparse grid) is gathered from compacted visibilities
    // (one float2 per image and baseline) instead of the records: its row
    // tiles hold few entries each, so the first-touch misses of the 32-byte
    // records are exposed in every tile (640 x 1024^2, 8256 baselines: row
    // pass 480 -> 350 us plus 36 us of compaction).  Dense grids amortise
    // them over thousands of entries per tile and lose (768 x 256^2, 32896
    // baselines: 412 -> 293 us plus 185 us).  BLINK_COMPACT=0 turns it off.
    std::vector<size_t> cp_off(n_sets, SIZE_MAX);
    size_t cp_bytes = 0;
    {
        const char* env = getenv("BLINK_COMPACT");
        const bool on = !(env && env[0] == '0');
        const bool dense = env && env[0] == '2';  // every set of a register side
        bp.compact_on = on;
        bp.compact_dense = dense;
        for (BatchGroup& g : bp.groups) {
            if (!on || !g.reg || g.n > 2048 || g.direct) continue;
            for (int i : g.sets) {
                const int64_t n_bl = (int64_t)sets[i].n_ant * (sets[i].n_ant + 1) / 2;
                if (!dense && n_bl > 64 * (int64_t)(g.n / 2 + 1)) continue;
                cp_off[i] = cp_bytes;
                cp_bytes += align16(sizeof(float2) * (size_t)sets[i].n_ch * sets[i].n_t *
                                    (size_t)n_bl);
                g.compact_max_bl = std::max(g.compact_max_bl, (int)n_bl);
            }
        }
    }
    // descriptor blob layout: SetDesc table (group order), units, image -> set tables
    size_t blob_bytes = align16(sizeof(SetDesc) * (size_t)n_sets);
    bp.units_off = blob_bytes;
    blob_bytes = align16(blob_bytes + sizeof(PlanUnit) * (size_t)n_units);
    for (BatchGroup& g : bp.groups) {
        if (g.sets.size() < 2 || g.transform != BLINK_TRANSFORM_FFT) continue;
        g.img_set_off = blob_bytes;
        blob_bytes = align16(blob_bytes + sizeof(int) * (size_t)g.n_img);
    }

    // averages of FFT sets: the weighted sum of one image's row-pass buffer and
    // the average pass's metadata
    std::vector<size_t> av_off(n_sets, SIZE_MAX);
    size_t av_bytes = 0;
    for (int i = 0; i < n_sets; i++) {
        // odd sides average their formed images
        if (!sets[i].average || is_dft(sets[i].transform) || sets[i].n_pixels % 2) continue;
        const size_t n = (size_t)sets[i].n_pixels;
        av_off[i] = av_bytes;
        av_bytes += align256(sizeof(float2) * (n / 2 + 1) * n) + align256(4 * (n + 2));
    }
    void *rs, *en, *is, *pa, *cl, *co, *cp = nullptr, *av = nullptr;
    if (av_bytes && (rc = batch_buffer(ctx, arena, "batch_average", av_bytes, &av))) return rc;
    if (cp_bytes && (rc = batch_buffer(ctx, arena, "batch_compact", cp_bytes, &cp))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_rows", rs_bytes, &rs))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_entries", en_bytes, &en))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_isums", 4 * (size_t)n_img_all, &is))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_plan_parts", partial_bytes, &pa))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_plan_cells", 4 * (size_t)n_cells, &cl))) return rc;
    if ((rc = batch_buffer(ctx, arena, "batch_bl_coef", sizeof(double) * (size_t)n_bl_all, &co)))
        return rc;
    if (arena) {
        if ((rc = batch_buffer(ctx, arena, "batch_desc", blob_bytes, &bp.blob_dev))) return rc;
        // the row-pass buffer takes what is left, split between the call's
        // stream and the second stream when there are several groups
        const int64_t left = (int64_t)arena->cap - (int64_t)align256(arena->used) - 256;
        inter_budget = std::max<int64_t>(0, bp.groups.size() > 1 ? left / 2 - 256 : left);
    }
    bp.inter_budget = inter_budget;

    // plans of a group: one per (set, channel)
    auto group_channels = [&](const BatchGroup& g) {
        int64_t c = 0;
        for (int i : g.sets) c += sets[i].n_ch;
        return c;
    };
    // row-pass chunks: images per row / column launch pair
    std::vector<size_t> group_inter(bp.groups.size(), 16);
    for (BatchGroup& g : bp.groups) {
        const int64_t n = g.n;
        const size_t gk = &g - bp.groups.data();
        if (is_dft(g.transform)) {
            for (int i : g.sets) {
                const int n_img = sets[i].n_ch * sets[i].n_t;
                const int64_t per = dft_inter_bytes(n);
                const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(n_img, inter_budget / per));
                group_inter[gk] = std::max(group_inter[gk], (size_t)(per * chunk));
            }
            continue;
        }
        const int64_t per = g.reg ? batch_inter_bytes(g.n) : 8 * (n / 2 + 1) * n;
        // chunks of the main group: a fixed count, or (register sides up to
        // 2048) the target size when the images need more than one chunk of it
        int chunks = 0;
        if ((int)gk == bp.main && g.n_img >= 2) {
            if (bp.overlap_chunks >= 0)
                chunks = bp.overlap_chunks;
            else if (g.reg && bp.overlap_bytes > 0 && per * g.n_img > bp.overlap_bytes &&
                     (g.n >= 2048 || g.n_img >= 2 * group_channels(g)))
                chunks = (int)((per * g.n_img + bp.overlap_bytes - 1) / bp.overlap_bytes);
        }
        const bool overlap = chunks >= 2;
        // with overlap the budget holds both buffers
        const int64_t budget = overlap ? inter_budget / 2 : inter_budget;
        g.chunk = (int)std::max<int64_t>(1, std::min<int64_t>(g.n_img, budget / per));
        if (overlap) {
            g.chunk = (int)std::min<int64_t>(g.chunk, (g.n_img + chunks - 1) / chunks);
            g.overlap = g.chunk < g.n_img;
        }
        g.cached_inter = bp.cached_mode >= 0 ? bp.cached_mode == 1
                                             : per * g.chunk <= kCachedChunkBytes;
        group_inter[gk] = (size_t)(per * g.chunk);
    }
    size_t inter_bytes = 16, main_bytes = align256(group_inter[bp.main]);
    {
        size_t aux = 0;
        for (size_t k = 0; k < bp.groups.size(); k++)
            if ((int)k != bp.main) aux = std::max(aux, group_inter[k]);
        inter_bytes = main_bytes * (bp.groups[bp.main].overlap ? 2 : 1) + aux;
    }
    void* inter;
    if ((rc = batch_buffer(ctx, arena, "batch_inter", inter_bytes, &inter))) return rc;
    bp.cells = (uint32_t*)cl;
    bp.inter = (float*)inter;
    bp.inter2 = bp.groups[bp.main].overlap ? (float*)((char*)inter + main_bytes) : nullptr;
    bp.inter_aux = (float*)((char*)inter + main_bytes * (bp.inter2 ? 2 : 1));
    bp.partial = (int*)pa;
    bp.sc.resize(n_sets);
    for (int i = 0; i < n_sets; i++) {
        bp.sc[i].row_start = (int*)((char*)rs + rs_off[i]);
        bp.sc[i].row_list =
            (int*)((char*)rs + rs_off[i] +
                   align16(sizeof(int) * (size_t)sets[i].n_ch * (sets[i].n_pixels + 1)));
        bp.sc[i].entries = (uint32_t*)((char*)en + en_off[i]);
        bp.sc[i].isums = (uint32_t*)is + is_off[i];
        bp.sc[i].coef = (double*)co + co_off[i];
        if (cp_off[i] != SIZE_MAX) bp.sc[i].compact = (float2*)((char*)cp + cp_off[i]);
        if (av_off[i] != SIZE_MAX) {
            const size_t n = (size_t)sets[i].n_pixels;
            bp.sc[i].avg_acc = (float*)((char*)av + av_off[i]);
            bp.sc[i].avg_meta = (unsigned*)((char*)av + av_off[i] +
                                            align256(sizeof(float2) * (n / 2 + 1) * n));
        }
        const int64_t per = dft_inter_bytes(sets[i].n_pixels);
        const int n_img = sets[i].n_ch * sets[i].n_t;
        bp.sc[i].chunk = (int)std::max<int64_t>(1, std::min<int64_t>(n_img, inter_budget / per));
    }
    for (int i = 0; i < n_sets; i++) {
        const int n = sets[i].n_pixels;
        if (arena) {
            // tables are never built here: the sides blink_create prepared
            if (!prepared_side(n) || !ctx->twiddles.count(n))
                return ctx->fail(BLINK_ERR_UNSUPPORTED,
                                 "set %d: grid side %d has no FFT table in this context (the "
                                 "no-allocation call images the sides blink_create prepares: "
                                 "128, 180, 256, 512, 1024, 2048, 4096)", i, n);
            continue;
        }
        const float* tw;
        if ((rc = twiddles(ctx, n, &tw))) return rc;
        if (sets[i].kernel_size > 0) {
            const double* table;
            if ((rc = kb_table(ctx, sets[i].kernel_size, sets[i].kernel_mode, &table))) return rc;
            bp.sc[i].kb_table = table;
        }
        if (is_dft(sets[i].transform)) {
            const uint16_t* pl;
            for (int stage = 1; stage <= 2; stage++)
                if ((rc = dft_planes(ctx, n, stage, dft_prec(sets[i].transform), &pl)))
                    return rc;
        }
    }

    // descriptor blob: SetDesc table (group order), units, image -> set tables
    std::vector<SetDesc> desc;
    std::vector<int> desc_of(n_sets);
    for (BatchGroup& g : bp.groups) {
        g.first_desc = (int)desc.size();
        int64_t base = 0;
        for (int i : g.sets) {
            const blink_image_set& st = sets[i];
            SetDesc d{};
            d.g.vis = (const float4*)st.vis;
            d.g.phase_coef = bp.sc[i].coef;
            d.w_m = st.w_m;
            d.cable_m = st.cable_m;
            d.coef_out = bp.sc[i].coef;
            d.g.freq_hz = st.freq_hz;
            d.g.corrections = st.corrections & 3u;
            d.g.pol = st.pol;
            d.g.n_bl = st.n_ant * (st.n_ant + 1) / 2;
            d.g.n_ch = st.n_ch;
            d.u_m = st.u_m;
            d.v_m = st.v_m;
            d.flags = st.bl_flags;
            d.du = st.delta_u;
            d.dv = st.delta_v;
            d.min_uv = st.min_uv;
            d.n_pixels = st.n_pixels;
            d.n_img = st.n_ch * st.n_t;
            d.img_base = base;
            d.row_start = bp.sc[i].row_start;
            d.row_list = bp.sc[i].row_list;
            d.entries = bp.sc[i].entries;
            d.isums = bp.sc[i].isums;
            d.images = (float2*)st.images;
            d.count_sums = st.count_sums;
            // the FFT kernels transform along v first, so that their second
            // pass writes whole image rows (the DFT stages keep rows v)
            d.uv_swap = g.transform == BLINK_TRANSFORM_FFT;
            d.compact = bp.sc[i].compact;
            d.uniform = st.weighting == BLINK_WEIGHTING_UNIFORM ? 1 : 0;
            d.kb_size = st.kernel_size;
            d.kb_table = bp.sc[i].kb_table;
            // (the FFT groups' direct row passes write the row starts instead)
            d.all_rows = (st.kernel_size > 0 && !g.direct) || (g.direct && is_dft(g.transform))
                             ? 1
                             : 0;
            d.cached_inter = g.cached_inter ? 1 : 0;
            // the windowed direct row pass's |u|, |v| maxima: in the row
            // lists, which no direct set uses (n_ch (N + 3) >= 2 kUvMaxParts)
            d.uv_max = g.direct && !is_dft(g.transform) && direct_uses_window(g.n)
                           ? (float*)bp.sc[i].row_list
                           : nullptr;
            base += d.n_img;
            desc_of[i] = (int)desc.size();
            desc.push_back(d);
        }
    }
    bp.blob.assign(blob_bytes, 0);
    memcpy(bp.blob.data(), desc.data(), sizeof(SetDesc) * desc.size());
    PlanUnit* units = (PlanUnit*)(bp.blob.data() + bp.units_off);
    int64_t u = 0, cell_off = 0;
    for (int i = 0; i < n_sets; i++) {
        if (set_direct[i]) continue;
        const int n_bl = sets[i].n_ant * (sets[i].n_ant + 1) / 2;
        for (int ch = 0; ch < sets[i].n_ch; ch++) {
            units[u++] = PlanUnit{desc_of[i], ch, cell_off};
            cell_off += n_bl;
        }
    }
    for (BatchGroup& g : bp.groups) {
        if (g.sets.size() < 2 || g.transform != BLINK_TRANSFORM_FFT) continue;
        int* t = (int*)(bp.blob.data() + g.img_set_off);
        int64_t k = 0;
        for (size_t j = 0; j < g.sets.size(); j++) {
            const int n_img = sets[g.sets[j]].n_ch * sets[g.sets[j]].n_t;
            for (int q = 0; q < n_img; q++) t[k++] = (int)j;
        }
    }
    return BLINK_OK;
}

// Device addresses of the blob's tables.
struct BatchTables {
    const SetDesc* sets;
    const PlanUnit* units;
    const unsigned char* base;
    const SetDesc* group_sets(const BatchGroup& g) const { return sets + g.first_desc; }
    const int* img_set(const BatchGroup& g) const {
        return g.sets.size() < 2 ? nullptr : (const int*)(base + g.img_set_off);
    }
};

static BatchTables tables_at(const BatchPlan& bp, const void* dblob) {
    BatchTables t;
    t.base = (const unsigned char*)dblob;
    t.sets = (const SetDesc*)dblob;
    t.units = (const PlanUnit*)(t.base + bp.units_off);
    return t;
}

// Default tiles (plan rows per row tile, output lines per second-pass block): the
// pairs measured fastest on MI355X for the 128-tile and 256-antenna arrays;
// the autotuner keeps them unless another pair is more than 3 % faster.
static std::pair<int, int> default_tiles(int n) {
    switch (n) {
        case 128: return {32, 64};
        case 256: return {8, 32};
        case 512: return {8, 16};
        case 1024: return {8, 16};
        case 2048: return {8, 16};
        case 4096: return {2, 8};
        default: return {0, 0};
    }
}

// Batch class of a group: its images per launch in factors of 4 (1, 2-3,
// 4-15, 16-63, ... images), so that a side re-tunes when the launches it
// makes change size materially.
static int batch_class(const BatchGroup& g) {
    const int64_t cnt = std::max<int64_t>(1, std::min<int64_t>(g.chunk, g.n_img));
    int c = 0;  // floor(log4(cnt))
    for (int64_t v = cnt; v >= 4; v >>= 2) c++;
    // direct groups tune their column block apart (row tile 0: none)
    return (cnt == 1 ? 0 : 1 + c) + (g.direct ? 32 : 0);
}

// Tiles for a register-side group: pinned, tuned for its batch class, or the
// defaults.
static bool has_tiles(const blink_ctx* ctx, const BatchGroup& g) {
    return ctx->pinned_tiles.count(g.n) || ctx->tuned_tiles.count({g.n, batch_class(g)});
}
static std::pair<int, int> tiles_for(const blink_ctx* ctx, const BatchGroup& g) {
    auto p = ctx->pinned_tiles.find(g.n);
    auto t = ctx->tuned_tiles.find({g.n, batch_class(g)});
    if (g.direct) {
        // a pinned pair pins the column block; the direct row pass has no tile
        std::pair<int, int> r{0, default_tiles(g.n).second};
        if (t != ctx->tuned_tiles.end()) r = t->second;
        if (p != ctx->pinned_tiles.end()) r.second = p->second.second;
        return r;
    }
    if (p != ctx->pinned_tiles.end()) return p->second;
    if (t != ctx->tuned_tiles.end()) return t->second;
    return default_tiles(g.n);
}

// Average GPU milliseconds of `reps` launches made by fn() on stream s.
}  // extern "C"
template <class F>
static int time_launches(blink_ctx* ctx, hipStream_t s, int reps, F fn, float* ms) {
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    hipError_t e = fn();  // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = fn();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return ctx->fail(BLINK_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    *ms = t / reps;
    return BLINK_OK;
}
extern "C" {

// Chooses the row-tile and column-block sizes of the register kernels for
// one grid side by timing every built variant on the group's first chunk.
// The plans must have been built (zero + plan launches) on s beforehand.
static int autotune_group(blink_ctx* ctx, const BatchPlan& bp, const BatchGroup& g,
                          const BatchTables& t, hipStream_t s) {
    const int n = g.n;
    int rows[8], cols[8];
    rows[0] = 0;  // direct groups: the one row kernel
    const int nr = g.direct ? 1 : batch_row_variants(n, rows, 8);
    const int nc = batch_col_variants(n, cols, 8);
    // timed on up to 2 GiB of row-pass buffer (all of a 640-image 1024^2
    // chunk, 30 images at 4096^2), which keeps the tuning under a second
    const int64_t cap = std::max<int64_t>(16, (2ll << 30) / batch_inter_bytes(n));
    const int cnt = (int)std::min<int64_t>({(int64_t)g.chunk, g.n_img, cap});
    const float* tw = ctx->twiddles[n];
    const SetDesc* sd = t.group_sets(g);
    const SetDesc& first = bp.host_desc(g);
    const int* is = t.img_set(g);
    const int ns = (int)g.sets.size();
    // direct groups: the one row pass runs once, and the column blocks are
    // timed alone (the row pass's own variation would hide their difference)
    if (g.direct) {
        HIP_TRY(launch_direct_prep(sd, ns, g.direct_work, s));
        HIP_TRY(launch_grid_rows_direct(first, sd, is, ns, n, 0, cnt, bp.inter, tw, s));
    }
    // Every (row tile, column block) pair is timed as the pipeline runs it
    // (row pass then column pass), in interleaved rounds; each pair keeps its
    // fastest round, so that clock ramp-up and neighbour noise do not favour
    // the pairs timed last.
    constexpr int kRounds = 3;
    float tp[8][8];
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) tp[r][c] = 1e30f;
    for (int round = 0; round < kRounds; round++)
        for (int r = 0; r < nr; r++)
            for (int c = 0; c < nc; c++) {
                float ms;
                int rc = time_launches(ctx, s, 2, [&] {
                    hipError_t e = g.direct ? hipSuccess
                                            : launch_grid_rows_reg(first, sd, is, ns, n, rows[r],
                                                                   0, cnt, bp.inter, tw, s);
                    if (e != hipSuccess) return e;
                    return launch_cols_reg(first, sd, is, ns, n, cols[c], 0, cnt, bp.inter, tw,
                                           s);
                }, &ms);
                if (rc) return rc;
                tp[r][c] = std::min(tp[r][c], ms);
            }
    int best_r = rows[0], best_c = cols[0];
    float best = 1e30f, t_default = 1e30f;
    const std::pair<int, int> dflt =
        g.direct ? std::pair<int, int>{0, default_tiles(n).second} : default_tiles(n);
    for (int r = 0; r < nr; r++)
        for (int c = 0; c < nc; c++) {
            if (tp[r][c] < best) best = tp[r][c], best_r = rows[r], best_c = cols[c];
            if (rows[r] == dflt.first && cols[c] == dflt.second) t_default = tp[r][c];
        }
    if (t_default <= 1.03f * best) best_r = dflt.first, best_c = dflt.second;
    ctx->tuned_tiles[{n, batch_class(g)}] = {best_r, best_c};
    if (!g.direct) ctx->last_tuned[n] = {best_r, best_c};  // reported: plan-path tiles
    return BLINK_OK;
}

// Zeroes the count sums and builds every plan of the call on s (none when
// every group takes the direct row pass).
static int enqueue_plans(blink_ctx* ctx, const BatchPlan& bp, const BatchTables& t,
                         hipStream_t s) {
    if (bp.n_units == 0) return BLINK_OK;
    LAUNCH("plan", s,
           launch_plan_units(t.sets, t.units, bp.n_units, bp.parts, bp.n_max, bp.partial,
                             bp.cells, s));
    return BLINK_OK;
}

// Work forked from s onto the aux stream: fork() makes the aux stream wait
// for s; join() (also on every error path after a successful fork) makes s
// wait for everything enqueued on the aux stream since, so that a capture on
// s never ends with the aux stream still part of it.
struct AuxFork {
    blink_ctx* ctx;
    hipStream_t s, a;
    hipEvent_t fork_ev, join_ev;
    bool forked = false;
    int fork() {
        hipError_t e = hipEventRecord(fork_ev, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(a, fork_ev, 0);
        if (e != hipSuccess)
            return ctx->fail(BLINK_ERR_HIP, "fork to the second stream failed: %s",
                             hipGetErrorString(e));
        forked = true;
        return BLINK_OK;
    }
    int join() {
        if (!forked) return BLINK_OK;
        forked = false;
        hipError_t e = hipEventRecord(join_ev, a);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, join_ev, 0);
        if (e != hipSuccess)
            return ctx->fail(BLINK_ERR_HIP, "join of the second stream failed: %s",
                             hipGetErrorString(e));
        return BLINK_OK;
    }
    // keeps the first error: a failure inside the forked work, else the join's
    int join_after(int rc) {
        const int rj = join();
        return rc ? rc : rj;
    }
};

static int enqueue_chunks(blink_ctx* ctx, const blink_image_set* sets, const BatchPlan& bp,
                          const BatchTables& t, const BatchGroup& g, float* inter,
                          float* inter2, hipStream_t s, hipStream_t p);

// Enqueues one group's launches on s, with row-pass buffer `inter` (and, for
// an overlapped group, `inter2` and the context's pass stream).
static int enqueue_group(blink_ctx* ctx, const blink_image_set* sets, const BatchPlan& bp,
                         const BatchTables& t, const BatchGroup& g, float* inter,
                         hipStream_t s, float* inter2 = nullptr) {
    const int n = g.n;
    if (is_dft(g.transform)) {
        // per set and chunk: half plane -> stage 1 (complex) -> stage 2 (real, scaled)
        const int prec = dft_prec(g.transform);
        if (g.direct)
            LAUNCH("direct_prep", s,
                   launch_direct_prep(t.group_sets(g), (int)g.sets.size(), g.direct_work, s));
        const uint16_t* pl1 = ctx->dft_planes[n * 8 + 2 + prec];
        const uint16_t* pl2 = ctx->dft_planes[n * 8 + 4 + prec];
        for (size_t j = 0; j < g.sets.size(); j++) {
            const int i = g.sets[j];
            const blink_image_set& st = sets[i];
            const SetScratch& sc = bp.sc[i];
            const int n_img = st.n_ch * st.n_t;
            GatherArgs ga{};
            ga.vis = (const float4*)st.vis;
            ga.phase_coef = sc.coef;
            ga.freq_hz = st.freq_hz;
            ga.corrections = st.corrections & 3u;
            ga.pol = st.pol;
            ga.n_bl = st.n_ant * (st.n_ant + 1) / 2;
            ga.n_ch = st.n_ch;
            float* half = inter;
            float* tbuf = inter + 2 * (int64_t)(n / 2 + 1) * n * sc.chunk;
            const SetDesc& d = ((const SetDesc*)bp.blob.data())[g.first_desc + j];
            for (int img0 = 0; img0 < n_img; img0 += sc.chunk) {
                const int cnt = std::min(sc.chunk, n_img - img0);
                if (g.direct)
                    LAUNCH("grid_half_rows", s, launch_grid_half_direct(d, cnt, img0, half, s));
                else
                    LAUNCH("grid_half_rows", s,
                           launch_grid_half_rows(ga, sc.row_start, sc.entries, n, cnt, img0,
                                                 half, sc.isums, s, g.uniform));
                LAUNCH("dft_rows", s, launch_dft_rows(half, pl1, tbuf, n, cnt, prec, s));
                LAUNCH("dft_cols", s,
                       launch_dft_cols(tbuf, pl2, st.images, n, cnt, img0, sc.isums,
                                       st.count_sums, prec, s));
            }
            // the mean of the formed images (image_averaging_gpu's arithmetic)
            if (st.average)
                LAUNCH("average", s,
                       launch_average(st.images, (int64_t)n * n, n_img, 0, st.average, s));
        }
        return BLINK_OK;
    }
    if (!g.overlap || !inter2) return enqueue_chunks(ctx, sets, bp, t, g, inter, inter, s, s);
    // overlapped: fork the pass stream from s, join it back at the end (also
    // after a failure, so that a capture on s never ends with it attached)
    AuxFork pf{ctx, s, ctx->pass_stream, ctx->pass_fork, ctx->pass_join};
    int rc = pf.fork();
    if (!rc) rc = enqueue_chunks(ctx, sets, bp, t, g, inter, inter2, s, ctx->pass_stream);
    return pf.join_after(rc);
}

// The FFT group's chunks: row passes on s, second passes and averages on p.
// With p != s, chunk k uses buffer k % 2 (inter, inter2): its second pass
// waits for its row pass, and the row pass of chunk k + 2 waits for it.
static int enqueue_chunks(blink_ctx* ctx, const blink_image_set* sets, const BatchPlan& bp,
                          const BatchTables& t, const BatchGroup& g, float* inter,
                          float* inter2, hipStream_t s, hipStream_t p) {
    const int n = g.n;
    const float* tw = ctx->twiddles[n];
    const bool two = p != s;
    const SetDesc* sd = t.group_sets(g);
    const SetDesc& first = bp.host_desc(g);
    const int* is = t.img_set(g);
    const int ns = (int)g.sets.size();
    FftPlan plan;
    make_fft_plan(n, plan);
    const std::pair<int, int> tl = tiles_for(ctx, g);
    // the direct row pass's count sums and phase coefficients
    if (g.direct) LAUNCH("direct_prep", s, launch_direct_prep(sd, ns, g.direct_work, s));
    // the second pass is skipped when no set of the group forms its images
    bool any_images = false;
    for (int i : g.sets) any_images |= sets[i].images != nullptr;
    auto second_pass = [&](const SetDesc& f, const SetDesc* tab, const int* img_set, int n_sets,
                           int64_t gi0, int cnt, const float* src) {
        return g.reg || g.kb_blocked
                   ? launch_cols_reg(f, tab, img_set, n_sets, n, tl.second, gi0, cnt, src, tw, p)
                   : launch_cols_half(f, tab, img_set, n_sets, plan, gi0, cnt, src, tw, p);
    };
    int k = 0;
    for (int64_t gi0 = 0; gi0 < g.n_img; gi0 += g.chunk, k++) {
        const int cnt = (int)std::min<int64_t>(g.chunk, g.n_img - gi0);
        float* buf = two && (k & 1) ? inter2 : inter;
        hipEvent_t rows_done = ctx->pass_ev[2 * (k % kPassEvents)];
        hipEvent_t pass_done = ctx->pass_ev[2 * (k % kPassEvents) + 1];
        // the buffer's previous chunk (k - 2) must have been read
        if (two && k >= 2)
            HIP_TRY(hipStreamWaitEvent(s, ctx->pass_ev[2 * ((k - 2) % kPassEvents) + 1], 0));
        if (g.direct && g.kb_size)
            LAUNCH("grid_rows_kb", s,
                   launch_grid_rows_direct(first, sd, is, ns, n, gi0, cnt, buf, tw, s));
        else if (g.direct && g.reg)
            LAUNCH("grid_rows_fft", s,
                   launch_grid_rows_direct(first, sd, is, ns, n, gi0, cnt, buf, tw, s));
        else if (g.direct)
            LAUNCH("grid_rows_fft", s,
                   launch_grid_rows_direct_fft(first, sd, is, ns, plan, gi0, cnt, buf, tw, s));
        else if (g.reg)
            LAUNCH("grid_rows_fft", s,
                   launch_grid_rows_reg(first, sd, is, ns, n, tl.first, gi0, cnt, buf, tw, s,
                                        g.uniform));
        else if (g.kb_size)
            LAUNCH("grid_rows_kb", s,
                   launch_grid_rows_kb(first, sd, is, ns, plan, gi0, cnt, buf, tw, g.kb_blocked,
                                       s));
        else
            LAUNCH("grid_rows_fft", s,
                   launch_grid_rows_fft(first, sd, is, ns, plan, gi0, cnt, buf, tw, s));
        if (two) {
            HIP_TRY(hipEventRecord(rows_done, s));
            HIP_TRY(hipStreamWaitEvent(p, rows_done, 0));
        }
        if (any_images) LAUNCH("fft_cols", p, second_pass(first, sd, is, ns, gi0, cnt, buf));
        // sets with an average: add this chunk's images of the set
        for (size_t j = 0; j < g.sets.size(); j++) {
            const SetScratch& sc = bp.sc[g.sets[j]];
            if (!sc.avg_acc) continue;
            const SetDesc& d = ((const SetDesc*)bp.blob.data())[g.first_desc + j];
            const int64_t a = std::max<int64_t>(gi0, d.img_base);
            const int64_t b = std::min<int64_t>(gi0 + cnt, d.img_base + d.n_img);
            if (a >= b) continue;
            LAUNCH("average_rows", p,
                   launch_average_rows(d, a - d.img_base, (int)(b - a), a - gi0, buf,
                                       sc.avg_acc, a > d.img_base, sc.avg_meta, p));
        }
        if (two) HIP_TRY(hipEventRecord(pass_done, p));
    }
    // the averages: one second pass each over the summed rows, scaled by 1/n_img
    for (size_t j = 0; j < g.sets.size(); j++) {
        const SetScratch& sc = bp.sc[g.sets[j]];
        if (!sc.avg_acc) continue;
        SetDesc d = ((const SetDesc*)bp.blob.data())[g.first_desc + j];
        d.images = (float2*)sets[g.sets[j]].average;
        d.isums = sc.avg_meta;
        d.row_start = (int*)(sc.avg_meta + 1);
        d.count_sums = nullptr;
        d.g.n_ch = 1;
        d.img_base = 0;
        d.n_img = 1;
        d.uniform = 0;
        LAUNCH("fft_cols", p, second_pass(d, nullptr, nullptr, 1, 0, 1, sc.avg_acc));
    }
    // odd sides: the mean of the formed images (image_averaging_gpu's arithmetic)
    if (n % 2)
        for (int i : g.sets) {
            const blink_image_set& st = sets[i];
            if (st.average)
                LAUNCH("average", p,
                       launch_average(st.images, (int64_t)n * n, st.n_ch * st.n_t, 0, st.average,
                                      p));
        }
    return BLINK_OK;
}

static int enqueue_compaction(blink_ctx* ctx, const BatchPlan& bp, const BatchTables& t,
                              hipStream_t a) {
    for (const BatchGroup& g : bp.groups)
        if (g.compact_max_bl)
            LAUNCH("vis_compact", a,
                   g.sets.size() == 1
                       ? launch_vis_compact(bp.host_desc(g), g.n_img, a)
                       : launch_vis_compact_sets(t.group_sets(g), t.img_set(g), g.n_img,
                                                 g.compact_max_bl, a));
    return BLINK_OK;
}

// Enqueues the whole call on s.  All scratch was reserved by build_batch, so
// this can be captured.  With several grid sides, the main group (most image
// bytes) runs on s and the others beside it, one after another, on the aux
// stream with their own row-pass buffer: the small sides' latency-bound
// launches overlap the large side's bandwidth-bound ones.
static int enqueue_batch(blink_ctx* ctx, const blink_image_set* sets, const BatchPlan& bp,
                         const BatchTables& t, hipStream_t s) {
    hipStream_t a = ctx->aux_stream;
    // the visibility compaction reads only the visibilities and the sets' own
    // inputs, so it runs on the aux stream beside the plan; the row passes
    // (on s, or on the aux stream after it) come after both
    bool compact = false;
    for (const BatchGroup& g : bp.groups) compact |= g.compact_max_bl != 0;
    AuxFork cf{ctx, s, a, ctx->compact_fork, ctx->compact_join};
    int rc = BLINK_OK;
    if (compact && !(rc = cf.fork())) rc = enqueue_compaction(ctx, bp, t, a);
    if (!rc) rc = enqueue_plans(ctx, bp, t, s);
    if ((rc = cf.join_after(rc))) return rc;
    if (bp.groups.size() == 1)
        return enqueue_group(ctx, sets, bp, t, bp.groups[0], bp.inter, s, bp.inter2);
    AuxFork gf{ctx, s, a, ctx->fork_event, ctx->join_event};
    if ((rc = gf.fork())) return rc;
    for (size_t k = 0; k < bp.groups.size() && !rc; k++)
        if ((int)k != bp.main) rc = enqueue_group(ctx, sets, bp, t, bp.groups[k], bp.inter_aux, a);
    if (rc) return gf.join_after(rc);
    // the main group runs on s while the aux stream works; join at the end
    rc = enqueue_group(ctx, sets, bp, t, bp.groups[bp.main], bp.inter, s, bp.inter2);
    return gf.join_after(rc);
}

// Autotunes the register kernels of every group whose side has no tiles yet.
static int autotune_batch(blink_ctx* ctx, const BatchPlan& bp, const BatchTables& t,
                          hipStream_t s) {
    if (!ctx->autotune) return BLINK_OK;
    bool need = false;
    for (const BatchGroup& g : bp.groups) need |= g.reg && !has_tiles(ctx, g);
    if (!need) return BLINK_OK;
    int rc = enqueue_plans(ctx, bp, t, s);
    if (rc) return rc;
    for (const BatchGroup& g : bp.groups) {
        if (!g.reg || has_tiles(ctx, g)) continue;
        // the timed row passes gather what the call will gather
        if (g.compact_max_bl)
            HIP_TRY(g.sets.size() == 1
                        ? launch_vis_compact(bp.host_desc(g), g.n_img, s)
                        : launch_vis_compact_sets(t.group_sets(g), t.img_set(g), g.n_img,
                                                  g.compact_max_bl, s));
        if ((rc = autotune_group(ctx, bp, g, t, s))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s));
    return BLINK_OK;
}

// Cache key of a batched call: every field of every set (by value, so that
// struct padding never enters it) except the visibility, image, count-sum and
// average addresses (their presence is part of the key), plus the environment
// switches that shape the call.  Calls that differ only in those addresses
// share up to kBufferVariants instantiated graphs (the cache's full key adds
// the addresses).
static std::string graph_key(const blink_image_set* sets, int n_sets, const BatchPlan& bp) {
    std::string k;
    auto put = [&](const auto& v) { k.append((const char*)&v, sizeof v); };
    put(n_sets);
    put(bp.inter_budget);
    put(bp.compact_on);
    put(bp.compact_dense);
    put(bp.cached_mode);
    put(bp.overlap_chunks);
    put(bp.overlap_bytes);
    put(bp.direct_mode);
    for (int i = 0; i < n_sets; i++) {
        const blink_image_set& st = sets[i];
        put(st.n_ant), put(st.n_ch), put(st.n_t);
        put(st.u_m), put(st.v_m), put(st.w_m), put(st.bl_flags), put(st.freq_hz), put(st.cable_m);
        put(st.corrections), put(st.delta_u), put(st.delta_v), put(st.min_uv);
        put(st.n_pixels), put(st.pol), put(st.transform), put(st.weighting);
        put(st.kernel_size), put(st.kernel_mode);
        const bool has_sums = st.count_sums != nullptr, has_images = st.images != nullptr,
                   has_average = st.average != nullptr;
        put(has_sums), put(has_images), put(has_average);
    }
    return k;
}

static std::vector<const void*> io_pointers(const blink_image_set* sets, int n_sets) {
    std::vector<const void*> p;
    p.reserve(3 * (size_t)n_sets);
    for (int i = 0; i < n_sets; i++) {
        p.push_back(sets[i].vis);
        p.push_back(sets[i].images);
        p.push_back(sets[i].count_sums);
        p.push_back(sets[i].average);
    }
    return p;
}

// Captures the call's launches on the context's capture stream.
static int capture_batch(blink_ctx* ctx, const blink_image_set* sets, const BatchPlan& bp,
                         const BatchTables& t, hipGraph_t* out) {
    hipStream_t cs = ctx->capture_stream;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return ctx->fail(BLINK_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    const int rc = enqueue_batch(ctx, sets, bp, t, cs);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        return ctx->fail(BLINK_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    *out = graph;
    return BLINK_OK;
}

// Argument checks shared by the batched entry points.
static int check_sets(blink_ctx* ctx, const blink_image_set* sets, int n_sets) {
    CHECK_ARG(sets && n_sets >= 1, "need at least one image set");
    for (int i = 0; i < n_sets; i++) {
        const blink_image_set& st = sets[i];
        int rc = check_vis_shape(ctx, st.n_ant, st.n_ch, st.n_t);
        if (rc) return rc;
        CHECK_ARG(st.vis && st.u_m && st.v_m && st.bl_flags && st.freq_hz,
                  "set %d: null pointer argument", i);
        CHECK_ARG(st.images || st.average, "set %d: neither images nor average is given", i);
        CHECK_ARG(!(st.corrections & BLINK_CORR_GEOMETRIC) || st.w_m, "set %d: w_m is null", i);
        CHECK_ARG(!(st.corrections & BLINK_CORR_CABLE) || st.cable_m, "set %d: cable_m is null", i);
        CHECK_ARG(st.pol >= 0 && st.pol <= 3, "set %d: pol must be 0..3", i);
        CHECK_ARG(st.weighting == BLINK_WEIGHTING_NATURAL || st.weighting == BLINK_WEIGHTING_UNIFORM,
                  "set %d: weighting must be NATURAL or UNIFORM", i);
        CHECK_ARG(st.kernel_size >= 0 && st.kernel_size <= BLINK_MAX_KERNEL_SIZE,
                  "set %d: kernel_size must be 0..%d, got %d", i, BLINK_MAX_KERNEL_SIZE,
                  st.kernel_size);
        if (st.kernel_size > 0) {
            CHECK_ARG(st.kernel_mode == BLINK_KERNEL_KAISER_BESSEL ||
                          st.kernel_mode == BLINK_KERNEL_KAISER_BESSEL_NO_SINC,
                      "set %d: unknown kernel mode %d", i, st.kernel_mode);
            if (st.transform != BLINK_TRANSFORM_FFT || st.weighting != BLINK_WEIGHTING_NATURAL)
                return ctx->fail(BLINK_ERR_UNSUPPORTED,
                                 "set %d: batched convolutional gridding runs with the FFT "
                                 "transform and natural weighting (blink_gridding_kernel "
                                 "takes the others)", i);
        }
        CHECK_ARG(((uintptr_t)st.vis & 15) == 0 && ((uintptr_t)st.images & 15) == 0 &&
                      ((uintptr_t)st.average & 15) == 0,
                  "set %d: vis, images and average must be 16-byte aligned", i);
        CHECK_ARG(st.transform == BLINK_TRANSFORM_FFT || is_dft(st.transform),
                  "set %d: unknown transform %d", i, st.transform);
        if (st.n_pixels % 2) {
            // odd sides: the FFT transform with nearest-cell gridding; a mean
            // is taken from the formed images
            if (st.transform != BLINK_TRANSFORM_FFT || st.kernel_size > 0)
                return ctx->fail(BLINK_ERR_UNSUPPORTED,
                                 "set %d: an odd grid side (%d) takes the FFT transform with "
                                 "nearest-cell gridding on the batched path", i, st.n_pixels);
            CHECK_ARG(st.images || !st.average,
                      "set %d: an odd grid side needs `images` (also with an average)", i);
        }
        if (is_dft(st.transform)) {
            if (st.n_pixels < 2 || st.n_pixels > BLINK_MAX_DFT_PIXELS)
                return ctx->fail(BLINK_ERR_UNSUPPORTED,
                                 "set %d: the DFT transform needs 2 <= N <= %d, got %d", i,
                                 BLINK_MAX_DFT_PIXELS, st.n_pixels);
            CHECK_ARG(st.images, "set %d: the DFT transforms need `images` (also with an average)",
                      i);
            continue;
        }
        FftPlan plan;
        if ((rc = check_fft_size(ctx, st.n_pixels, plan))) return rc;
        // above 4096 the generic row tile holds one row and its FFT scratch;
        // the per-cell counts of uniform weighting no longer fit beside them
        if (st.weighting == BLINK_WEIGHTING_UNIFORM && st.n_pixels > 4096)
            return ctx->fail(BLINK_ERR_UNSUPPORTED,
                             "set %d: uniform weighting on the batched path needs N <= 4096, got "
                             "%d (blink_gridding + blink_image take larger sides)", i,
                             st.n_pixels);
    }
    return BLINK_OK;
}

BLINK_API int blink_image_batch(blink_ctx* ctx, const blink_image_set* sets, int n_sets,
                                void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    if (int rc = check_sets(ctx, sets, n_sets)) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    BatchPlan bp;
    if ((rc = build_batch(ctx, sets, n_sets, bp))) return rc;
    hipStream_t s = (hipStream_t)stream;

    if (!ctx->use_graphs || ctx->prof_on) {
        // eager: the descriptor blob goes to a context buffer that no queued
        // work may still read
        void* dblob;
        if ((rc = ctx_scratch(ctx, "batch_desc", bp.blob.size(), &dblob))) return rc;
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy(dblob, bp.blob.data(), bp.blob.size(), hipMemcpyHostToDevice));
        const BatchTables t = tables_at(bp, dblob);
        if ((rc = autotune_batch(ctx, bp, t, s))) return rc;
        return enqueue_batch(ctx, sets, bp, t, s);
    }

    // cache key: the call's fields, then its buffer addresses
    const std::string call = graph_key(sets, n_sets, bp);
    const std::vector<const void*> ptrs = io_pointers(sets, n_sets);
    std::string key = call;
    key.append((const char*)ptrs.data(), sizeof(const void*) * ptrs.size());
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        // the same call with other buffers: a new graph while fewer than
        // kBufferVariants exist, else the least recently used one re-pointed
        auto lru = ctx->graphs.end();
        int variants = 0;
        for (auto g = ctx->graphs.begin(); g != ctx->graphs.end(); ++g)
            if (g->second.call == call) {
                variants++;
                if (lru == ctx->graphs.end() || g->second.last_use < lru->second.last_use) lru = g;
            }
        if (variants >= kBufferVariants) {
            GraphEntry moved = lru->second;
            ctx->graphs.erase(lru);
            it = ctx->graphs.emplace(key, moved).first;
        }
    }
    if (it != ctx->graphs.end()) {
        GraphEntry& ge = it->second;
        if (ge.ptrs != ptrs) {
            // same call with other visibility / image / count-sum buffers:
            // refresh the descriptor blob in stream order, capture the call
            // again and update the instantiated graph's kernel arguments
            // (no new instantiation).  The graph's previous replay, and the
            // copy before it, must have finished with the blob and its source.
            HIP_TRY(hipEventSynchronize(ge.done));
            ge.host_blob = bp.blob;
            HIP_TRY(hipMemcpyAsync(ge.blob, ge.host_blob.data(), ge.host_blob.size(),
                                   hipMemcpyHostToDevice, s));
            hipGraph_t graph = nullptr;
            if ((rc = capture_batch(ctx, sets, bp, tables_at(bp, ge.blob), &graph))) {
                // the blob no longer matches the graph: drop both
                release_graph(ge);
                ctx->graphs.erase(it);
                return rc;
            }
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            hipError_t e = hipGraphExecUpdate(ge.exec, graph, &err_node, &res);
            if (e != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(ge.exec);
                ge.exec = nullptr;
                e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGraphDestroy(graph);
                    release_graph(ge);
                    ctx->graphs.erase(it);
                    return ctx->fail(BLINK_ERR_HIP, "graph instantiation failed: %s",
                                     hipGetErrorString(e));
                }
            }
            (void)hipGraphDestroy(ge.graph);
            ge.graph = graph;
            ge.ptrs = ptrs;
        }
    } else {
        GraphEntry ge;
        ge.call = call;
        HIP_TRY(hipEventCreateWithFlags(&ge.done, hipEventDisableTiming));
        hipError_t e = hipMalloc(&ge.blob, bp.blob.size());
        if (e == hipSuccess)
            e = hipMemcpy(ge.blob, bp.blob.data(), bp.blob.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            release_graph(ge);
            return ctx->fail(BLINK_ERR_HIP, "descriptor upload failed: %s", hipGetErrorString(e));
        }
        const BatchTables t = tables_at(bp, ge.blob);
        if ((rc = autotune_batch(ctx, bp, t, s)) || (rc = capture_batch(ctx, sets, bp, t, &ge.graph))) {
            release_graph(ge);
            return rc;
        }
        e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            release_graph(ge);
            return ctx->fail(BLINK_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        ge.ptrs = ptrs;
        if (ctx->graphs.size() >= kMaxGraphs) {
            auto lru = ctx->graphs.begin();
            for (auto g = ctx->graphs.begin(); g != ctx->graphs.end(); ++g)
                if (g->second.last_use < lru->second.last_use) lru = g;
            // a graph still executing is destroyed only after its last replay
            // completes
            HIP_TRY(hipEventSynchronize(lru->second.done));
            release_graph(lru->second);
            ctx->graphs.erase(lru);
        }
        it = ctx->graphs.emplace(key, ge).first;
    }
    it->second.last_use = ++ctx->graph_clock;
    HIP_TRY(hipGraphLaunch(it->second.exec, s));
    HIP_TRY(hipEventRecord(it->second.done, s));
    return BLINK_OK;
}

// ------------------------------------------------ no-allocation batched call

BLINK_API int blink_image_batch_noalloc(blink_ctx* ctx, const blink_image_set* sets, int n_sets,
                                        void* stream) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    if (int rc = check_sets(ctx, sets, n_sets)) return rc;
    for (int i = 0; i < n_sets; i++) {
        if (is_dft(sets[i].transform))
            return ctx->fail(BLINK_ERR_UNSUPPORTED,
                             "set %d: the no-allocation call runs the FFT transform only", i);
        if (sets[i].kernel_size > 0)
            return ctx->fail(BLINK_ERR_UNSUPPORTED,
                             "set %d: the no-allocation call grids to the nearest cell only", i);
    }
    if (!ctx->noalloc_base)
        return ctx->fail(BLINK_ERR_HIP, "the fixed scratch region has no address on device %d: %s",
                         ctx->device, hipGetErrorString(ctx->noalloc_status));
    int rc = set_device(ctx);
    if (rc) return rc;
    Arena arena;
    arena.base = ctx->noalloc_base;
    arena.cap = BLINK_NOALLOC_SCRATCH_BYTES;
    BatchPlan bp;
    if ((rc = build_batch(ctx, sets, n_sets, bp, &arena))) return rc;
    hipStream_t s = (hipStream_t)stream;
    // the previous call's descriptor copy has read its host source before
    // that source is replaced
    HIP_TRY(hipEventSynchronize(ctx->noalloc_copied));
    ctx->noalloc_blob = bp.blob;
    HIP_TRY(hipMemcpyAsync(bp.blob_dev, ctx->noalloc_blob.data(), ctx->noalloc_blob.size(),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(ctx->noalloc_copied, s));
    // no profiling event pairs (the event pool may grow)
    const bool prof = ctx->prof_on;
    ctx->prof_on = false;
    rc = enqueue_batch(ctx, sets, bp, tables_at(bp, bp.blob_dev), s);
    ctx->prof_on = prof;
    return rc;
}

// ------------------------------------------------------ measurement hooks

BLINK_API int blink_profile_enable(blink_ctx* ctx, int enable) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    ctx->prof_on = enable != 0;
    return BLINK_OK;
}

BLINK_API int blink_profile_get(blink_ctx* ctx, const char* kernel, int64_t* launches,
                                double* total_ms) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(kernel && launches && total_ms, "null pointer argument");
    prof_collect(ctx);
    auto it = ctx->prof_totals.find(kernel);
    *launches = it == ctx->prof_totals.end() ? 0 : it->second.first;
    *total_ms = it == ctx->prof_totals.end() ? 0.0 : it->second.second;
    return BLINK_OK;
}

BLINK_API int blink_profile_kernels(blink_ctx* ctx, char* buf, int buf_len) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(buf && buf_len > 0, "need an output buffer");
    prof_collect(ctx);
    std::string names;
    for (auto& kv : ctx->prof_totals) {
        if (!names.empty()) names += ",";
        names += kv.first;
    }
    CHECK_ARG((int)names.size() < buf_len, "buffer too small (%d bytes needed)",
              (int)names.size() + 1);
    memcpy(buf, names.c_str(), names.size() + 1);
    return BLINK_OK;
}

BLINK_API int blink_profile_reset(blink_ctx* ctx) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    prof_collect(ctx);
    ctx->prof_totals.clear();
    return BLINK_OK;
}

BLINK_API int blink_set_batch_tiles(blink_ctx* ctx, int n_pixels, int row_tile, int col_tile) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(batch_reg_supported(n_pixels), "side %d uses the generic kernels (no tiles)",
              n_pixels);
    if (row_tile == 0 && col_tile == 0) {
        ctx->pinned_tiles.erase(n_pixels);
    } else {
        int t[8];
        const int nr = batch_row_variants(n_pixels, t, 8);
        CHECK_ARG(std::find(t, t + nr, row_tile) != t + nr, "no row tile %d built for side %d",
                  row_tile, n_pixels);
        const int nc = batch_col_variants(n_pixels, t, 8);
        CHECK_ARG(std::find(t, t + nc, col_tile) != t + nc, "no column block %d built for side %d",
                  col_tile, n_pixels);
        ctx->pinned_tiles[n_pixels] = {row_tile, col_tile};
    }
    drop_graphs(ctx);  // captured graphs hold the previous tiles
    return BLINK_OK;
}

BLINK_API int blink_batch_tiles(blink_ctx* ctx, int n_pixels, int* row_tile, int* col_tile) {
    if (!ctx) return BLINK_ERR_INVALID_ARGUMENT;
    CHECK_ARG(row_tile && col_tile, "null pointer argument");
    auto pin = ctx->pinned_tiles.find(n_pixels);
    auto tun = ctx->last_tuned.find(n_pixels);
    if (pin != ctx->pinned_tiles.end() || tun != ctx->last_tuned.end()) {
        const std::pair<int, int> t = pin != ctx->pinned_tiles.end() ? pin->second : tun->second;
        *row_tile = t.first;
        *col_tile = t.second;
    } else if (batch_reg_supported(n_pixels)) {
        const std::pair<int, int> t = default_tiles(n_pixels);
        *row_tile = -t.first;
        *col_tile = -t.second;
    } else {
        *row_tile = *col_tile = 0;
    }
    return BLINK_OK;
}

}  // extern "C"
