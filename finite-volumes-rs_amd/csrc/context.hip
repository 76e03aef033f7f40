// context.hip -- context, device memory accounting, vectors, shared vector kernels.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <atomic>
#include <mutex>

#include "internal.hpp"
#include "reduce.hpp"

namespace fvg {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

void release_work_cache(fvg_ctx* ctx) {
    if (ctx->work_cache.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->work_cache) dev_free(ctx, b.first, (size_t)b.second * sizeof(double));
    ctx->work_cache.clear();
}

int dev_alloc(fvg_ctx* ctx, size_t bytes, void** out) {
    *out = nullptr;
    if (ctx->budget && ctx->in_use + bytes > ctx->budget) release_work_cache(ctx);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        return fail(FVG_E_BUDGET, "allocation of " + std::to_string(bytes) + " bytes exceeds the HBM budget (" +
                                      std::to_string(ctx->in_use) + " of " + std::to_string(ctx->budget) +
                                      " in use)");
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess && !ctx->work_cache.empty()) {
        (void)hipGetLastError();
        release_work_cache(ctx);
        e = hipMalloc(out, bytes);
    }
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(FVG_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    ctx->in_use += bytes;
    return FVG_OK;
}

void dev_free(fvg_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    ctx->in_use = ctx->in_use >= bytes ? ctx->in_use - bytes : 0;
}

// ---- kernels ------------------------------------------------------------------------

__global__ void k_fill(int64_t n, double* __restrict__ x, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_copy(int64_t n, const double* __restrict__ s, double* __restrict__ d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// Each block sums a contiguous tile of kBlock*kItems elements (item-major per thread,
// then the block tree); the tile decomposition depends only on n.
constexpr int kItems = 8;

__global__ void __launch_bounds__(kBlock) k_dot_partial(int64_t n, const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    const int64_t tiles = gridDim.x;
    const int64_t per = (n + tiles - 1) / tiles;
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < n ? lo + per : n;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) s += a[i] * b[i];
    s = block_sum<kBlock>(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_reduce_final(const double* __restrict__ partials, int count,
                                                         double* __restrict__ out) {
    __shared__ double lds[kBlock / 64];
    double s = sum_partials<kBlock>(partials, count, lds);
    if (threadIdx.x == 0) *out = s;
}

int reduce_blocks_for(int64_t n) {
    int64_t g = (n + (int64_t)kBlock * kItems - 1) / ((int64_t)kBlock * kItems);
    if (g < 1) g = 1;
    if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
    return (int)g;
}

int launch_dot(fvg_ctx* ctx, int64_t n, const double* a, const double* b, double* out_dev) {
    int g = reduce_blocks_for(n);
    hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, ctx->stream, n, a, b, ctx->partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, g, out_dev);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int launch_fill(fvg_ctx* ctx, int64_t n, double* x, double value) {
    if (n <= 0) return FVG_OK;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, x, value);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int launch_copy(fvg_ctx* ctx, int64_t n, const double* src, double* dst) {
    if (n <= 0) return FVG_OK;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, src, dst);
    FVG_HIP(hipGetLastError());
    return FVG_OK;
}

int read_scalars(fvg_ctx* ctx, const double* dev, double* host, int count) {
    FVG_HIP(hipMemcpyAsync(ctx->pinned, dev, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    FVG_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, ctx->pinned, sizeof(double) * count);
    return FVG_OK;
}

// Timing records an event pair around one launch group in every `timing_every` groups
// (events reused from a pool), so that the markers do not slow the stream they measure.
int timing_begin(fvg_ctx* ctx) {
    if (!ctx->timing) return -1;
    if (ctx->timing_skip > 0) {
        ctx->timing_skip -= 1;
        return -1;
    }
    ctx->timing_skip = ctx->timing_every - 1;
    fvg_ctx::Span s{};
    for (hipEvent_t* e : {&s.a, &s.b}) {
        if (!ctx->event_pool.empty()) {
            *e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            return -1;
        }
    }
    (void)hipEventRecord(s.a, ctx->stream);
    ctx->spans.push_back(s);
    return (int)ctx->spans.size() - 1;
}

void timing_end(fvg_ctx* ctx, int token, int kind, int64_t launches) {
    if (token < 0 || token >= (int)ctx->spans.size()) return;
    auto& s = ctx->spans[(size_t)token];
    (void)hipEventRecord(s.b, ctx->stream);
    s.kind = kind;
    s.launches = launches;
}

static int collect_spans(fvg_ctx* ctx) {
    if (ctx->spans.empty()) return FVG_OK;
    FVG_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& s : ctx->spans) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
            if (s.kind == kSpanSpmv) {
                ctx->stats.spmv_ms += ms;
                ctx->stats.spmv_launches += s.launches;
                ctx->stats.spmv_calls += 1;
            } else {
                ctx->stats.sweep_ms += ms;
                ctx->stats.sweep_launches += s.launches;
                ctx->stats.sweep_calls += 1;
            }
        }
        ctx->event_pool.push_back(s.a);
        ctx->event_pool.push_back(s.b);
    }
    ctx->spans.clear();
    return FVG_OK;
}

// contexts alive per device in this process
constexpr int kMaxDevices = 64;
static std::atomic<int> g_live[kMaxDevices];

// Launch mode of a kernel whose workgroups wait on one another (1 ordinary, 2 cooperative).
// Such workgroups deadlock only if some of them never become resident.  The caller sizes
// the grid to fit the device, so once the stream's previous kernel drains they all become
// resident, unless other kernels hold compute units for as long as the waits last: kernels
// of another context of this process (which may themselves wait on one another), or a
// batch of this context in flight on its own stream (a persistent kernel running up to
// hundreds of milliseconds).  Only with neither is the ordinary launch, without the
// cooperative launch's hand-off to another queue and back, chosen by default.  After a
// wait of this context gave up, launches stay cooperative.  Kernels of other processes are
// not visible here: a wait they delay past its bound ends the call with FVG_E_TIMEOUT,
// after which this context launches cooperatively.
int sweep_launch_rule(int64_t tune, int live_contexts, int batches_in_flight, bool timed_out_before) {
    if (tune == 1 || tune == 2) return (int)tune;
    return live_contexts == 1 && batches_in_flight == 0 && !timed_out_before ? 1 : 2;
}

hipError_t launch_coresident(fvg_ctx* ctx, const void* fn, unsigned groups, unsigned threads, void** args,
                             size_t lds) {
    const int live = ctx->device < kMaxDevices ? g_live[ctx->device].load(std::memory_order_relaxed) : 2;
    const int mode = sweep_launch_rule(ctx->tune_sweep_launch, live, ctx->batches_in_flight, ctx->sweep_timed_out);
    ctx->last_sweep_launch = mode;
    return mode == 2 ? hipLaunchCooperativeKernel(fn, dim3(groups), dim3(threads), args, lds, ctx->stream)
                     : hipLaunchKernel(fn, dim3(groups), dim3(threads), args, lds, ctx->stream);
}

}  // namespace fvg

using namespace fvg;

extern "C" {

const char* fvg_last_error(void) { return g_last_error.c_str(); }

const char* fvg_status_string(int s) {
    switch (s) {
    case FVG_OK: return "ok";
    case FVG_E_INVALID: return "invalid argument";
    case FVG_E_HIP: return "HIP runtime error";
    case FVG_E_BUDGET: return "HBM budget exceeded";
    case FVG_E_ZERO_PIVOT: return "zero pivot in ILU";
    case FVG_E_NOT_SYMMETRIC: return "matrix not symmetric";
    case FVG_E_MISSING_ENTRY: return "sparse matrix entry not found";
    case FVG_E_MESH: return "invalid mesh";
    case FVG_E_TIMEOUT: return "device wait timed out";
    case FVG_E_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

int fvg_ctx_create(int device, uint64_t budget, fvg_ctx** out) {
    if (!out) return fail(FVG_E_INVALID, "null output");
    *out = nullptr;
    int count = 0;
    FVG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(FVG_E_INVALID, "no such device");
    FVG_HIP(hipSetDevice(device));
    auto* c = new fvg_ctx();
    c->device = device;
    (void)hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    c->budget = budget;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(FVG_E_HIP, "hipStreamCreate failed");
    }
    int s = dev_alloc_n(c, (int64_t)kMaxReduceBlocks * kPartialSlots, &c->partials);
    c->partials_cap = kMaxReduceBlocks;
    if (s == FVG_OK) s = dev_alloc_n(c, 64, &c->scalar);
    if (s == FVG_OK && hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault) != hipSuccess)
        s = fail(FVG_E_HIP, "hipHostMalloc failed");
    if (s == FVG_OK && hipHostMalloc((void**)&c->poll_host, 64, hipHostMallocDefault) != hipSuccess)
        s = fail(FVG_E_HIP, "hipHostMalloc failed");
    for (int k = 0; s == FVG_OK && k < 2; ++k)
        if (hipEventCreateWithFlags(&c->poll_evt[k], hipEventDisableTiming) != hipSuccess)
            s = fail(FVG_E_HIP, "hipEventCreate failed");
    if (s != FVG_OK) {
        fvg_ctx_destroy(c);
        return s;
    }
    if (device < fvg::kMaxDevices) fvg::g_live[device].fetch_add(1);
    c->counted = true;
    *out = c;
    return FVG_OK;
}

int fvg_ctx_destroy(fvg_ctx* c) {
    if (!c) return FVG_OK;
    if (c->counted && c->device < fvg::kMaxDevices) fvg::g_live[c->device].fetch_sub(1);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& s : c->spans) {
        (void)hipEventDestroy(s.a);
        (void)hipEventDestroy(s.b);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    release_work_cache(c);
    dev_free_n(c, c->partials, c->partials_cap * kPartialSlots);
    dev_free_n(c, c->scalar, 64);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->poll_host) (void)hipHostFree(c->poll_host);
    for (hipEvent_t e : c->poll_evt)
        if (e) (void)hipEventDestroy(e);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->jobs) dev_free(c, c->jobs, c->jobs_cap);
    if (c->jobs_host) (void)hipHostFree(c->jobs_host);
    if (c->jobs_evt) (void)hipEventDestroy(c->jobs_evt);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FVG_OK;
}

int fvg_ctx_sync(fvg_ctx* c) {
    if (!c) return fail(FVG_E_INVALID, "null context");
    FVG_HIP(hipStreamSynchronize(c->stream));
    return FVG_OK;
}

int fvg_ctx_bytes_in_use(const fvg_ctx* c, uint64_t* bytes) {
    if (!c || !bytes) return fail(FVG_E_INVALID, "null argument");
    *bytes = c->in_use;
    return FVG_OK;
}

int fvg_ctx_stream(const fvg_ctx* c, void** stream) {
    if (!c || !stream) return fail(FVG_E_INVALID, "null argument");
    *stream = (void*)c->stream;
    return FVG_OK;
}

int fvg_ctx_set_tuning(fvg_ctx* c, int32_t knob, int64_t value) {
    if (!c || value < 0) return fail(FVG_E_INVALID, "bad tuning argument");
    switch (knob) {
    case FVG_TUNE_SWEEP_MODE:
        if (value > 2) return fail(FVG_E_INVALID, "sweep mode must be 0, 1 or 2");
        c->tune_sweep_mode = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_WAVES:
        c->tune_sweep_waves = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_POLL:
        if (value > 2) return fail(FVG_E_INVALID, "poll variant must be 0, 1 or 2");
        c->tune_sweep_poll = value;
        return FVG_OK;
    case FVG_TUNE_INNER_EXACT:
        if (value > 1) return fail(FVG_E_INVALID, "inner sweep mode must be 0 or 1");
        c->tune_inner_exact = value;
        return FVG_OK;
    case FVG_TUNE_INNER_SWEEPS:
        if (value > 64) return fail(FVG_E_INVALID, "at most 64 Jacobi sweeps");
        c->tune_inner_sweeps = value;
        return FVG_OK;
    case FVG_TUNE_GRAPHS:
        if (value > 1) return fail(FVG_E_INVALID, "launch mode must be 0 or 1");
        c->tune_graphs = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_BLOCKS:
        if (value > 4096) return fail(FVG_E_INVALID, "at most 4096 sweep blocks");
        c->tune_sweep_blocks = value;
        return FVG_OK;
    case FVG_TUNE_SPIN_LIMIT:
        if (value > (1 << 30)) return fail(FVG_E_INVALID, "spin limit above 2^30");
        c->tune_spin_limit = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_LAUNCH:
        if (value > 2) return fail(FVG_E_INVALID, "sweep launch must be 0, 1 or 2");
        c->tune_sweep_launch = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_WITHHOLD:
        if (value > 1) return fail(FVG_E_INVALID, "withhold must be 0 or 1");
        c->tune_withhold = value;
        return FVG_OK;
    case FVG_TUNE_SOLVE_ORDER:
        if (value > 1) return fail(FVG_E_INVALID, "solve order must be 0 or 1");
        c->tune_solve_order = value;
        return FVG_OK;
    case FVG_TUNE_ASSEMBLE:
        if (value > 2) return fail(FVG_E_INVALID, "assemble kernel must be 0, 1 or 2");
        c->tune_assemble = value;
        return FVG_OK;
    case FVG_TUNE_CG_FUSION:
        if (value > 1) return fail(FVG_E_INVALID, "CG fusion must be 0 or 1");
        c->tune_cg_fusion = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_GAP:
        if (value == 1 || value > 8) return fail(FVG_E_INVALID, "sweep gap must be 0 or 2 to 8");
        c->tune_sweep_gap = value;
        return FVG_OK;
    case FVG_TUNE_SWEEP_BOXES:
        if (value > 1024) return fail(FVG_E_INVALID, "sweep boxes must be 0, 1 or 2 to 1024");
        c->tune_sweep_boxes = value;
        return FVG_OK;
    }
    return fail(FVG_E_INVALID, "unknown tuning knob");
}

int fvg_ctx_get_tuning(const fvg_ctx* c, int32_t knob, int64_t* value) {
    if (!c || !value) return fail(FVG_E_INVALID, "null argument");
    switch (knob) {
    case FVG_TUNE_SWEEP_MODE: *value = c->tune_sweep_mode; return FVG_OK;
    case FVG_TUNE_SWEEP_WAVES: *value = c->tune_sweep_waves; return FVG_OK;
    case FVG_TUNE_SWEEP_POLL: *value = c->tune_sweep_poll; return FVG_OK;
    case FVG_TUNE_INNER_EXACT: *value = c->tune_inner_exact; return FVG_OK;
    case FVG_TUNE_INNER_SWEEPS: *value = c->tune_inner_sweeps; return FVG_OK;
    case FVG_TUNE_GRAPHS: *value = c->tune_graphs; return FVG_OK;
    case FVG_TUNE_SWEEP_BLOCKS: *value = c->tune_sweep_blocks; return FVG_OK;
    case FVG_TUNE_SPIN_LIMIT: *value = c->tune_spin_limit; return FVG_OK;
    case FVG_TUNE_SWEEP_LAUNCH: *value = c->tune_sweep_launch; return FVG_OK;
    case FVG_TUNE_SWEEP_WITHHOLD: *value = c->tune_withhold; return FVG_OK;
    case FVG_TUNE_SOLVE_ORDER: *value = c->tune_solve_order; return FVG_OK;
    case FVG_TUNE_ASSEMBLE: *value = c->tune_assemble; return FVG_OK;
    case FVG_TUNE_CG_FUSION: *value = c->tune_cg_fusion; return FVG_OK;
    case FVG_TUNE_SWEEP_GAP: *value = c->tune_sweep_gap; return FVG_OK;
    case FVG_TUNE_SWEEP_BOXES: *value = c->tune_sweep_boxes; return FVG_OK;
    }
    return fail(FVG_E_INVALID, "unknown tuning knob");
}

int fvg_ctx_sweep_launch_state(const fvg_ctx* c, int32_t* last_launch, int32_t* batches_in_flight) {
    if (!c) return fail(FVG_E_INVALID, "null context");
    if (last_launch) *last_launch = c->last_sweep_launch;
    if (batches_in_flight) *batches_in_flight = c->batches_in_flight;
    return FVG_OK;
}

int fvg_sweep_launch_rule(int64_t tune, int32_t live_contexts, int32_t batches_in_flight, int32_t timed_out_before) {
    if (tune < 0 || tune > 2 || live_contexts < 0 || batches_in_flight < 0) return 0;
    return fvg::sweep_launch_rule(tune, live_contexts, batches_in_flight, timed_out_before != 0);
}

int fvg_ctx_enable_timing(fvg_ctx* c, int enable) {
    if (!c || enable < 0) return fail(FVG_E_INVALID, "bad timing argument");
    c->timing = enable != 0;
    c->timing_every = enable > 0 ? enable : 1;
    c->timing_skip = 0;
    return FVG_OK;
}

int fvg_ctx_kernel_stats(fvg_ctx* c, fvg_kernel_stats* out, int reset) {
    if (!c || !out) return fail(FVG_E_INVALID, "null argument");
    FVG_TRY(collect_spans(c));
    *out = c->stats;
    if (reset) c->stats = fvg_kernel_stats{};
    return FVG_OK;
}

int fvg_vec_create(fvg_ctx* c, int64_t n, fvg_vec** out) {
    if (!c || !out || n < 0) return fail(FVG_E_INVALID, "bad vector arguments");
    auto* v = new fvg_vec();
    v->n = n;
    int s = dev_alloc_n(c, n, &v->d);
    if (s != FVG_OK) {
        delete v;
        return s;
    }
    if (n > 0) {
        s = launch_fill(c, n, v->d, 0.0);
        if (s != FVG_OK) {
            dev_free_n(c, v->d, n);
            delete v;
            return s;
        }
    }
    *out = v;
    return FVG_OK;
}

int fvg_vec_wrap(fvg_ctx* c, int64_t n, double* ptr, fvg_vec** out) {
    if (!c || !out || n < 0 || (!ptr && n > 0)) return fail(FVG_E_INVALID, "bad vector arguments");
    auto* v = new fvg_vec();
    v->n = n;
    v->d = ptr;
    v->owned = false;
    *out = v;
    return FVG_OK;
}

int fvg_vec_destroy(fvg_ctx* c, fvg_vec* v) {
    if (!v) return FVG_OK;
    if (v->owned && c) {
        (void)hipStreamSynchronize(c->stream);
        dev_free_n(c, v->d, v->n);
    }
    delete v;
    return FVG_OK;
}

int fvg_vec_len(const fvg_vec* v, int64_t* n) {
    if (!v || !n) return fail(FVG_E_INVALID, "null argument");
    *n = v->n;
    return FVG_OK;
}

int fvg_vec_device_ptr(const fvg_vec* v, double** p) {
    if (!v || !p) return fail(FVG_E_INVALID, "null argument");
    *p = v->d;
    return FVG_OK;
}

// Host transfers go through a pinned staging buffer, so the copy engines move the data
// (a copy from or to pageable memory may otherwise need compute units, which long-running
// kernels can hold).
static constexpr size_t kStageMax = size_t(32) << 20;

static int stage_reserve(fvg_ctx* c, size_t bytes) {
    bytes = std::min(bytes, kStageMax);
    if (c->stage_cap >= bytes) return FVG_OK;
    if (c->stage) {
        FVG_HIP(hipStreamSynchronize(c->stream));
        (void)hipHostFree(c->stage);
        c->stage = nullptr;
        c->stage_cap = 0;
    }
    size_t cap = size_t(1) << 16;
    while (cap < bytes) cap <<= 1;
    FVG_HIP(hipHostMalloc(&c->stage, cap, hipHostMallocDefault));
    c->stage_cap = cap;
    return FVG_OK;
}

int fvg_vec_upload(fvg_ctx* c, fvg_vec* v, const double* host) {
    if (!c || !v || (!host && v->n)) return fail(FVG_E_INVALID, "null argument");
    const size_t total = sizeof(double) * (size_t)v->n;
    if (total == 0) return FVG_OK;
    FVG_TRY(stage_reserve(c, total));
    const size_t chunk = c->stage_cap;
    for (size_t off = 0; off < total; off += chunk) {
        const size_t len = std::min(chunk, total - off);
        FVG_HIP(hipStreamSynchronize(c->stream));   // the previous chunk has left the buffer
        std::memcpy(c->stage, reinterpret_cast<const char*>(host) + off, len);
        FVG_HIP(hipMemcpyAsync(reinterpret_cast<char*>(v->d) + off, c->stage, len, hipMemcpyHostToDevice, c->stream));
    }
    FVG_HIP(hipStreamSynchronize(c->stream));
    return FVG_OK;
}

int fvg_vec_upload_many(fvg_ctx* c, int32_t count, fvg_vec* const* vecs, const double* const* hosts) {
    if (!c || count < 0 || (count && (!vecs || !hosts))) return fail(FVG_E_INVALID, "null argument");
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        if (!vecs[i] || (!hosts[i] && vecs[i]->n)) return fail(FVG_E_INVALID, "null vector or data");
        total += sizeof(double) * (size_t)vecs[i]->n;
    }
    if (total == 0) return FVG_OK;
    if (total > kStageMax) {   // too large to stage at once: one vector at a time
        for (int i = 0; i < count; ++i) FVG_TRY(fvg_vec_upload(c, vecs[i], hosts[i]));
        return FVG_OK;
    }
    FVG_TRY(stage_reserve(c, total));
    FVG_HIP(hipStreamSynchronize(c->stream));   // earlier copies have left the staging buffer
    size_t off = 0;
    for (int i = 0; i < count; ++i) {
        const size_t len = sizeof(double) * (size_t)vecs[i]->n;
        if (len == 0) continue;
        std::memcpy(reinterpret_cast<char*>(c->stage) + off, hosts[i], len);
        FVG_HIP(hipMemcpyAsync(vecs[i]->d, reinterpret_cast<char*>(c->stage) + off, len, hipMemcpyHostToDevice,
                               c->stream));
        off += len;
    }
    return FVG_OK;   // the copies complete in stream order; the staging buffer is reused only after a sync
}

int fvg_vec_download(fvg_ctx* c, const fvg_vec* v, double* host) {
    if (!c || !v || (!host && v->n)) return fail(FVG_E_INVALID, "null argument");
    const size_t total = sizeof(double) * (size_t)v->n;
    if (total == 0) return FVG_OK;
    FVG_TRY(stage_reserve(c, total));
    const size_t chunk = c->stage_cap;
    for (size_t off = 0; off < total; off += chunk) {
        const size_t len = std::min(chunk, total - off);
        FVG_HIP(hipMemcpyAsync(c->stage, reinterpret_cast<const char*>(v->d) + off, len, hipMemcpyDeviceToHost,
                               c->stream));
        FVG_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(reinterpret_cast<char*>(host) + off, c->stage, len);
    }
    return FVG_OK;
}

int fvg_vec_fill(fvg_ctx* c, fvg_vec* v, double value) {
    if (!c || !v) return fail(FVG_E_INVALID, "null argument");
    return launch_fill(c, v->n, v->d, value);
}

int fvg_vec_dot(fvg_ctx* c, const fvg_vec* a, const fvg_vec* b, double* out) {
    if (!c || !a || !b || !out) return fail(FVG_E_INVALID, "null argument");
    if (a->n != b->n) return fail(FVG_E_INVALID, "vector lengths differ");
    FVG_TRY(launch_dot(c, a->n, a->d, b->d, c->scalar));
    return read_scalars(c, c->scalar, out, 1);
}

}  // extern "C"
