// Runtime layer of libbitruss_hip.so: contexts, resident device graphs and
// the C-ABI entry points (argument checks, timing, result copies).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "bt_internal.h"

namespace {
thread_local char g_last_hip_error[512] = "";
}

void bt_set_hip_error(hipError_t e, const char* what, const char* file, int line) {
    std::snprintf(g_last_hip_error, sizeof(g_last_hip_error), "%s: %s (%s:%d)",
                  hipGetErrorString(e), what, file, line);
    std::fprintf(stderr, "[bitruss_hip] HIP error %s\n", g_last_hip_error);
}

int bt_buf_reserve(bt_buf& b, size_t bytes) {
    if (bytes <= b.bytes && b.ptr) return BT_OK;
    size_t const want = bytes ? bytes : 256;
    if (b.owner) {
        if (b.ptr) bt_ctx_release(b.owner, b.ptr, b.bytes);
        b.ptr = nullptr;
        b.bytes = 0;
        size_t got = 0;
        BT_TRY(bt_ctx_alloc(b.owner, want, &b.ptr, &got));
        b.bytes = got;
        return BT_OK;
    }
    if (b.ptr) {
        BT_HIP_TRY(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    BT_HIP_TRY(hipMalloc(&b.ptr, want));
    b.bytes = want;
    return BT_OK;
}

void bt_buf_free(bt_buf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

double bt_timer::ms(int a, int b) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, ctx->ev[a], ctx->ev[b]) != hipSuccess) return -1.0;
    return double(t);
}

int bt_read_ctl(bt_dgraph* dg, int first, int count, uint64_t* out) {
    bt_ctx* c = dg->ctx;
    BT_HIP_TRY(hipMemcpyAsync(c->h_pinned, dg->ctl + first, sizeof(uint64_t) * count,
                              hipMemcpyDeviceToHost, c->stream));
    BT_HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_pinned, sizeof(uint64_t) * count);
    return BT_OK;
}

int bt_ctx_alloc(bt_ctx* c, size_t bytes, void** p, size_t* got) {
    {
        std::lock_guard<std::mutex> lk(c->cache_mu);
        auto it = c->cache.lower_bound(bytes);
        if (it != c->cache.end() && it->first <= 2 * bytes) {
            *p = it->second;
            *got = it->first;  // the block keeps its own size
            c->cache_bytes -= it->first;
            c->cache.erase(it);
            return BT_OK;
        }
    }
    if (hipMalloc(p, bytes) != hipSuccess) {
        // the device may be full of this context's cached blocks: give them
        // back and try once more
        (void)hipGetLastError();
        BT_TRY(bt_ctx_trim_cache(c));
        BT_HIP_TRY(hipMalloc(p, bytes));
    }
    *got = bytes;
    return BT_OK;
}

int bt_ctx_trim_cache(bt_ctx* c) {
    // cached blocks may still be read by work queued on the stream
    BT_HIP_TRY(hipStreamSynchronize(c->stream));
    std::multimap<size_t, void*> drop;
    {
        std::lock_guard<std::mutex> lk(c->cache_mu);
        drop.swap(c->cache);
        c->cache_bytes = 0;
    }
    for (auto const& kv : drop) BT_HIP_TRY(hipFree(kv.second));
    return BT_OK;
}

void bt_ctx_release(bt_ctx* c, void* p, size_t bytes) {
    if (!p) return;
    if (!c) {
        (void)hipFree(p);
        return;
    }
    constexpr size_t kCacheCap = size_t(32) << 30;  // blocks kept per context at most
    {
        std::lock_guard<std::mutex> lk(c->cache_mu);
        if (c->cache_bytes + bytes <= kCacheCap) {
            c->cache.emplace(bytes, p);
            c->cache_bytes += bytes;
            return;
        }
    }
    (void)hipFree(p);
}

namespace {
struct CuPool {
    std::mutex mu;
    std::condition_variable cv;
    int used = 0;
};
CuPool& cu_pool(int device) {
    static CuPool pools[64];
    return pools[static_cast<unsigned>(device) % 64];
}
}  // namespace

bt_cu_lease::bt_cu_lease(int device, int total_cus, int want) : device_(device), want_(want) {
    CuPool& p = cu_pool(device_);
    std::unique_lock<std::mutex> lk(p.mu);
    p.cv.wait(lk, [&] { return p.used == 0 || p.used + want_ <= total_cus; });
    p.used += want_;
}

bt_cu_lease::~bt_cu_lease() {
    CuPool& p = cu_pool(device_);
    {
        std::lock_guard<std::mutex> lk(p.mu);
        p.used -= want_;
    }
    p.cv.notify_all();
}

namespace {

// Graph arrays come from the context's block cache, so dropping a graph
// does not wait for the whole device (as hipFree does) while another
// context's kernels run.
template <typename T>
int dalloc(bt_dgraph* dg, T** p, size_t count) {
    size_t const bytes = sizeof(T) * (count ? count : 1);
    void* q = nullptr;
    size_t got = 0;
    BT_TRY(bt_ctx_alloc(dg->ctx, bytes, &q, &got));
    dg->blocks.emplace_back(q, got);
    *p = static_cast<T*>(q);
    return BT_OK;
}

int check_ctx(bt_ctx* ctx) {
    if (!ctx) return BT_ERR_INVALID;
    BT_HIP_TRY(hipSetDevice(ctx->device));
    return BT_OK;
}

int check_graph(bt_ctx* ctx, bt_dgraph* dg) {
    if (!ctx || !dg || dg->ctx != ctx) return BT_ERR_INVALID;
    return check_ctx(ctx);
}

int copy_out(bt_dgraph* dg, const uint64_t* dsrc, uint64_t* host) {
    if (!host || dg->m == 0) return BT_OK;
    BT_HIP_TRY(hipMemcpyAsync(host, dsrc, sizeof(uint64_t) * dg->m, hipMemcpyDeviceToHost,
                              dg->ctx->stream));
    BT_HIP_TRY(hipStreamSynchronize(dg->ctx->stream));
    return BT_OK;
}

}  // namespace

extern "C" {

int bt_device_count(int* count) {
    if (!count) return BT_ERR_INVALID;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return BT_OK;
}

int bt_device_info(int device, bt_device_props* out) {
    if (!out) return BT_ERR_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return BT_ERR_NO_DEVICE;
    hipDeviceProp_t p{};
    BT_HIP_TRY(hipGetDeviceProperties(&p, device));
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    out->memory_bus_width_bits = p.memoryBusWidth;
    out->memory_clock_khz = p.memoryClockRate;
    out->total_global_mem = p.totalGlobalMem;
    out->multiprocessor_count = p.multiProcessorCount;
    out->shared_mem_per_block = p.sharedMemPerBlock;
    out->max_threads_per_multiprocessor = p.maxThreadsPerMultiProcessor;
    out->warp_size = p.warpSize;
    out->l2_cache_bytes = p.l2CacheSize;
    return BT_OK;
}

int bt_ctx_create(int device, bt_ctx** out) {
    if (!out) return BT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return BT_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return BT_ERR_NO_DEVICE;
    BT_HIP_TRY(hipSetDevice(device));
    bt_ctx* c = new (std::nothrow) bt_ctx();
    if (!c) return BT_ERR_NOMEM;
    c->device = device;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BT_ERR_DEVICE;
    }
    c->cu_count = prop.multiProcessorCount;
    for (int k = 0; k < BT_TUNE_CLASSES; k++) bt_tune_default(k, &c->tune[k]);
    c->coop = prop.cooperativeLaunch;
    c->lds_per_block = prop.sharedMemPerBlock;
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            bt_ctx_destroy(c);
            return BT_ERR_DEVICE;
        }
    }
    for (bt_buf* b : {&c->sort_tmp, &c->pool, &c->pinned_dummy, &c->ix.keys0, &c->ix.keys1, &c->ix.vals0,
                      &c->ix.vals1, &c->ix.bid, &c->ix.bstart, &c->ix.bk, &c->ix.bclaim, &c->ix.inc_off,
                      &c->ix.inc_cur, &c->ix.inc, &c->ix.touched, &c->ix.area, &c->ix.bgraph, &c->ix.glev,
                      &c->ix.live, &c->ix.seg, &c->ix.c2s, &c->ix.hdirty, &c->ix.flip})
        b->owner = c;
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_pinned), 64 * sizeof(uint64_t)) != hipSuccess) {
        bt_ctx_destroy(c);
        return BT_ERR_DEVICE;
    }
    *out = c;
    return BT_OK;
}

int bt_ctx_destroy(bt_ctx* c) {
    if (!c) return BT_ERR_INVALID;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // batches and graphs still alive: their memory goes back with the
    // context's, and their handles stay valid for bt_*_free (which then
    // only deletes them)
    std::vector<bt_batch*> batches;
    {
        std::lock_guard<std::mutex> lk(c->live_mu);
        batches.assign(c->live_batches.begin(), c->live_batches.end());
        c->live_batches.clear();
    }
    for (bt_batch* b : batches) bt_batch_detach(b);
    std::vector<bt_dgraph*> graphs;
    {
        std::lock_guard<std::mutex> lk(c->live_mu);
        graphs.assign(c->live_graphs.begin(), c->live_graphs.end());
        c->live_graphs.clear();
    }
    for (bt_dgraph* dg : graphs) {
        for (auto const& b : dg->blocks) bt_ctx_release(c, b.first, b.second);
        dg->blocks.clear();
        dg->ctx = nullptr;
    }
    bt_buf_free(c->sort_tmp);
    bt_buf_free(c->pool);
    bt_buf_free(c->pinned_dummy);
    bt_buf* ixb[] = {&c->ix.keys0, &c->ix.keys1, &c->ix.vals0, &c->ix.vals1, &c->ix.bid, &c->ix.bstart,
                     &c->ix.bk, &c->ix.bclaim, &c->ix.inc_off, &c->ix.inc_cur, &c->ix.inc, &c->ix.touched, &c->ix.area, &c->ix.bgraph, &c->ix.glev,
                     &c->ix.live, &c->ix.seg, &c->ix.c2s, &c->ix.hdirty, &c->ix.flip};
    for (bt_buf* b : ixb) bt_buf_free(*b);
    for (auto const& kv : c->cache) (void)hipFree(kv.second);
    c->cache.clear();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BT_OK;
}

int bt_ctx_set_option(bt_ctx* ctx, int option, int64_t value) {
    if (!ctx) return BT_ERR_INVALID;
    switch (option) {
        case BT_OPT_MSP_PATH:
            if (value < BT_MSP_PATH_AUTO || value > BT_MSP_PATH_INDEX) return BT_ERR_INVALID;
            ctx->msp_path = int(value);
            return BT_OK;
        case BT_OPT_INDEX_MEM_PCT:
            if (value < 1 || value > 95) return BT_ERR_INVALID;
            ctx->index_mem_frac = double(value) / 100.0;
            return BT_OK;
        case BT_OPT_BATCH_LEVELS:
            if (value != BT_BATCH_LEVELS_PER_GRAPH && value != BT_BATCH_LEVELS_UNION) return BT_ERR_INVALID;
            ctx->batch_levels = int(value);
            return BT_OK;
        case BT_OPT_PEEL_BLOCKS:
            if (value < 0 || value > ctx->cu_count) return BT_ERR_INVALID;
            ctx->peel_blocks = int(value);
            return BT_OK;
        case BT_OPT_BATCH_LANES:
            if (value < 1 || value > 64) return BT_ERR_INVALID;
            ctx->batch_lanes = int(value);
            return BT_OK;
        case BT_OPT_PEEL_SLICE_MS:
            if (value < 0 || value > 3600000) return BT_ERR_INVALID;
            ctx->peel_slice_ms = int(value);
            return BT_OK;
        case BT_OPT_BATCH_WEDGES:
            if (value < 0) return BT_ERR_INVALID;
            ctx->batch_wedges = uint64_t(value);
            return BT_OK;
        case BT_OPT_PEEL_MAX_LEVEL:
            if (value < 0) return BT_ERR_INVALID;
            ctx->peel_max_level = uint64_t(value);
            return BT_OK;
        case BT_OPT_PEEL_C2CAP:
            if (value < 0 || value > 128) return BT_ERR_INVALID;
            ctx->peel_c2cap = int(value);
            return BT_OK;
        case BT_OPT_PEEL_WINDOW_EDGES:
            if (value < 0) return BT_ERR_INVALID;
            ctx->peel_window_edges = uint64_t(value);
            return BT_OK;
        case BT_OPT_HIDX_GRAPH_ROUNDS:
            if (value < 0 || value > 1024) return BT_ERR_INVALID;
            ctx->hidx_graph_rounds = int(value);
            return BT_OK;
        case BT_OPT_HIDX_SORTED:
            if (value < 0 || value > 1) return BT_ERR_INVALID;
            ctx->hidx_sorted = int(value);
            return BT_OK;
        case BT_OPT_PEEL_BIG_MIDDLES:
            if (value < 1 || value > 0x7FFFFFFF) return BT_ERR_INVALID;
            ctx->peel_big_middles = int(value);
            return BT_OK;
        default:
            return BT_ERR_INVALID;
    }
}

int bt_ctx_trim(bt_ctx* ctx) {
    BT_TRY(check_ctx(ctx));
    BT_HIP_TRY(hipSetDevice(ctx->device));
    BT_HIP_TRY(hipStreamSynchronize(ctx->stream));
    bt_buf* bufs[] = {&ctx->sort_tmp, &ctx->pool, &ctx->ix.keys0, &ctx->ix.keys1, &ctx->ix.vals0, &ctx->ix.vals1,
                      &ctx->ix.bid, &ctx->ix.bstart, &ctx->ix.bk, &ctx->ix.bclaim, &ctx->ix.inc_off,
                      &ctx->ix.inc_cur, &ctx->ix.inc, &ctx->ix.touched, &ctx->ix.area, &ctx->ix.bgraph,
                      &ctx->ix.glev, &ctx->ix.live, &ctx->ix.seg, &ctx->ix.c2s, &ctx->ix.hdirty,
                      &ctx->ix.flip};
    for (bt_buf* b : bufs) bt_buf_free(*b);
    ctx->ix.R = 0;
    ctx->ix.NB = 0;
    return bt_ctx_trim_cache(ctx);
}

int bt_ctx_last_timings(const bt_ctx* ctx, bt_timings* out) {
    if (!ctx || !out) return BT_ERR_INVALID;
    *out = ctx->last;
    return BT_OK;
}

int bt_graph_free(bt_dgraph* dg) {
    if (!dg) return BT_ERR_INVALID;
    if (dg->ctx) {
        (void)hipSetDevice(dg->ctx->device);
        (void)hipStreamSynchronize(dg->ctx->stream);
        std::lock_guard<std::mutex> lk(dg->ctx->live_mu);
        dg->ctx->live_graphs.erase(dg);
    }
    for (auto const& b : dg->blocks) bt_ctx_release(dg->ctx, b.first, b.second);
    delete dg;
    return BT_OK;
}

int bt_graph_upload(bt_ctx* ctx, const bt_host_graph* g, bt_dgraph** out) {
    if (!out) return BT_ERR_INVALID;
    *out = nullptr;
    BT_TRY(check_ctx(ctx));
    if (!g) return BT_ERR_INVALID;
    // host-side validation: the kernels index with these arrays
    BT_TRY(bt_graph_check(g));
    bt_dgraph* dg = new (std::nothrow) bt_dgraph();
    if (!dg) return BT_ERR_NOMEM;
    dg->ctx = ctx;
    dg->u_num = g->u_num;
    dg->l_num = g->l_num;
    dg->n = g->n;
    dg->m = g->m;
    size_t const n = g->n, m = g->m;
    int st = BT_OK;
    auto A = [&](int s) {
        if (st == BT_OK) st = s;
    };
    A(dalloc(dg, &dg->nbr, 2 * m));
    A(dalloc(dg, &dg->off, n + 1));
    A(dalloc(dg, &dg->deg, n));
    A(dalloc(dg, &dg->edges, 2 * m));
    A(dalloc(dg, &dg->eids, 2 * m));
    A(dalloc(dg, &dg->rank, n));
    A(dalloc(dg, &dg->r_off, n + 1));
    A(dalloc(dg, &dg->r_nbr, 2 * m));
    A(dalloc(dg, &dg->r_eid, 2 * m));
    A(dalloc(dg, &dg->r_lo, n));
    A(dalloc(dg, &dg->r_up, n));
    A(dalloc(dg, &dg->r_p, 2 * m));
    A(dalloc(dg, &dg->r_w, n + 1));
    A(dalloc(dg, &dg->sup, m));
    A(dalloc(dg, &dg->res, m));
    A(dalloc(dg, &dg->aux, m));
    A(dalloc(dg, &dg->state, m));
    A(dalloc(dg, &dg->list0, m + n + 1));
    A(dalloc(dg, &dg->list1, m + n + 1));
    A(dalloc(dg, &dg->ctl, 128));
    A(dalloc(dg, &dg->r_key, 2 * m + n));
    A(dalloc(dg, &dg->r_key2, 2 * m + n));
    A(dalloc(dg, &dg->r_val2, 2 * m + n));
    if (st != BT_OK) {
        bt_graph_free(dg);
        return st == BT_ERR_DEVICE ? BT_ERR_BUDGET : st;
    }
    hipStream_t s = ctx->stream;
    auto H2D = [&](void* d, const void* h, size_t bytes) {
        if (st != BT_OK || bytes == 0) return;
        if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s) != hipSuccess) st = BT_ERR_DEVICE;
    };
    H2D(dg->nbr, g->neighbors, 8 * m);
    H2D(dg->off, g->offsets, 4 * (n + 1));
    H2D(dg->deg, g->degrees, 4 * n);
    H2D(dg->edges, g->edges, 8 * m);
    H2D(dg->eids, g->edge_ids, 8 * m);
    if (st == BT_OK && hipStreamSynchronize(s) != hipSuccess) st = BT_ERR_DEVICE;
    if (st != BT_OK) {
        bt_graph_free(dg);
        return st;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->live_mu);
        ctx->live_graphs.insert(dg);
    }
    *out = dg;
    return BT_OK;
}

int bt_bfc(bt_ctx* ctx, bt_dgraph* dg, uint64_t* total) {
    BT_TRY(check_graph(ctx, dg));
    uint64_t t = 0;
    BT_TRY(bt_count_impl(dg, false, &t, nullptr));
    if (total) *total = t;
    return BT_OK;
}

int bt_ebfc(bt_ctx* ctx, bt_dgraph* dg, uint64_t* support_out, uint64_t* total,
            uint64_t* support_max) {
    BT_TRY(check_graph(ctx, dg));
    uint64_t t = 0, smax = 0;
    BT_TRY(bt_count_impl(dg, true, &t, &smax));
    BT_TRY(copy_out(dg, dg->sup, support_out));
    if (total) *total = t;
    if (support_max) *support_max = smax;
    return BT_OK;
}

int bt_hidx(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* rounds) {
    BT_TRY(check_graph(ctx, dg));
    uint32_t r = 0;
    BT_TRY(bt_hidx_impl(dg, &r));
    BT_TRY(copy_out(dg, dg->res, bitruss_out));
    if (rounds) *rounds = r;
    return BT_OK;
}

namespace {
int msp_family(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels, int rule) {
    BT_TRY(check_graph(ctx, dg));
    uint32_t l = 0;
    int const st = bt_msp_impl(dg, &l, rule);
    if (st != BT_OK && st != BT_PARTIAL) return st;
    BT_TRY(copy_out(dg, dg->res, bitruss_out));
    if (levels) *levels = l;
    return st;
}
}  // namespace

int bt_msp(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels) {
    return msp_family(ctx, dg, bitruss_out, levels, 0);
}

int bt_mspp(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels) {
    return msp_family(ctx, dg, bitruss_out, levels, 1);
}

int bt_peel(bt_ctx* ctx, bt_dgraph* dg, uint64_t* bitruss_out, uint32_t* levels) {
    return msp_family(ctx, dg, bitruss_out, levels, 2);
}

}  // extern "C"
