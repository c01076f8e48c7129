// Counting launch settings per graph-size class, and the autotuner that
// measures them.
//
// The reference fixes its counting launches at compile time (BLK_NUMS and
// BLK_DIM, src/util/config.h:9-10). Here the bucket bounds, the LDS table
// sizes that go with them and the workgroup sizes are chosen per size class
// (floor(log4(edges per graph))) from a table measured on MI355X
// (bt_tune_table.h); bt_autotune times every candidate on a given graph,
// checks that all of them give the same supports, and keeps the fastest for
// the graph's class. The settings never change a result: every candidate
// counts the same wedges, only the kernel that counts a root differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bt_common.h"
#include "bt_internal.h"
#include "bt_tune_table.h"

using namespace bt;

namespace {

constexpr uint32_t kWaveWedges[] = {256, 512, 1024, 2048};
constexpr uint32_t kBlockWedges[] = {4096, 8192, 16384};
constexpr uint32_t kThreads[] = {256, 512};

std::vector<bt_count_config> candidates() {
    std::vector<bt_count_config> v;
    for (uint32_t a : kWaveWedges)
        for (uint32_t b : kBlockWedges)
            for (uint32_t t : kThreads)
                for (uint32_t h : kThreads) v.push_back(bt_count_config{a, b, t, h});
    return v;
}

// Sets flag[0] when a and b differ anywhere.
__global__ void k_differs(uint32_t m, const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                          unsigned long long* flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        if (a[i] != b[i]) {
            atomicOr(flag, 1ull);
            return;
        }
}

}  // namespace

bool bt_count_config_valid(const bt_count_config& c) {
    auto in = [](uint32_t x, const uint32_t* a, size_t n) { return std::find(a, a + n, x) != a + n; };
    return in(c.wave_wedges, kWaveWedges, 4) && in(c.block_wedges, kBlockWedges, 3) && in(c.block_threads, kThreads, 2) &&
           in(c.hub_threads, kThreads, 2);
}

bt_count_config bt_count_settings(const bt_dgraph* dg) {
    bt_ctx const* c = dg->ctx;
    if (c->count_force) return *c->count_force;
    return c->tune[bt_tune_class(dg->m, dg->ngraph ? dg->ngraph : 1)];
}

extern "C" int bt_tune_class(uint64_t edges, uint32_t graphs) {
    uint64_t const per = edges / (graphs ? graphs : 1u);
    int c = 0;
    while (c + 1 < BT_TUNE_CLASSES && (per >> (2 * (c + 1))) != 0) c++;
    return c;
}

extern "C" int bt_tune_default(int size_class, bt_count_config* out) {
    if (!out || size_class < 0 || size_class >= BT_TUNE_CLASSES) return BT_ERR_INVALID;
    *out = kTuneTable[size_class];
    return BT_OK;
}

extern "C" int bt_tune_candidates(bt_count_config* out, int cap, int* n) {
    if (!n || cap < 0) return BT_ERR_INVALID;
    std::vector<bt_count_config> const v = candidates();
    *n = int(v.size());
    if (out)
        for (int i = 0; i < cap && i < int(v.size()); i++) out[i] = v[size_t(i)];
    return BT_OK;
}

extern "C" int bt_tune_get(const bt_ctx* ctx, int size_class, bt_count_config* out) {
    if (!ctx || !out || size_class < 0 || size_class >= BT_TUNE_CLASSES) return BT_ERR_INVALID;
    *out = ctx->tune[size_class];
    return BT_OK;
}

extern "C" int bt_tune_set(bt_ctx* ctx, int size_class, const bt_count_config* cfg) {
    if (!ctx || !cfg || size_class < 0 || size_class >= BT_TUNE_CLASSES || !bt_count_config_valid(*cfg))
        return BT_ERR_INVALID;
    ctx->tune[size_class] = *cfg;
    return BT_OK;
}

extern "C" int bt_autotune(bt_ctx* ctx, bt_dgraph* dg, int reps, bt_count_config* best, double* best_ms,
                           double* ms) {
    if (!ctx || !dg || dg->ctx != ctx || reps < 1) return BT_ERR_INVALID;
    BT_HIP_TRY(hipSetDevice(ctx->device));
    std::vector<bt_count_config> const cand = candidates();
    uint32_t const m = dg->m;
    hipStream_t s = ctx->stream;
    unsigned long long* flag = reinterpret_cast<unsigned long long*>(dg->ctl + 20);
    double top = 0;
    size_t top_i = 0;
    int st = BT_OK;
    for (size_t i = 0; i < cand.size() && st == BT_OK; i++) {
        ctx->count_force = &cand[i];
        double t_best = 0;
        for (int r = 0; r < reps && st == BT_OK; r++) {
            st = bt_count_impl(dg, true, nullptr, nullptr, false);
            double const t = ctx->last.count_ms;
            if (r == 0 || t < t_best) t_best = t;
        }
        if (st != BT_OK) break;
        // the first candidate's supports are the reference for the others
        // (aux is free outside the h-index rounds)
        if (m && i == 0) {
            BT_HIP_TRY(hipMemcpyAsync(dg->aux, dg->sup, sizeof(uint64_t) * m, hipMemcpyDeviceToDevice, s));
        } else if (m) {
            BT_HIP_TRY(hipMemsetAsync(flag, 0, sizeof(uint64_t), s));
            k_differs<<<grid_for(m, 256, unsigned(ctx->cu_count) * 8), 256, 0, s>>>(m, dg->sup, dg->aux, flag);
            BT_HIP_TRY(hipGetLastError());
            uint64_t differs = 0;
            BT_HIP_TRY(hipMemcpyAsync(&differs, flag, sizeof(differs), hipMemcpyDeviceToHost, s));
            BT_HIP_TRY(hipStreamSynchronize(s));
            if (differs) st = BT_ERR_INTERNAL;
        }
        if (ms) ms[i] = t_best;
        if (i == 0 || t_best < top) {
            top = t_best;
            top_i = i;
        }
    }
    ctx->count_force = nullptr;
    if (st != BT_OK) return st;
    ctx->tune[bt_tune_class(m, dg->ngraph ? dg->ngraph : 1)] = cand[top_i];
    if (best) *best = cand[top_i];
    if (best_ms) *best_ms = top;
    return BT_OK;
}
