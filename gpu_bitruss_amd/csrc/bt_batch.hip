// Batches: many independent graphs decomposed as one disjoint union.
//
// The union puts the upper vertices of every graph first, then the lower
// vertices, each graph's block of ids in input order; edge ids are offset
// per graph. Inside one graph every comparison the algorithms make (vertex
// priority (degree, id), edge ids in the attribution rule, the upper/lower
// side of an edge) keeps its outcome, and butterflies never cross graphs, so
// every per-edge result of the union restricted to graph i is graph i's
// plain result. The per-graph scalars are reductions over the graph's edge
// range (Σ supports / 4 butterflies, max support, max bitruss number) or,
// for hidx, the round after the graph's last change, recorded by the round
// kernels through dg->egraph / dg->graph_last.
//
// One union means one launch sequence for all its graphs: the counting
// buckets the roots of all graphs together by wedge count, and msp runs
// every graph's sub-rounds inside the same persistent launch, so the
// latency of a phase is paid once for all of them. The graphs are spread
// over lanes (contexts with their own stream and a share of the CUs) by
// wedge count, largest first onto the lightest lane, and the lanes run side
// by side: a phase of a union is short and bound by its slowest wave, so
// several smaller unions on parts of the device finish sooner than one
// large union on all of it. A lane whose bloom records would not fit one
// index (32-bit record indices, the memory budget) holds several unions,
// run one after the other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/bitruss_hip.h"
#include "bt_common.h"
#include "bt_internal.h"

using namespace bt;

// One disjoint union of graphs gid[0..] (batch indices, ascending) of a batch.
struct bt_chunk {
    bt_dgraph* dg = nullptr;
    std::vector<size_t> gid;
    int lane = 0;
    std::vector<uint64_t> e_off;     // gid.size() + 1 edge offsets inside the union
    uint64_t* d_eoff = nullptr;      // device copy
    uint32_t* egraph = nullptr;      // device, m: graph of each edge (chunk-local)
    uint32_t* glast = nullptr;       // device, per graph
    uint64_t* red = nullptr;         // device, per graph: sum, max
    std::vector<std::pair<void*, size_t>> blocks;  // device blocks of the arrays above
};

struct bt_batch {
    bt_ctx* ctx = nullptr;
    size_t k = 0;
    int nlanes = 1;
    std::vector<bt_chunk> chunks;
    std::vector<bt_ctx*> lane_ctx;               // lanes beyond the caller's context (created on first run)
    std::vector<std::vector<uint64_t>> h_out;    // per lane: host staging of a union result
};

namespace {

constexpr int kRB = 256;

__global__ void k_fill_egraph(uint32_t k, const uint64_t* __restrict__ eoff, uint32_t* __restrict__ egraph) {
    for (uint32_t g = blockIdx.x; g < k; g += gridDim.x)
        for (uint64_t e = eoff[g] + threadIdx.x; e < eoff[g + 1]; e += blockDim.x) egraph[e] = g;
}

// per graph: sum and max of a[e] over its edge range (one block per graph)
__global__ __launch_bounds__(kRB) void k_seg_reduce(uint32_t k, const uint64_t* __restrict__ eoff,
                                                    const uint64_t* __restrict__ a, uint64_t* __restrict__ red) {
    __shared__ uint64_t s_sum[kRB / kWave], s_max[kRB / kWave];
    for (uint32_t g = blockIdx.x; g < k; g += gridDim.x) {
        uint64_t sum = 0, mx = 0;
        for (uint64_t e = eoff[g] + threadIdx.x; e < eoff[g + 1]; e += blockDim.x) {
            uint64_t const v = a[e];
            sum += v;
            mx = v > mx ? v : mx;
        }
        sum = wave_sum_u64(sum);
        mx = wave_max_u64(mx);
        int const lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
        if (lane == 0) {
            s_sum[w] = sum;
            s_max[w] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t ts = 0, tm = 0;
            for (int i = 0; i < kRB / kWave; i++) {
                ts += s_sum[i];
                tm = s_max[i] > tm ? s_max[i] : tm;
            }
            red[2 * g] = ts;
            red[2 * g + 1] = tm;
        }
        __syncthreads();
    }
}

void chunk_free(bt_chunk& c, bt_ctx* ctx) {
    if (c.dg) bt_graph_free(c.dg);  // synchronises the stream
    for (auto const& b : c.blocks) bt_ctx_release(ctx, b.first, b.second);
    c = bt_chunk{};
}

template <typename T>
int chunk_alloc(bt_ctx* ctx, bt_chunk& c, T** p, size_t count) {
    void* q = nullptr;
    size_t got = 0;
    BT_TRY(bt_ctx_alloc(ctx, sizeof(T) * std::max<size_t>(count, 1), &q, &got));
    c.blocks.emplace_back(q, got);
    *p = static_cast<T*>(q);
    return BT_OK;
}

int chunk_upload(bt_ctx* ctx, const bt_host_graph* graphs, bt_chunk& c) {
    size_t const k = c.gid.size();
    std::vector<bt_host_graph> parts(k);
    for (size_t i = 0; i < k; i++) parts[i] = graphs[c.gid[i]];
    bt_host_graph u{};
    c.e_off.assign(k + 1, 0);
    int st = bt_graph_union(parts.data(), k, &u, c.e_off.data());
    if (st == BT_OK) st = bt_graph_upload(ctx, &u, &c.dg);
    bt_graph_release(&u);
    if (st != BT_OK) return st;
    size_t const kk = std::max<size_t>(k, 1), m = c.dg->m;
    hipStream_t s = ctx->stream;
    BT_TRY(chunk_alloc(ctx, c, &c.d_eoff, kk + 1));
    BT_TRY(chunk_alloc(ctx, c, &c.egraph, m));
    BT_TRY(chunk_alloc(ctx, c, &c.glast, kk));
    BT_TRY(chunk_alloc(ctx, c, &c.red, 2 * kk));
    BT_HIP_TRY(hipMemcpyAsync(c.d_eoff, c.e_off.data(), sizeof(uint64_t) * (k + 1), hipMemcpyHostToDevice, s));
    if (k) k_fill_egraph<<<unsigned(std::min<size_t>(k, 65535)), 256, 0, s>>>(uint32_t(k), c.d_eoff, c.egraph);
    BT_HIP_TRY(hipGetLastError());
    BT_HIP_TRY(hipStreamSynchronize(s));
    c.dg->egraph = c.egraph;
    c.dg->graph_last = c.glast;
    c.dg->ngraph = uint32_t(k);
    return BT_OK;
}

// Runs algo on one chunk and scatters its per-graph results.
int chunk_run(bt_ctx* ctx, bt_chunk& c, int algo, uint64_t* const* outs, bt_batch_result* results,
              std::vector<uint64_t>& h_out) {
    bt_dgraph* dg = c.dg;
    hipStream_t s = ctx->stream;
    size_t const k = c.gid.size();
    uint32_t rounds = 0;
    const uint64_t* dres = nullptr;
    if (algo == BT_ALGO_BFC || algo == BT_ALGO_EBFC) {
        // per-graph totals need per-edge supports: sum / 4 per graph
        BT_TRY(bt_count_impl(dg, true, nullptr, nullptr));
        dres = dg->sup;
    } else if (algo == BT_ALGO_HIDX) {
        if (k) BT_HIP_TRY(hipMemsetAsync(c.glast, 0, sizeof(uint32_t) * k, s));
        BT_TRY(bt_hidx_impl(dg, &rounds));
        dres = dg->res;
    } else {
        BT_TRY(bt_msp_impl(dg, nullptr));
        dres = dg->res;
    }
    std::vector<uint64_t> red;
    std::vector<uint32_t> glast;
    if (results && k) {
        k_seg_reduce<<<unsigned(std::min<size_t>(k, 65535)), kRB, 0, s>>>(uint32_t(k), c.d_eoff, dres, c.red);
        BT_HIP_TRY(hipGetLastError());
        red.resize(2 * k);
        BT_HIP_TRY(hipMemcpyAsync(red.data(), c.red, sizeof(uint64_t) * 2 * k, hipMemcpyDeviceToHost, s));
        if (algo == BT_ALGO_HIDX) {
            glast.resize(k);
            BT_HIP_TRY(hipMemcpyAsync(glast.data(), c.glast, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s));
        }
    }
    bool const want_out = outs && dg->m && algo != BT_ALGO_BFC;
    if (want_out) {
        h_out.resize(dg->m);
        BT_HIP_TRY(hipMemcpyAsync(h_out.data(), dres, sizeof(uint64_t) * dg->m, hipMemcpyDeviceToHost, s));
    }
    BT_HIP_TRY(hipStreamSynchronize(s));
    if (want_out) {
        for (size_t i = 0; i < k; i++)
            if (outs[c.gid[i]] && c.e_off[i + 1] > c.e_off[i])
                std::memcpy(outs[c.gid[i]], h_out.data() + c.e_off[i], sizeof(uint64_t) * (c.e_off[i + 1] - c.e_off[i]));
    }
    if (results) {
        for (size_t i = 0; i < k; i++) {
            bt_batch_result& r = results[c.gid[i]];
            std::memset(&r, 0, sizeof(r));
            uint64_t const sum = red[2 * i], mx = red[2 * i + 1];
            if (algo == BT_ALGO_BFC || algo == BT_ALGO_EBFC) {
                r.total = sum / 4;  // every butterfly adds 1 to each of its 4 edges
                r.support_max = mx;
            } else if (algo == BT_ALGO_MSP) {
                if (mx > 0xFFFFFFFFull) return BT_ERR_OVERFLOW;
                r.levels = uint32_t(mx);
            } else {
                // the plain run stops after the first round without a change
                r.rounds = glast[i] + 1;
            }
        }
    }
    return BT_OK;
}

}  // namespace

void bt_batch_detach(bt_batch* b) {
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    for (bt_chunk& c : b->chunks) chunk_free(c, b->ctx);
    b->chunks.clear();
    for (bt_ctx* lc : b->lane_ctx) bt_ctx_destroy(lc);
    b->lane_ctx.clear();
    b->ctx = nullptr;
}

extern "C" {

int bt_batch_free(bt_batch* b) {
    if (!b) return BT_ERR_INVALID;
    if (b->ctx) {
        std::lock_guard<std::mutex> lk(b->ctx->live_mu);
        b->ctx->live_batches.erase(b);
    }
    bt_batch_detach(b);
    delete b;
    return BT_OK;
}

int bt_batch_upload(bt_ctx* ctx, const bt_host_graph* graphs, size_t k, bt_batch** out) {
    if (!out) return BT_ERR_INVALID;
    *out = nullptr;
    if (!ctx || (k && !graphs) || k >= (1ull << 31)) return BT_ERR_INVALID;
    BT_HIP_TRY(hipSetDevice(ctx->device));
    for (size_t i = 0; i < k; i++) BT_TRY(bt_graph_check(&graphs[i]));
    bt_batch* b = new (std::nothrow) bt_batch();
    if (!b) return BT_ERR_NOMEM;
    b->ctx = ctx;
    b->k = k;
    // Lanes: graphs by priority wedges (which bound the bloom records),
    // largest first onto the lane with the fewest so far.
    std::vector<uint64_t> wedges(k, 0);
    for (size_t i = 0; i < k; i++) BT_TRY(bt_graph_wedges(&graphs[i], &wedges[i]));
    int const nl = int(std::max<size_t>(1, std::min<size_t>(size_t(std::max(ctx->batch_lanes, 1)), k)));
    b->nlanes = nl;
    std::vector<std::vector<size_t>> lane_g(nl);
    {
        std::vector<size_t> order(k);
        std::iota(order.begin(), order.end(), size_t(0));
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return wedges[x] > wedges[y]; });
        std::vector<uint64_t> load(nl, 0);
        for (size_t i : order) {
            int const l = int(std::min_element(load.begin(), load.end()) - load.begin());
            lane_g[l].push_back(i);
            load[l] += wedges[i] + graphs[i].m;
        }
        for (auto& v : lane_g) std::sort(v.begin(), v.end());
    }
    // A lane's graphs share a union while their wedges stay within what one
    // bloom index may hold: record indices are 32-bit, and the indexes of
    // the lanes' unions are built side by side, at most 52 bytes per wedge
    // (16 B emitted, 36 B sorted and indexed per record) within 85% of the
    // free device memory after the unions' graphs (about 140 B per edge).
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete b;
        return BT_ERR_DEVICE;
    }
    uint64_t m_all = 0;
    for (size_t i = 0; i < k; i++) m_all += graphs[i].m;
    double const room = std::max(0.0, double(free_b) * 0.85 - 140.0 * double(m_all));
    uint64_t const by_mem = uint64_t(room / 52.0 / nl);
    uint64_t limit = std::max<uint64_t>(1, std::min<uint64_t>((1ull << 31) - 1, by_mem));
    if (ctx->batch_wedges) limit = std::min(limit, ctx->batch_wedges);
    for (int l = 0; l < nl; l++) {
        std::vector<size_t> const& gl = lane_g[l];
        uint64_t acc = 0, acc_m = 0;
        size_t i0 = 0;
        for (size_t i = 0; i <= gl.size(); i++) {
            uint64_t const w = i < gl.size() ? wedges[gl[i]] : 0;
            uint64_t const gm = i < gl.size() ? graphs[gl[i]].m : 0;
            bool const close = i == gl.size() || (i > i0 && (acc + w > limit || 2 * (acc_m + gm) >= (1ull << 32)));
            if (close && i > i0) {
                bt_chunk c;
                c.gid.assign(gl.begin() + i0, gl.begin() + i);
                c.lane = l;
                b->chunks.push_back(std::move(c));
                i0 = i;
                acc = acc_m = 0;
            }
            acc += w;
            acc_m += gm;
        }
    }
    for (bt_chunk& c : b->chunks) {
        int const st = chunk_upload(ctx, graphs, c);
        if (st != BT_OK) {
            bt_batch_free(b);
            return st;
        }
    }
    {
        std::lock_guard<std::mutex> lk(ctx->live_mu);
        ctx->live_batches.insert(b);
    }
    *out = b;
    return BT_OK;
}

int bt_batch_run(bt_ctx* ctx, bt_batch* b, int algo, uint64_t* const* outs, bt_batch_result* results) {
    if (!ctx || !b || b->ctx != ctx) return BT_ERR_INVALID;
    if (algo < BT_ALGO_BFC || algo > BT_ALGO_MSP) return BT_ERR_INVALID;
    BT_HIP_TRY(hipSetDevice(ctx->device));
    int const nl = b->nlanes;
    b->h_out.resize(nl);
    if (nl == 1) {
        for (bt_chunk& c : b->chunks) BT_TRY(chunk_run(ctx, c, algo, outs, results, b->h_out[0]));
        return BT_OK;
    }
    // lane contexts: each persistent peel takes its share of the CUs
    int const cus = ctx->peel_blocks > 0 && ctx->peel_blocks < ctx->cu_count ? ctx->peel_blocks : ctx->cu_count;
    // shares are whole multiples of the 8 XCDs: the blocks of a launch go
    // round-robin over the XCDs, so a multiple of 8 takes the same number
    // of CUs on each, and shares that add up to at most the CU count fit
    // every XCD together
    int const share = std::max(kXcds, cus / nl / kXcds * kXcds);
    if (b->lane_ctx.empty()) {
        for (int l = 0; l < nl; l++) {
            bt_ctx* lc = nullptr;
            int const st = bt_ctx_create(ctx->device, &lc);
            if (st != BT_OK) return st;
            b->lane_ctx.push_back(lc);
        }
    }
    for (bt_ctx* lc : b->lane_ctx) {
        lc->msp_path = ctx->msp_path;
        lc->index_mem_frac = ctx->index_mem_frac;
        lc->batch_levels = ctx->batch_levels;
        lc->peel_slice_ms = ctx->peel_slice_ms;
        lc->peel_c2cap = ctx->peel_c2cap;
        lc->peel_window_edges = ctx->peel_window_edges;
        lc->peel_big_middles = ctx->peel_big_middles;
        lc->hidx_graph_rounds = ctx->hidx_graph_rounds;
        lc->hidx_sorted = ctx->hidx_sorted;
        std::copy(ctx->tune, ctx->tune + BT_TUNE_CLASSES, lc->tune);
        lc->peel_blocks = share < lc->cu_count ? share : 0;
    }
    std::vector<int> status(nl, BT_OK);
    auto run_lane = [&](int l) {
        bt_ctx* lc = b->lane_ctx[l];
        if (hipSetDevice(lc->device) != hipSuccess) {
            status[l] = BT_ERR_DEVICE;
            return;
        }
        for (bt_chunk& c : b->chunks) {
            if (c.lane != l) continue;
            c.dg->ctx = lc;  // the union runs on the lane's stream and index
            int const st = chunk_run(lc, c, algo, outs, results, b->h_out[l]);
            c.dg->ctx = ctx;
            if (st != BT_OK) {
                status[l] = st;
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < nl; l++) th.emplace_back(run_lane, l);
    run_lane(0);
    for (auto& t : th) t.join();
    for (int st : status) BT_TRY(st);
    return BT_OK;
}

int bt_batch_chunks(const bt_batch* b, size_t* chunks) {
    if (!b || !chunks) return BT_ERR_INVALID;
    *chunks = b->chunks.size();
    return BT_OK;
}

}  // extern "C"
