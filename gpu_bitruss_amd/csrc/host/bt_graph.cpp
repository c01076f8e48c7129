// Host-side graph layer: CSR build with the reference's process_graph
// semantics, .bin read/write, validation and the seeded generators.
//
// Reference behaviour followed (file:line in ShunyangLi/GPU_Bitruss):
//   process_graph      src/graph/graph.cpp:74-242
//   graph_to_bin       src/graph/graph.cpp:293-318
//   load_graph_bin     src/graph/graph.cpp:335-375
//   check_graph        src/graph/graph.cpp:13-52
#include "bt_graph.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

template <typename T>
T* alloc_array(uint64_t count) {
    // malloc(0) may return NULL; always hand out a valid pointer.
    void* p = std::malloc(sizeof(T) * (count ? count : 1));
    return static_cast<T*>(p);
}

void zero_graph(bt_host_graph* g) { std::memset(g, 0, sizeof(*g)); }

int alloc_graph(bt_host_graph* g, uint32_t n, uint32_t m) {
    g->neighbors = alloc_array<uint32_t>(2ull * m);
    g->offsets = alloc_array<uint32_t>(uint64_t(n) + 1);
    g->degrees = alloc_array<uint32_t>(n);
    g->edges = alloc_array<uint32_t>(2ull * m);
    g->edge_ids = alloc_array<uint32_t>(2ull * m);
    if (!g->neighbors || !g->offsets || !g->degrees || !g->edges || !g->edge_ids) {
        bt_graph_release(g);
        return BT_ERR_NOMEM;
    }
    return BT_OK;
}

// The lower-side edge-id lookup of process_graph (graph.cpp:196-228):
// signed-int binary search with the `mid == 0` early exit.
int64_t ref_binary_search(const uint32_t* nbr, uint32_t len, uint32_t value) {
    int left = 0;
    int right = int(len) - 1;
    while (left <= right) {
        int mid = left + ((right - left) >> 1);
        if (nbr[mid] == value) return mid;
        if (nbr[mid] < value) {
            left = mid + 1;
        } else {
            if (mid == 0) break;
            right = mid - 1;
        }
    }
    return -1;
}

// ---------------------------------------------------------------- PRNG ---
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro256ss {
    uint64_t s[4];
    explicit Xoshiro256ss(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t const result = rotl(s[1] * 5, 7) * 9;
        uint64_t const t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform double in [0, 1)
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n), n > 0, by rejection (no modulo bias)
    uint64_t below(uint64_t n) {
        uint64_t const limit = (~0ull) - ((~0ull) % n);
        uint64_t x;
        do {
            x = next();
        } while (x >= limit);
        return x % n;
    }
};

// Chung-Lu expected-degree weights: w_i = min(cap, c*(i+1)^(-alpha)) with
// alpha = 1/(gamma-1), c found by bisection so that sum w_i = target.
std::vector<double> chung_lu_weights(uint32_t count, double gamma, double target, double cap) {
    std::vector<double> w(count);
    double const alpha = 1.0 / (gamma - 1.0);
    std::vector<double> base(count);
    for (uint32_t i = 0; i < count; i++) base[i] = std::pow(double(i) + 1.0, -alpha);
    auto sum_for = [&](double c) {
        double s = 0.0;
        for (uint32_t i = 0; i < count; i++) s += std::min(cap, c * base[i]);
        return s;
    };
    if (double(count) * cap <= target) {
        for (auto& x : w) x = cap;
        return w;
    }
    double lo = 0.0, hi = 1.0;
    while (sum_for(hi) < target) hi *= 2.0;
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (sum_for(mid) < target) lo = mid; else hi = mid;
    }
    for (uint32_t i = 0; i < count; i++) w[i] = std::min(cap, hi * base[i]);
    return w;
}

int fill_edge_list(const std::vector<uint32_t>& us, const std::vector<uint32_t>& vs, bt_edge_list* out) {
    out->count = us.size();
    out->upper = alloc_array<uint32_t>(us.size());
    out->lower = alloc_array<uint32_t>(vs.size());
    if (!out->upper || !out->lower) {
        bt_edge_list_release(out);
        return BT_ERR_NOMEM;
    }
    if (!us.empty()) {
        std::memcpy(out->upper, us.data(), us.size() * sizeof(uint32_t));
        std::memcpy(out->lower, vs.data(), vs.size() * sizeof(uint32_t));
    }
    return BT_OK;
}

int gen_uniform_impl(uint32_t nu, uint32_t nv, uint64_t m, Xoshiro256ss& rng,
                     std::vector<uint32_t>& us, std::vector<uint32_t>& vs) {
    uint64_t const cells = uint64_t(nu) * uint64_t(nv);
    if (m > cells) return BT_ERR_INVALID;
    us.reserve(m);
    vs.reserve(m);
    if (cells <= (1ull << 31)) {
        std::vector<uint64_t> bits((cells + 63) / 64, 0);
        while (us.size() < m) {
            uint64_t const u = rng.below(nu), v = rng.below(nv);
            uint64_t const key = u * nv + v;
            uint64_t const mask = 1ull << (key & 63);
            if (bits[key >> 6] & mask) continue;
            bits[key >> 6] |= mask;
            us.push_back(uint32_t(u));
            vs.push_back(uint32_t(v));
        }
    } else {
        std::unordered_set<uint64_t> seen;
        seen.reserve(m * 2);
        while (us.size() < m) {
            uint64_t const u = rng.below(nu), v = rng.below(nv);
            if (!seen.insert(u * nv + v).second) continue;
            us.push_back(uint32_t(u));
            vs.push_back(uint32_t(v));
        }
    }
    return BT_OK;
}

void gen_chung_lu_impl(uint32_t nu, uint32_t nv, double gamma_u, double gamma_v, double m_target,
                       double cap_u, double cap_v, Xoshiro256ss& rng,
                       std::vector<uint32_t>& us, std::vector<uint32_t>& vs) {
    std::vector<double> const wu = chung_lu_weights(nu, gamma_u, m_target, cap_u);
    std::vector<double> const wv = chung_lu_weights(nv, gamma_v, m_target, cap_v);
    double const S = m_target;
    us.reserve(size_t(m_target * 1.05) + 16);
    vs.reserve(size_t(m_target * 1.05) + 16);
    // Miller & Hagberg (2011), "Efficient generation of networks with given
    // expected degrees": both weight sequences are non-increasing, so for a
    // fixed u the edge probability is non-increasing in v and candidates are
    // reached by geometric skips with acceptance q/p.
    for (uint32_t u = 0; u < nu; u++) {
        uint64_t v = 0;
        double p = std::min(1.0, wu[u] * wv[0] / S);
        while (v < nv && p > 0.0) {
            if (p < 1.0) {
                double const r = 1.0 - rng.uniform();  // (0, 1]
                double const skip = std::floor(std::log(r) / std::log1p(-p));
                if (skip >= double(nv)) break;
                v += uint64_t(skip);
            }
            if (v >= nv) break;
            double const q = std::min(1.0, wu[u] * wv[v] / S);
            double const r2 = rng.uniform();
            if (r2 < q / p) {
                us.push_back(u);
                vs.push_back(uint32_t(v));
            }
            p = q;
            v += 1;
        }
    }
}

inline uint32_t log_uniform_u32(Xoshiro256ss& rng, double lo, double hi) {
    double const x = std::exp(std::log(lo) + rng.uniform() * (std::log(hi) - std::log(lo)));
    return uint32_t(std::llround(x));
}

}  // namespace

extern "C" {

const char* bt_status_str(int status) {
    switch (status) {
        case BT_OK: return "ok";
        case BT_ERR_INVALID: return "invalid argument or malformed graph";
        case BT_ERR_IO: return "file cannot be opened, read or written";
        case BT_ERR_NOMEM: return "host allocation failed";
        case BT_ERR_DEVICE: return "HIP runtime call failed";
        case BT_ERR_NO_DEVICE: return "no HIP device, or the device id is not available";
        case BT_ERR_BUDGET: return "request exceeds the device-memory budget";
        case BT_ERR_OVERFLOW: return "count does not fit its type";
        case BT_ERR_INTERNAL: return "internal invariant failed";
        case BT_ERR_STATE: return "call not valid in the current state";
        case BT_PARTIAL: return "partial result: msp stopped at the level bound (BT_OPT_PEEL_MAX_LEVEL)";
        case BT_ERR_BUSY: return "persistent peel blocks were not co-resident (grid barrier timed out; CUs held by other work)";
        case BT_ERR_SCRATCH: return "graph exceeds the static scratch of the allocation-free entry point";
        default: return "unknown status";
    }
}

void bt_graph_release(bt_host_graph* g) {
    if (!g) return;
    std::free(g->neighbors);
    std::free(g->offsets);
    std::free(g->degrees);
    std::free(g->edges);
    std::free(g->edge_ids);
    zero_graph(g);
}

void bt_edge_list_release(bt_edge_list* e) {
    if (!e) return;
    std::free(e->upper);
    std::free(e->lower);
    e->upper = e->lower = nullptr;
    e->count = 0;
}

int bt_graph_from_edges(const uint32_t* raw_upper, const uint32_t* raw_lower, uint64_t count,
                        bt_host_graph* out) {
    if (!out || (count && (!raw_upper || !raw_lower))) return BT_ERR_INVALID;
    zero_graph(out);
    if (count > 0x7FFFFFFFull) return BT_ERR_OVERFLOW;  // 2m must fit u32
    uint32_t max_u = 0, max_v = 0;
    for (uint64_t i = 0; i < count; i++) {
        if (raw_upper[i] >= BT_MAX_IDS || raw_lower[i] >= BT_MAX_IDS) return BT_ERR_INVALID;
        max_u = std::max(max_u, raw_upper[i]);
        max_v = std::max(max_v, raw_lower[i]);
    }
    uint32_t const NONE = 0xFFFFFFFFu;
    // first-appearance relabel (graph.cpp:92-114)
    std::vector<uint32_t> upper_ids(count ? size_t(max_u) + 1 : 0, NONE);
    std::vector<uint32_t> lower_ids(count ? size_t(max_v) + 1 : 0, NONE);
    std::vector<uint32_t> nu_of(count), nl_of(count);
    uint32_t u_num = 0, l_num = 0;
    for (uint64_t i = 0; i < count; i++) {
        uint32_t& a = upper_ids[raw_upper[i]];
        if (a == NONE) a = u_num++;
        uint32_t& b = lower_ids[raw_lower[i]];
        if (b == NONE) b = l_num++;
        nu_of[i] = a;
        nl_of[i] = b;
    }
    uint32_t const n = u_num + l_num;
    uint32_t const m = uint32_t(count);
    if (alloc_graph(out, n, m) != BT_OK) return BT_ERR_NOMEM;
    out->u_num = u_num;
    out->l_num = l_num;
    out->u_max = u_num;
    out->n = n;
    out->m = m;
    out->support_max = 0;

    // degrees and offsets (graph.cpp:157-165)
    std::fill(out->degrees, out->degrees + n, 0u);
    for (uint64_t i = 0; i < count; i++) {
        out->degrees[nu_of[i]]++;
        out->degrees[u_num + nl_of[i]]++;
    }
    out->offsets[0] = 0;
    for (uint32_t x = 0; x < n; x++) out->offsets[x + 1] = out->offsets[x] + out->degrees[x];
    // fill lists (lower ids shifted by u_num, graph.cpp:118-131), then sort
    // each list ascending (graph.cpp:134-143)
    std::vector<uint32_t> cursor(out->offsets, out->offsets + n);
    for (uint64_t i = 0; i < count; i++) {
        uint32_t const a = nu_of[i], b = u_num + nl_of[i];
        out->neighbors[cursor[a]++] = b;
        out->neighbors[cursor[b]++] = a;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t x = 0; x < int64_t(n); x++)
        std::sort(out->neighbors + out->offsets[x], out->neighbors + out->offsets[x + 1]);

    // edge ids in (upper id, neighbour) order (graph.cpp:178-194)
    uint32_t e = 0;
    for (uint32_t u = 0; u < u_num; u++) {
        for (uint32_t i = out->offsets[u]; i < out->offsets[u + 1]; i++) {
            out->edge_ids[i] = e;
            out->edges[2ull * e] = u;
            out->edges[2ull * e + 1] = out->neighbors[i];
            e++;
        }
    }
    // lower-side ids by binary search in the upper list (graph.cpp:196-228)
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t vv = u_num; vv < int64_t(n); vv++) {
        uint32_t const v = uint32_t(vv);
        for (uint32_t i = out->offsets[v]; i < out->offsets[v + 1]; i++) {
            uint32_t const u = out->neighbors[i];
            int64_t const pos = ref_binary_search(out->neighbors + out->offsets[u], out->degrees[u], v);
            if (pos < 0) {
                bad = 1;
                out->edge_ids[i] = 0;
            } else {
                out->edge_ids[i] = out->edge_ids[out->offsets[u] + uint32_t(pos)];
            }
        }
    }
    if (bad) {
        bt_graph_release(out);
        return BT_ERR_INTERNAL;
    }
    return BT_OK;
}

int bt_graph_read_text(const char* path, bt_host_graph* out) {
    if (!path || !out) return BT_ERR_INVALID;
    zero_graph(out);
    FILE* f = std::fopen(path, "rb");
    if (!f) return BT_ERR_IO;
    std::vector<uint32_t> us, vs;
    std::vector<char> line(1 << 16);
    std::string acc;
    int status = BT_OK;
    while (true) {
        acc.clear();
        bool got = false;
        while (std::fgets(line.data(), int(line.size()), f)) {
            got = true;
            acc += line.data();
            if (!acc.empty() && acc.back() == '\n') break;
        }
        if (!got) break;
        if (!acc.empty() && acc[0] == '%') continue;  // graph.cpp:93
        const char* p = acc.c_str();
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
        if (*p == '\0') continue;  // blank line
        char* end = nullptr;
        long long a = std::strtoll(p, &end, 10);
        if (end == p) { status = BT_ERR_INVALID; break; }
        p = end;
        long long b = std::strtoll(p, &end, 10);
        if (end == p) { status = BT_ERR_INVALID; break; }
        if (a < 0 || b < 0 || a >= (long long)BT_MAX_IDS || b >= (long long)BT_MAX_IDS) {
            status = BT_ERR_INVALID;
            break;
        }
        us.push_back(uint32_t(a));
        vs.push_back(uint32_t(b));
    }
    std::fclose(f);
    if (status != BT_OK) return status;
    return bt_graph_from_edges(us.data(), vs.data(), us.size(), out);
}

int bt_graph_write_bin(const char* path, const bt_host_graph* g) {
    if (!path || !g) return BT_ERR_INVALID;
    FILE* f = std::fopen(path, "wb");
    if (!f) return BT_ERR_IO;
    uint32_t header[5] = {g->u_num, g->l_num, g->u_max, g->m, g->support_max};
    bool ok = std::fwrite(header, sizeof(uint32_t), 5, f) == 5;
    auto put = [&](const uint32_t* a, uint64_t cnt) {
        if (!ok || cnt == 0) return;
        ok = std::fwrite(a, sizeof(uint32_t), cnt, f) == cnt;
    };
    uint64_t const m = g->m, n = uint64_t(g->u_num) + g->l_num;
    std::vector<uint32_t> zeros(m, 0);
    put(g->neighbors, 2 * m);
    put(g->offsets, n + 1);
    put(g->degrees, n);
    put(zeros.data(), m);  // edge_support, always zero when written (graph.cpp:234, 312)
    put(g->edges, 2 * m);
    put(g->edge_ids, 2 * m);
    put(zeros.data(), m);  // bitruss, always zero when written (graph.cpp:233, 315)
    ok = (std::fclose(f) == 0) && ok;
    return ok ? BT_OK : BT_ERR_IO;
}

int bt_graph_read_bin(const char* path, bt_host_graph* out) {
    if (!path || !out) return BT_ERR_INVALID;
    zero_graph(out);
    FILE* f = std::fopen(path, "rb");
    if (!f) return BT_ERR_IO;
    uint32_t header[5];
    if (std::fread(header, sizeof(uint32_t), 5, f) != 5) {
        std::fclose(f);
        return BT_ERR_INVALID;
    }
    uint64_t const u_num = header[0], l_num = header[1], m = header[3];
    uint64_t const n = u_num + l_num;
    if (n > 0xFFFFFFFEull || 2 * m > 0xFFFFFFFFull) {
        std::fclose(f);
        return BT_ERR_INVALID;
    }
    // exact size check: 20 + 4*(8m + 2n + 1) bytes
    std::fseek(f, 0, SEEK_END);
    long long const size = std::ftell(f);
    std::fseek(f, 20, SEEK_SET);
    if (size != (long long)(20 + 4 * (8 * m + 2 * n + 1))) {
        std::fclose(f);
        return BT_ERR_INVALID;
    }
    if (alloc_graph(out, uint32_t(n), uint32_t(m)) != BT_OK) {
        std::fclose(f);
        return BT_ERR_NOMEM;
    }
    out->u_num = uint32_t(u_num);
    out->l_num = uint32_t(l_num);
    out->u_max = header[2];
    out->m = uint32_t(m);
    out->support_max = header[4];
    out->n = uint32_t(n);
    bool ok = true;
    auto get = [&](uint32_t* a, uint64_t cnt) {
        if (!ok || cnt == 0) return;
        ok = std::fread(a, sizeof(uint32_t), cnt, f) == cnt;
    };
    auto skip = [&](uint64_t cnt) {
        if (!ok) return;
        ok = std::fseek(f, long(cnt * sizeof(uint32_t)), SEEK_CUR) == 0;
    };
    get(out->neighbors, 2 * m);
    get(out->offsets, n + 1);
    get(out->degrees, n);
    skip(m);  // edge_support
    get(out->edges, 2 * m);
    get(out->edge_ids, 2 * m);
    skip(m);  // bitruss
    std::fclose(f);
    if (!ok) {
        bt_graph_release(out);
        return BT_ERR_IO;
    }
    int const st = bt_graph_check(out);
    if (st != BT_OK) {
        bt_graph_release(out);
        return st;
    }
    return BT_OK;
}

int bt_graph_check(const bt_host_graph* g) {
    if (!g) return BT_ERR_INVALID;
    uint64_t const n = uint64_t(g->u_num) + g->l_num, m = g->m;
    if (g->n != n) return BT_ERR_INVALID;
    if (n && (!g->offsets || !g->degrees)) return BT_ERR_INVALID;
    if (m && (!g->neighbors || !g->edges || !g->edge_ids)) return BT_ERR_INVALID;
    if (g->offsets && g->offsets[0] != 0) return BT_ERR_INVALID;
    for (uint64_t x = 0; x < n; x++) {
        if (g->offsets[x + 1] < g->offsets[x]) return BT_ERR_INVALID;
        if (g->offsets[x + 1] - g->offsets[x] != g->degrees[x]) return BT_ERR_INVALID;
    }
    if (n && g->offsets[n] != 2 * m) return BT_ERR_INVALID;
    if (!n && m) return BT_ERR_INVALID;
    // each upper list holds lower ids and vice versa; sorted; edge ids valid
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t xx = 0; xx < int64_t(n); xx++) {
        uint32_t const x = uint32_t(xx);
        bool const upper = x < g->u_num;
        for (uint32_t i = g->offsets[x]; i < g->offsets[x + 1]; i++) {
            uint32_t const y = g->neighbors[i];
            if (y >= n || (upper ? (y < g->u_num) : (y >= g->u_num))) { bad = 1; break; }
            if (i > g->offsets[x] && g->neighbors[i - 1] > y) { bad = 1; break; }
            uint32_t const e = g->edge_ids[i];
            if (e >= m) { bad = 1; break; }
            uint32_t const eu = g->edges[2ull * e], ev = g->edges[2ull * e + 1];
            // check_graph: the id seen from x names the pair (x, y)
            if (upper ? (eu != x || ev != y) : (eu != y || ev != x)) { bad = 1; break; }
        }
    }
    return bad ? BT_ERR_INVALID : BT_OK;
}

int bt_graph_wedges(const bt_host_graph* g, uint64_t* out) {
    if (!g || !out) return BT_ERR_INVALID;
    // W = #{(u, v, w): v in N(u), w in N(v), v < u, w < u} in vertex priority
    // (degree, id) (cpu_bfc.cpp:72): for each v and each neighbour u above
    // v, the neighbours of v below u
    auto below = [&](uint32_t x, uint32_t y) {
        return g->degrees[x] < g->degrees[y] || (g->degrees[x] == g->degrees[y] && x < y);
    };
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        std::vector<uint32_t> nb;
#pragma omp for schedule(dynamic, 256)
        for (int64_t vv = 0; vv < int64_t(g->n); vv++) {
            uint32_t const v = uint32_t(vv);
            nb.assign(g->neighbors + g->offsets[v], g->neighbors + g->offsets[v + 1]);
            std::sort(nb.begin(), nb.end(), below);
            for (size_t i = 0; i < nb.size(); i++)
                if (below(v, nb[i])) total += i;  // nb[0..i) are below nb[i]
        }
    }
    *out = total;
    return BT_OK;
}

int bt_graph_union(const bt_host_graph* gs, uint64_t k, bt_host_graph* u, uint64_t* e_off) {
    if (!u || !e_off || (k && !gs)) return BT_ERR_INVALID;
    std::memset(u, 0, sizeof(*u));
    std::vector<uint64_t> uo(k + 1, 0), lo(k + 1, 0);
    e_off[0] = 0;
    for (uint64_t i = 0; i < k; i++) {
        int const st = bt_graph_check(&gs[i]);
        if (st != BT_OK) return st;
        uo[i + 1] = uo[i] + gs[i].u_num;
        lo[i + 1] = lo[i] + gs[i].l_num;
        e_off[i + 1] = e_off[i] + gs[i].m;
    }
    uint64_t const U = uo[k], L = lo[k], m = e_off[k], n = U + L;
    if (2 * m >= (1ull << 32) || n >= (1ull << 32) - 1) return BT_ERR_OVERFLOW;
    u->u_num = uint32_t(U);
    u->l_num = uint32_t(L);
    u->u_max = U ? uint32_t(U - 1) : 0;
    u->m = uint32_t(m);
    u->n = uint32_t(n);
    u->neighbors = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<uint64_t>(2 * m, 1)));
    u->edge_ids = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<uint64_t>(2 * m, 1)));
    u->edges = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<uint64_t>(2 * m, 1)));
    u->offsets = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (n + 1)));
    u->degrees = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * std::max<uint64_t>(n, 1)));
    if (!u->neighbors || !u->edge_ids || !u->edges || !u->offsets || !u->degrees) {
        bt_graph_release(u);
        return BT_ERR_NOMEM;
    }
    // union id of graph i's vertex x: uppers of all graphs, then lowers
    auto uid = [&](uint64_t i, uint32_t x) -> uint32_t {
        return x < gs[i].u_num ? uint32_t(uo[i] + x) : uint32_t(U + lo[i] + (x - gs[i].u_num));
    };
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < int64_t(k); ii++) {
        uint64_t const i = uint64_t(ii);
        for (uint32_t x = 0; x < gs[i].n; x++) u->degrees[uid(i, x)] = gs[i].degrees[x];
    }
    u->offsets[0] = 0;
    for (uint64_t x = 0; x < n; x++) u->offsets[x + 1] = u->offsets[x] + u->degrees[x];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < int64_t(k); ii++) {
        uint64_t const i = uint64_t(ii);
        bt_host_graph const& g = gs[i];
        uint32_t const eb = uint32_t(e_off[i]);
        for (uint32_t x = 0; x < g.n; x++) {
            uint32_t const dst = u->offsets[uid(i, x)];
            for (uint32_t j = 0; j < g.degrees[x]; j++) {
                u->neighbors[dst + j] = uid(i, g.neighbors[g.offsets[x] + j]);
                u->edge_ids[dst + j] = eb + g.edge_ids[g.offsets[x] + j];
            }
        }
        for (uint32_t e = 0; e < g.m; e++) {
            u->edges[2ull * (eb + e)] = uid(i, g.edges[2ull * e]);
            u->edges[2ull * (eb + e) + 1] = uid(i, g.edges[2ull * e + 1]);
        }
    }
    return BT_OK;
}

int bt_gen_uniform(uint32_t nu, uint32_t nv, uint64_t m, uint64_t seed, bt_edge_list* out) {
    if (!out) return BT_ERR_INVALID;
    out->count = 0;
    out->upper = out->lower = nullptr;
    if ((nu == 0 || nv == 0) && m) return BT_ERR_INVALID;
    Xoshiro256ss rng(seed);
    std::vector<uint32_t> us, vs;
    int const st = gen_uniform_impl(nu, nv, m, rng, us, vs);
    if (st != BT_OK) return st;
    return fill_edge_list(us, vs, out);
}

int bt_gen_chung_lu(uint32_t nu, uint32_t nv, double gamma_u, double gamma_v, double m_target,
                    double cap_u, double cap_v, uint64_t seed, bt_edge_list* out) {
    if (!out) return BT_ERR_INVALID;
    out->count = 0;
    out->upper = out->lower = nullptr;
    if (nu == 0 || nv == 0 || !(gamma_u > 1.0) || !(gamma_v > 1.0) || !(m_target > 0.0) ||
        !(cap_u > 0.0) || !(cap_v > 0.0))
        return BT_ERR_INVALID;
    Xoshiro256ss rng(seed);
    std::vector<uint32_t> us, vs;
    gen_chung_lu_impl(nu, nv, gamma_u, gamma_v, m_target, cap_u, cap_v, rng, us, vs);
    return fill_edge_list(us, vs, out);
}

int bt_gen_workload(int workload, uint64_t index, bt_edge_list* out) {
    if (!out) return BT_ERR_INVALID;
    switch (workload) {
        case BT_WL_SMALL_UNIFORM:
            return bt_gen_uniform(1000, 1500, 10000, 0, out);
        case BT_WL_POWERLAW_2M:
            return bt_gen_chung_lu(200000, 100000, 2.3, 2.3, 2.0e6, 5000.0, 5000.0, 1, out);
        case BT_WL_SKEWED_60M:
            return bt_gen_chung_lu(4000000, 1000000, 2.5, 2.0, 6.0e7, 2.0e5, 2.0e5, 2, out);
        case BT_WL_MIX_GRAPH: {
            if (index >= 4096) return BT_ERR_INVALID;
            uint64_t const seed = 1000 + index;
            Xoshiro256ss rng(seed);
            uint32_t const nu = log_uniform_u32(rng, 50.0, 5000.0);
            uint32_t const nv = log_uniform_u32(rng, 50.0, 5000.0);
            double const dens = std::exp(std::log(0.005) + rng.uniform() * (std::log(0.05) - std::log(0.005)));
            uint64_t m = uint64_t(std::llround(dens * double(nu) * double(nv)));
            if (m == 0) m = 1;
            std::vector<uint32_t> us, vs;
            if (seed % 2 == 0) {
                int const st = gen_uniform_impl(nu, nv, m, rng, us, vs);
                if (st != BT_OK) return st;
            } else {
                gen_chung_lu_impl(nu, nv, 2.2, 2.2, double(m), double(nv), double(nu), rng, us, vs);
            }
            return fill_edge_list(us, vs, out);
        }
        case BT_WL_LARGE_REQUEST: {
            Xoshiro256ss rng(0xB17255ull ^ (index * 0x9E3779B97F4A7C15ull));
            uint32_t const nu = uint32_t(100000 + rng.below(400001));
            uint32_t const nv = uint32_t(50000 + rng.below(150001));
            double const m = double(1000000 + rng.below(4000001));
            std::vector<uint32_t> us, vs;
            gen_chung_lu_impl(nu, nv, 2.2, 2.2, m, double(nv), double(nu), rng, us, vs);
            return fill_edge_list(us, vs, out);
        }
        default:
            return BT_ERR_INVALID;
    }
}

}  // extern "C"
