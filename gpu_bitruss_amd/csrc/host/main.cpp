// coh — command-line host with the reference's flags (main.cu:27-133,
// src/util/utility.h:75-114), calling libbitruss_hip.so through its C-ABI.
//
//   coh --graph X.txt --bin X.bin          convert a text edge list
//   coh --gpu --algo bfc|ebfc|hidx|msp --graph X.bin [--device N]
//   coh --device_info [--device N]
//
// Additions beyond the reference flags: --dump F writes the per-edge uint64
// result (supports for ebfc, bitruss numbers for hidx/msp) as raw
// little-endian uint64; --json prints one JSON summary line.
// `--cpu` is not served: the CPU algorithms of the reference are restated
// only as a test oracle outside this library.
#include <sys/stat.h>

#include <clocale>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "bitruss_hip.h"

namespace {

const char* kVersion = "1.3.1";


// Log line in the reference's layout (log.cpp:101-116):
// "<date> <time> LEVEL file:line: message".
void logf(const char* level, int line, const char* fmt, ...) {
    char tbuf[64];
    std::time_t t = std::time(nullptr);
    std::tm lt{};
    localtime_r(&t, &lt);
    tbuf[std::strftime(tbuf, sizeof(tbuf), "%Y-%m-%d %H:%M:%S", &lt)] = '\0';
    std::fprintf(stdout, "%s %-5s main.cpp:%d: ", tbuf, level, line);
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stdout, fmt, ap);
    va_end(ap);
    std::fputc('\n', stdout);
    std::fflush(stdout);
}
#define LOG_INFO(...) logf("INFO", __LINE__, __VA_ARGS__)
#define LOG_WARN(...) logf("WARN", __LINE__, __VA_ARGS__)
#define LOG_ERROR(...) logf("ERROR", __LINE__, __VA_ARGS__)

void usage() {
    std::printf(
        "Usage: bitruss [--help] [--version] [--device VAR] [--device_info] [--graph VAR] [--bin VAR] "
        "[--cpu] [--gpu] [--algo VAR] [--threads VAR] [--dump VAR] [--json]\n\n"
        "Optional arguments:\n"
        "  -h, --help     shows help message and exits\n"
        "  -v, --version  prints version information and exits\n"
        "  --device       GPU Device ID (must be a positive integer) [nargs=0..1] [default: 0]\n"
        "  --device_info  Display GPU device properties\n"
        "  --graph        Graph file path [nargs=0..1] [default: \"/\"]\n"
        "  --bin          Output binary file path\n"
        "  --cpu          Run CPU algorithms (not served by this build)\n"
        "  --gpu          Run GPU algorithms\n"
        "  --algo         Algorithm to run including: bfc, ebfc, hidx, msp (also mspp, peel) [nargs=0..1] "
        "[default: \"msp\"]\n"
        "  --threads      Number of threads (must be a positive integer) [nargs=0..1] [default: 1]\n"
        "  --dump         Write the per-edge uint64 result to this file\n"
        "  --json         Print a JSON summary line\n");
}

struct Args {
    bool device_used = false, bin_used = false, graph_used = false, threads_used = false;
    int device = 0, threads = 1;
    bool device_info = false, cpu = false, gpu = false, json = false;
    std::string graph = "/", bin, algo = "msp", dump;
};

bool parse_int(const char* s, int* out) {
    char* end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') return false;
    *out = int(v);
    return true;
}

int parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; i++) {
        std::string const k = argv[i];
        auto need = [&](std::string& dst) {
            if (i + 1 >= argc) return false;
            dst = argv[++i];
            return true;
        };
        if (k == "-h" || k == "--help") { usage(); std::exit(0); }
        if (k == "-v" || k == "--version") { std::printf("%s\n", kVersion); std::exit(0); }
        std::string v;
        if (k == "--device") {
            if (!need(v) || !parse_int(v.c_str(), &a.device)) return 1;
            a.device_used = true;
        } else if (k == "--threads") {
            if (!need(v) || !parse_int(v.c_str(), &a.threads)) return 1;
            a.threads_used = true;
        } else if (k == "--graph") {
            if (!need(a.graph)) return 1;
            a.graph_used = true;
        } else if (k == "--bin") {
            if (!need(a.bin)) return 1;
            a.bin_used = true;
        } else if (k == "--algo") {
            if (!need(a.algo)) return 1;
        } else if (k == "--dump") {
            if (!need(a.dump)) return 1;
        } else if (k == "--device_info") {
            a.device_info = true;
        } else if (k == "--cpu") {
            a.cpu = true;
        } else if (k == "--gpu") {
            a.gpu = true;
        } else if (k == "--json") {
            a.json = true;
        } else {
            std::fprintf(stdout, "Unknown argument: %s\n", k.c_str());
            return 1;
        }
    }
    return 0;
}

void device_info(int id) {
    bt_device_props p{};
    int const st = bt_device_info(id, &p);
    if (st != BT_OK) {
        LOG_ERROR("error: %s", bt_status_str(st));
        std::exit(EXIT_FAILURE);
    }
    double const peak = 2.0 * p.memory_clock_khz * (double(p.memory_bus_width_bits) / 8) / 1.0e6;
    std::printf("%-32s | %s\n", "Property", "Info");
    std::printf("%-32s | %d\n", "Device Number", id);
    std::printf("%-32s | %s\n", "Device name", p.name);
    std::printf("%-32s | %s\n", "Architecture", p.arch);
    std::printf("%-32s | %d\n", "Memory Bus Width (bits)", p.memory_bus_width_bits);
    std::printf("%-32s | %f\n", "Peak Memory Bandwidth (GB/s)", peak);
    std::printf("%-32s | %llu\n", "Total global memory (bytes)", (unsigned long long)p.total_global_mem);
    std::printf("%-32s | %f\n", "Total global memory (GB)", double(p.total_global_mem) / (1024.0 * 1024.0 * 1024.0));
    std::printf("%-32s | %d\n", "Number of CUs", p.multiprocessor_count);
    std::printf("%-32s | %llu\n", "Shared memory per block (bytes)", (unsigned long long)p.shared_mem_per_block);
    std::printf("%-32s | %d\n", "Max threads per CU", p.max_threads_per_multiprocessor);
    std::printf("%-32s | %d\n", "Total max threads", p.max_threads_per_multiprocessor * p.multiprocessor_count);
    std::printf("%-32s | %d\n", "Wavefront size", p.warp_size);
}

[[noreturn]] void fail(const char* what, int st) {
    LOG_ERROR("error: %s failed: %s", what, bt_status_str(st));
    std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
    std::setlocale(LC_ALL, "");  // main.cu:33-35 (digit grouping in %'d)
    Args a;
    if (parse(argc, argv, a)) {
        LOG_ERROR("error: invalid arguments");
        usage();
        return EXIT_FAILURE;
    }
    int device_count = 0;
    bt_device_count(&device_count);
    if (device_count == 0) LOG_WARN("no gpu devices supporting HIP.");
    if (a.device_used && (a.device < 0 || a.device >= device_count)) {
        LOG_ERROR("error: gpu device id %d is not available", a.device);
        return EXIT_FAILURE;
    }
    if (a.device_info) {
        if (device_count == 0) LOG_WARN("no gpu devices supporting HIP.");
        else device_info(a.device);
    }
    if (!a.graph_used) return 0;

    bt_host_graph g{};
    if (a.bin_used) {  // main.cu:79-87: convert and exit
        LOG_INFO("processing graph file: %s", a.graph.c_str());
        int st = bt_graph_read_text(a.graph.c_str(), &g);
        if (st != BT_OK) {
            LOG_ERROR("Cannot convert %s: %s", a.graph.c_str(), bt_status_str(st));
            return EXIT_FAILURE;
        }
        LOG_INFO("graph with upper: %'u, lower: %'u, vertices: %'u, edges: %'u", g.u_num, g.l_num, g.n, g.m);
        LOG_INFO("convert graph to binary file: %s", a.bin.c_str());
        st = bt_graph_write_bin(a.bin.c_str(), &g);
        bt_graph_release(&g);
        if (st != BT_OK) {
            LOG_ERROR("Unable to open file: %s", a.bin.c_str());
            return EXIT_FAILURE;
        }
        return 0;
    }
    if (a.graph.find(".bin") == std::string::npos) {  // graph.cpp:271-274
        LOG_ERROR("graph file should be binary file");
        return EXIT_FAILURE;
    }
    LOG_INFO("loading graph from binary file: %s", a.graph.c_str());
    int st = bt_graph_read_bin(a.graph.c_str(), &g);
    if (st != BT_OK) {
        LOG_ERROR("Unable to open file: %s (%s)", a.graph.c_str(), bt_status_str(st));
        return EXIT_FAILURE;
    }
    LOG_INFO("graph with upper: %'u, lower: %'u, vertices: %'u, edges: %'u", g.u_num, g.l_num, g.n, g.m);

    if (a.cpu) {
        LOG_ERROR("error: --cpu is not served by this build; the reference CPU algorithms exist here only "
                  "as the test oracle");
        bt_graph_release(&g);
        return EXIT_FAILURE;
    }
    if (!a.gpu) {
        bt_graph_release(&g);
        return 0;
    }
    if (device_count == 0) {
        LOG_ERROR("error: no gpu devices supporting HIP to run algorithms.");
        return EXIT_FAILURE;
    }
    bt_ctx* ctx = nullptr;
    if ((st = bt_ctx_create(a.device, &ctx)) != BT_OK) fail("bt_ctx_create", st);
    bt_dgraph* dg = nullptr;
    if ((st = bt_graph_upload(ctx, &g, &dg)) != BT_OK) fail("bt_graph_upload", st);
    std::vector<uint64_t> out(g.m ? g.m : 1, 0);
    std::string algo = a.algo;
    // main.cu:126-127 sends any other name to msp; mspp and peel are the
    // reference's two further algorithms (bitruss_mspp, bitruss_peeling_cpu),
    // which its dispatch does not reach
    if (algo != "ebfc" && algo != "bfc" && algo != "hidx" && algo != "mspp" && algo != "peel") algo = "msp";
    uint64_t total = 0, smax = 0;
    uint32_t iters = 0;
    bool have_array = true;
    if (algo == "ebfc") {
        if ((st = bt_ebfc(ctx, dg, out.data(), &total, &smax)) != BT_OK) fail("bt_ebfc", st);
        LOG_INFO("butterfly counting with vertex priority on gpu, and the max edge support is %'llu",
                 (unsigned long long)smax);
        LOG_INFO("total butterfly counting with vertex priority on gpu: %'llu", (unsigned long long)total);
    } else if (algo == "bfc") {
        have_array = false;
        if ((st = bt_bfc(ctx, dg, &total)) != BT_OK) fail("bt_bfc", st);
        LOG_INFO("butterfly counting with vertex priority on gpu, and btf number: %'llu", (unsigned long long)total);
    } else if (algo == "hidx") {
        LOG_INFO("running bitruss decomposition based on h-index on GPU");
        if ((st = bt_hidx(ctx, dg, out.data(), &iters)) != BT_OK) fail("bt_hidx", st);
        uint64_t kmax = 0;
        for (uint32_t e = 0; e < g.m; e++) kmax = out[e] > kmax ? out[e] : kmax;
        LOG_INFO("k-bitruss max value %llu, with %u rounds", (unsigned long long)kmax, iters);
    } else if (algo == "mspp") {
        LOG_INFO("running bitruss decomposition online peeling (baseline)");  // bitruss_msp.cu:603
        if ((st = bt_mspp(ctx, dg, out.data(), &iters)) != BT_OK) fail("bt_mspp", st);
        LOG_INFO("(baseline) bitruss msp with %'u levels", iters);  // bitruss_msp.cu:693
    } else if (algo == "peel") {
        LOG_INFO("running bitruss decomposition by peeling");
        if ((st = bt_peel(ctx, dg, out.data(), &iters)) != BT_OK) fail("bt_peel", st);
        LOG_INFO("k-bitruss max value %u by peeling", iters);
    } else {
        LOG_INFO("running bitruss decomposition online peeling");
        if ((st = bt_msp(ctx, dg, out.data(), &iters)) != BT_OK) fail("bt_msp", st);
        LOG_INFO("bitruss msp with %'u levels", iters);
    }
    bt_timings t{};
    bt_ctx_last_timings(ctx, &t);
    if (a.json) {
        std::printf("{\"algo\": \"%s\", \"m\": %u, \"n\": %u, \"total\": %llu, \"support_max\": %llu, "
                    "\"iters\": %u, \"total_ms\": %.4f, \"prep_ms\": %.4f, \"count_ms\": %.4f, "
                    "\"peel_ms\": %.4f, \"sub_rounds\": %llu, \"wedges_count\": %llu}\n",
                    algo.c_str(), g.m, g.n, (unsigned long long)total, (unsigned long long)smax, iters,
                    t.total_ms, t.prep_ms, t.count_ms, t.peel_ms, (unsigned long long)t.sub_rounds,
                    (unsigned long long)t.wedges_count);
    }
    if (!a.dump.empty() && have_array) {
        FILE* f = std::fopen(a.dump.c_str(), "wb");
        if (!f || std::fwrite(out.data(), sizeof(uint64_t), g.m, f) != g.m) {
            LOG_ERROR("cannot write %s", a.dump.c_str());
            return EXIT_FAILURE;
        }
        std::fclose(f);
    }
    bt_graph_free(dg);
    bt_ctx_destroy(ctx);
    bt_graph_release(&g);
    return 0;
}
