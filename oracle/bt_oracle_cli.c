/*
 * bt_oracle_cli.c — command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the CPU baseline and the checker):
 *   bt_oracle_cli --graph X.bin --algo bfc|ebfc|hidx|msp [--threads N]
 *                 [--dump out.u64]
 * mirrors the reference's `--cpu --algo X --threads N` path (main.cu:96-110,
 * including `ebfc` forced to 32 threads, main.cu:98, and the unknown-algo
 * fallback to hidx, main.cu:106). It prints one JSON line with the result
 * summary, the algorithm time (excluding the .bin load) and the work
 * counters. --dump writes the per-edge uint64 array (supports for ebfc,
 * bitruss numbers for hidx/msp) as raw little-endian uint64.
 */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bt_oracle.h"

static uint32_t* read_u32(FILE* f, uint64_t count) {
    uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * (count ? count : 1));
    if (!a) return NULL;
    if (count && fread(a, sizeof(uint32_t), count, f) != count) {
        free(a);
        return NULL;
    }
    return a;
}

int main(int argc, char** argv) {
    const char* graph = NULL;
    const char* algo = "msp";
    const char* dump = NULL;
    int threads = 1;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--graph") && i + 1 < argc) graph = argv[++i];
        else if (!strcmp(argv[i], "--algo") && i + 1 < argc) algo = argv[++i];
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
        else if (!strcmp(argv[i], "--cpu")) continue;
        else {
            fprintf(stderr, "usage: %s --graph X.bin --algo bfc|ebfc|hidx|msp [--threads N] [--dump F]\n",
                    argv[0]);
            return 1;
        }
    }
    if (!graph || !strstr(graph, ".bin")) {
        fprintf(stderr, "graph file should be binary file\n");
        return 1;
    }
    FILE* f = fopen(graph, "rb");
    if (!f) {
        fprintf(stderr, "Unable to open file: %s\n", graph);
        return 1;
    }
    uint32_t h[5];
    if (fread(h, 4, 5, f) != 5) { fclose(f); return 1; }
    const uint64_t m = h[3], n = (uint64_t)h[0] + h[1];
    uint32_t* nbr = read_u32(f, 2 * m);
    uint32_t* off = read_u32(f, n + 1);
    uint32_t* deg = read_u32(f, n);
    fseek(f, (long)(4 * m), SEEK_CUR);
    uint32_t* edges = read_u32(f, 2 * m);
    uint32_t* eids = read_u32(f, 2 * m);
    fclose(f);
    if (!nbr || !off || !deg || !edges || !eids) {
        fprintf(stderr, "truncated graph file\n");
        return 1;
    }
    or_graph g = {h[0], h[1], (uint32_t)n, (uint32_t)m, nbr, off, deg, edges, eids};
    uint64_t* out = (uint64_t*)calloc(m ? m : 1, sizeof(uint64_t));
    or_stats st;
    memset(&st, 0, sizeof(st));
    uint64_t total = 0, smax = 0;
    uint32_t iters = 0;
    int rc;
    const char* kind;
    double t0 = omp_get_wtime();
    if (!strcmp(algo, "ebfc")) {
        threads = 32; /* main.cu:98 */
        kind = "ebfc";
        rc = or_ebfc(&g, threads, out, &total, &smax, &st);
    } else if (!strcmp(algo, "bfc")) {
        kind = "bfc";
        rc = or_bfc(&g, threads, &total, &st);
    } else if (!strcmp(algo, "msp")) {
        kind = "msp";
        rc = or_msp(&g, threads, out, &iters, &st);
    } else {
        kind = "hidx"; /* main.cu:104-106 */
        rc = or_hidx(&g, threads, out, &iters, &st);
    }
    double t1 = omp_get_wtime();
    if (rc) {
        fprintf(stderr, "oracle failed with status %d\n", rc);
        return 1;
    }
    uint64_t kmax = 0;
    if (strcmp(kind, "bfc") && strcmp(kind, "ebfc"))
        for (uint64_t e = 0; e < m; e++)
            if (out[e] > kmax) kmax = out[e];
    printf("{\"algo\": \"%s\", \"threads\": %d, \"m\": %llu, \"n\": %llu, \"seconds\": %.6f, "
           "\"total\": %llu, \"support_max\": %llu, \"iters\": %u, \"kmax\": %llu, "
           "\"w_cnt\": %llu, \"w_peel\": %llu, \"i_peel\": %llu, \"b_peel\": %llu, "
           "\"d_peel\": %llu, \"sum_curr\": %llu, \"sub_rounds\": %llu, "
           "\"w_h\": %llu, \"i_h\": %llu, \"b_h\": %llu}\n",
           kind, threads, (unsigned long long)m, (unsigned long long)n, t1 - t0,
           (unsigned long long)total, (unsigned long long)smax, iters,
           (unsigned long long)kmax, (unsigned long long)st.w_cnt,
           (unsigned long long)st.w_peel, (unsigned long long)st.i_peel,
           (unsigned long long)st.b_peel, (unsigned long long)st.d_peel,
           (unsigned long long)st.sum_curr, (unsigned long long)st.sub_rounds,
           (unsigned long long)st.w_h, (unsigned long long)st.i_h, (unsigned long long)st.b_h);
    if (dump && strcmp(kind, "bfc")) {
        FILE* o = fopen(dump, "wb");
        if (!o || fwrite(out, sizeof(uint64_t), m, o) != m) {
            fprintf(stderr, "cannot write %s\n", dump);
            return 1;
        }
        fclose(o);
    }
    free(out); free(nbr); free(off); free(deg); free(edges); free(eids);
    return 0;
}
