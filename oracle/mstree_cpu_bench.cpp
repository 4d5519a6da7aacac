// mstree_cpu_bench.cpp — CPU baseline harness for the Deep Dive sweep, built against the reference's own
// mineshaft code (engine/tools/mstree: swamphut.h for the hut placement and Java random, mineshaft_tree.h and
// mineshaft_tree_iter.h for the layout walk). Test infrastructure: the benchmark's CPU-baseline leg runs it; the
// product never does.
//
// For S = s_first + i (mod 2^48), i < count, and every hut region of the rect: the hut chunk
// (swampHutGetInRegion), the carver seed X = (cx*a ^ cz*b ^ S) with a, b the first two nextLong of Random(S)
// (as mstree.cpp:41-44 computes it), the start test (mstree.cpp:10-15, restated here because the reference keeps
// it file-static), and for starts the walk (mtRunInit/mtRunStep/mtRunFinish) with the reference's abortBelow cut
// at -12 - floor_max. Prints one JSON line with the counts and the rate. Threads split the S range.
//
//   mstree_cpu_bench --s-first S --count N --rx0 a --rx1 b --rz0 c --rz1 d --floor F --threads T [--time-limit sec]
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "mineshaft_tree.h"
#include "mineshaft_tree_iter.h"
#include "swamphut.h"

namespace {

bool start_passes(uint64_t X) {
    uint64_t s = (X ^ MT_MULT) & MT_MASK;
    s = (s * MT_MULT + 0xBULL) & MT_MASK;
    const uint64_t a = s >> 22;
    s = (s * MT_MULT + 0xBULL) & MT_MASK;
    const uint64_t b = s >> 21;
    return (double)((a << 27) + b) * 0x1.0p-53 < 0.004;
}

struct Counts {
    uint64_t seeds = 0, tests = 0, starts = 0, status[4] = {0, 0, 0, 0}, spots = 0;
};

}  // namespace

int main(int argc, char **argv) {
    uint64_t s_first = 0, count = 1 << 16;
    int rx0 = -2, rx1 = 2, rz0 = -2, rz1 = 2, floor_max = -55, threads = 1;
    double limit = 0;
    for (int i = 1; i + 1 < argc; i += 2) {
        const char *k = argv[i], *v = argv[i + 1];
        if (!strcmp(k, "--s-first")) s_first = strtoull(v, nullptr, 10);
        else if (!strcmp(k, "--count")) count = strtoull(v, nullptr, 10);
        else if (!strcmp(k, "--rx0")) rx0 = atoi(v);
        else if (!strcmp(k, "--rx1")) rx1 = atoi(v);
        else if (!strcmp(k, "--rz0")) rz0 = atoi(v);
        else if (!strcmp(k, "--rz1")) rz1 = atoi(v);
        else if (!strcmp(k, "--floor")) floor_max = atoi(v);
        else if (!strcmp(k, "--threads")) threads = atoi(v);
        else if (!strcmp(k, "--time-limit")) limit = atof(v);
        else { fprintf(stderr, "unknown option %s\n", k); return 2; }
    }
    if (threads < 1) threads = 1;
    const SwampHutConfig cfg = swampHutConfigModern();
    const int cut = -12 - floor_max;
    std::vector<Counts> per(threads);
    std::atomic<uint64_t> next{0};
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t chunk = 4096;
    auto worker = [&](int t) {
        MtTree *tree = new MtTree();
        MtRun *run = new MtRun();
        Counts &c = per[t];
        for (;;) {
            const uint64_t i0 = next.fetch_add(chunk);
            if (i0 >= count) break;
            if (limit > 0 && std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > limit) break;
            const uint64_t i1 = i0 + chunk < count ? i0 + chunk : count;
            for (uint64_t i = i0; i < i1; i++) {
                const int64_t s48 = (int64_t)((s_first + i) & MT_MASK);
                JavaRandom jr;
                jr.setSeed(s48);
                const int64_t a = javaNextLong(jr), b = javaNextLong(jr);
                c.seeds++;
                for (int rz = rz0; rz < rz1; rz++) {
                    for (int rx = rx0; rx < rx1; rx++) {
                        const ChunkPos p = swampHutGetInRegion(s48, rx, rz, cfg);
                        const uint64_t X = (uint64_t)(((int64_t)p.x * a) ^ ((int64_t)p.z * b) ^ s48) & MT_MASK;
                        c.tests++;
                        if (!start_passes(X)) continue;
                        c.starts++;
                        mtSeed(*tree, X);
                        mtRunInit(*tree, *run, 0, 0, cut, 0x7fffffff);
                        while (mtRunStep(*tree, *run)) {}
                        MtResult r;
                        mtRunFinish(*tree, &r);
                        c.status[r.status & 3]++;
                        if (r.status == 0 && r.roomFloor <= floor_max) c.spots++;
                    }
                }
            }
        }
        delete tree;
        delete run;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) th.emplace_back(worker, t);
    for (auto &x : th) x.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    Counts tot;
    for (const Counts &c : per) {
        tot.seeds += c.seeds;
        tot.tests += c.tests;
        tot.starts += c.starts;
        tot.spots += c.spots;
        for (int k = 0; k < 4; k++) tot.status[k] += c.status[k];
    }
    printf("{\"structure_seeds\": %llu, \"tests\": %llu, \"starts\": %llu, \"walks\": [%llu, %llu, %llu], "
           "\"spots\": %llu, \"seconds\": %.4f, \"seeds_per_s\": %.1f, \"threads\": %d}\n",
           (unsigned long long)tot.seeds, (unsigned long long)tot.tests, (unsigned long long)tot.starts,
           (unsigned long long)tot.status[0], (unsigned long long)tot.status[1], (unsigned long long)tot.status[2],
           (unsigned long long)tot.spots, secs, tot.seeds / secs, threads);
    return 0;
}
