// deepdive_scan — the Deep Dive on libwhf_hip (MI355X): mineshaft rooms under swamp-hut chunks over a range of
// 48-bit structure seeds, and optionally the 2^16 world seeds of each deep room through the phase-1 check.
//
// Outputs match the reference's Deep Dive files where the data is the same:
//   --dump-hits FILE  one line per spot, "S,cx,cz,X,1": the solver's hits format (hut_lookup_v4_source.cu:1341),
//                     which hunt.py turns into its spots file (hunt.py:1090-1097);
//   --out FILE        the expansion's finds, header "seed,chunkX,chunkZ,layoutX,roomFloor" (the terrain stage's
//                     file starts "seed,chunkX,chunkZ,layoutX,...", hut_lookup_v4_source.cu:897; its airHi/airLo
//                     columns come from the reference's layout catalogue, which this engine does not use, so the
//                     room floor stands in their place).
//
// --spots FILE runs the terrain stage of the reference's --spots mode instead of the sweep (hut_lookup_v4_source.cu:
// 1262-1292, started by hunt.py per chunk of solver hits, hunt.py:954-957): "S cx cz" lines (others skipped), each
// spot's start test, and every start expanded to its 2^16 world seeds through the phase-1 chain (--max-height);
// --out gets the reference's survivor file, "seed,chunkX,chunkZ,layoutX,airHi,airLo" with air columns 0 as the
// reference writes them in this mode. The reference's other options on that command line (--catalog, --batch,
// --ero-split, --ero-t1/-c0/-t3/-c1, --quiet) are accepted and have no effect here. --phase2-max-y T keeps only the
// survivors lysh's hut verdict accepts at avg_y <= T (hunt.py's finisher asks lysh for that first, hunt.py:281-287).
// A spot that is not a mineshaft start is reported and skipped and makes the exit status 1, as in the reference.
//
// --solve-only --solve5 runs the solver of the reference's hut_lookup_v4 (hut_lookup_v4_source.cu:1293-1403) on
// hunt.py's command line (hunt.py:947-952): --catalog FILE (text "X airHi airLo" lines; its layouts are the
// bundles when there is no --plan, with --regime XZ and --fused-tier T), --plan FILE ("X XZ class" lines),
// --first-index i --count n (groups; a bundle is 2^20 of them), --dump-hits FILE ("S,cx,cz,X,1" lines) and
// --quiet; --task-batch, --fused-blk/-grid/-gpb/-slots/-threads and --out are accepted and checked as the reference
// checks them, and shape nothing here. Every hit is re-verified on the host (carver seed, hut position, start
// test); a failure prints "*** HOST CHECK FAILED" and makes the exit status 1. stdout ends with a "solve:" line.
// Exit status 0 on success, 2 for a bad command line, 1 for a device or output error.
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "whf.h"

namespace {

void usage() {
    std::printf(
        "usage: deepdive_scan --s-first S --count N --min-region-x a --max-region-x b --min-region-z c --max-region-z d\n"
        "                     [--floor F] [--expand] [--max-height H] [--version V] [--dump-hits FILE] [--out FILE]\n"
        "                     [--stats]\n"
        "  structure seeds S + i (mod 2^48), i < N; hut regions [a,b) x [c,d); keep mineshaft rooms with floor <= F\n"
        "  (default -55); --expand checks each kept room's 2^16 world seeds with the phase-1 chain (maxHeight H,\n"
        "  default -50; version V, default 1.19.x~26.1)\n"
        "       deepdive_scan --spots FILE [--max-height H] [--version V] [--phase2-max-y T] [--out FILE] [--stats]\n"
        "  the terrain stage of the spots in FILE (\"S cx cz\" lines); --out gets seed,chunkX,chunkZ,layoutX,airHi,airLo\n"
        "       deepdive_scan --catalog FILE --solve-only --solve5 [--plan FILE | --regime XZ [--fused-tier T]]\n"
        "                     --first-index i --count n [--dump-hits FILE] [--quiet]\n"
        "  the solver: structure seeds whose hut chunk carries a layout of the plan (groups [i, i + n))\n");
}

// ---- the solver's command line ----

constexpr uint64_t kM48 = 0xFFFFFFFFFFFFull, kLA = 0x5DEECE66Dull;

int32_t lcg_next(uint64_t &st, int bits) {
    st = (st * kLA + 0xB) & kM48;
    return (int32_t)(st >> (48 - bits));
}

// The host's own check of a solver hit (the reference's, hut_lookup_v4_source.cu:1336-1341): the chunk's carver
// seed is the layout, the region's hut is at the chunk, and the layout is a mineshaft start (nextDouble() < 0.004).
bool host_check(uint64_t S, int32_t cx, int32_t cz, uint64_t X) {
    uint64_t st = (S ^ kLA) & kM48;
    const int64_t h1 = lcg_next(st, 32), h2 = lcg_next(st, 32), h3 = lcg_next(st, 32), h4 = lcg_next(st, 32);
    const uint64_t a = ((uint64_t)h1 << 32) + (uint64_t)h2, b = ((uint64_t)h3 << 32) + (uint64_t)h4;
    if (((((uint64_t)((int64_t)cx * (int64_t)a)) ^ ((uint64_t)((int64_t)cz * (int64_t)b)) ^ S) & kM48) != X) return false;
    const int32_t rx = cx >> 5, rz = cz >> 5;
    uint64_t r = (S + 14357620ull + (uint64_t)((int64_t)rx * 341873128712ll) + (uint64_t)((int64_t)rz * 132897987541ll));
    st = (r ^ kLA) & kM48;
    int v[2];
    for (int k = 0; k < 2; k++)
        for (;;) {
            const int32_t bits = lcg_next(st, 31), val = bits % 24;
            if ((int32_t)((uint32_t)bits - (uint32_t)val + 23u) >= 0) { v[k] = val; break; }
        }
    if (32 * rx + v[0] != cx || 32 * rz + v[1] != cz) return false;
    st = (X ^ kLA) & kM48;
    const int64_t hi = lcg_next(st, 26), lo = lcg_next(st, 27);
    return (double)((hi << 27) + lo) * 0x1.0p-53 < 0.004;
}

struct SolveArgs {
    std::string catalog, plan, hits_path;
    long long first = -1, count = 0;
    int regime_x = -1, regime_z = -1, tier = 1;
    bool quiet = false;
};

int run_solve(const SolveArgs &a) {
    std::vector<whf_solve_bundle> bundles;
    if (!a.plan.empty()) {
        FILE *pf = std::fopen(a.plan.c_str(), "r");
        if (!pf) { std::fprintf(stderr, "cannot read plan %s\n", a.plan.c_str()); return 1; }
        char ln[1024];
        while (std::fgets(ln, sizeof ln, pf)) {
            if (ln[0] == '#' || ln[0] == '\n') continue;
            unsigned long long x = 0;
            char rg[8] = {0}, cls[16] = "odd";
            if (std::sscanf(ln, "%llu %7s %15s", &x, rg, cls) < 2 || std::strlen(rg) != 2 || rg[0] < '0' || rg[0] > '2' ||
                rg[1] < '0' || rg[1] > '2') {
                std::fprintf(stderr, "bad plan line: %s", ln);
                std::fclose(pf);
                return 2;
            }
            const int tier = !std::strcmp(cls, "odd") ? 1 : !std::strcmp(cls, "u2") ? 2 : !std::strcmp(cls, "eq") ? 3 : 0;
            if (!tier) {
                std::fprintf(stderr, "plan class '%s' unknown (odd | u2 | eq)\n", cls);
                std::fclose(pf);
                return 2;
            }
            bundles.push_back(whf_solve_bundle{x & kM48, rg[0] - '0', rg[1] - '0', tier, 0});
        }
        std::fclose(pf);
    } else {
        if (a.regime_x < 0) { std::fprintf(stderr, "--solve5 needs --regime XZ (or --plan)\n"); return 2; }
        FILE *cf = std::fopen(a.catalog.c_str(), "r");
        if (!cf) { std::fprintf(stderr, "cannot read %s\n", a.catalog.c_str()); return 1; }
        char ln[1024];
        while (std::fgets(ln, sizeof ln, cf)) {
            unsigned long long x = 0;
            double hi = 0, lo = 0;
            if (ln[0] == '#' || std::sscanf(ln, "%llu %lf %lf", &x, &hi, &lo) != 3) continue;
            bundles.push_back(whf_solve_bundle{x & kM48, a.regime_x, a.regime_z, a.tier, 0});
        }
        std::fclose(cf);
    }
    if (bundles.empty()) { std::fprintf(stderr, "no bundles\n"); return 2; }
    if (a.first < 0 || a.count < 1 || a.first + a.count > ((long long)bundles.size() << 20)) {
        std::fprintf(stderr, "group range %lld..%lld outside the %zu bundles x 2^20 groups\n", a.first, a.first + a.count,
                     bundles.size());
        return 2;
    }
    FILE *hf = nullptr;
    if (!a.hits_path.empty()) {
        hf = std::fopen(a.hits_path.c_str(), "w");
        if (!hf) { std::fprintf(stderr, "cannot write %s\n", a.hits_path.c_str()); return 1; }
    }
    whf_ctx *ctx = nullptr;
    whf_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    int rc = whf_open(&cfg, &ctx);
    if (rc != WHF_OK) {
        std::fprintf(stderr, "whf_open: %s\n", whf_strerror(rc));
        if (hf) std::fclose(hf);
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    // one call per chunk of groups, so that progress shows and the hit buffer stays bounded
    const long long kStep = 4096;
    std::vector<whf_solve_hit> hits(1u << 22);
    std::vector<uint64_t> b_hits(bundles.size(), 0), b_groups(bundles.size(), 0), b_tasks(bundles.size(), 0);
    uint64_t tot_hits = 0, tot_bad = 0, tot_tasks = 0, tot_groups = 0;
    double device_ms = 0.0;
    for (long long g = a.first; g < a.first + a.count && rc == WHF_OK; g += kStep) {
        const long long n = std::min(kStep, a.first + a.count - g);
        whf_solve_sink sink{hits.data(), hits.size(), 0};
        whf_solve_stats st;
        rc = whf_solve(ctx, bundles.data(), bundles.size(), (uint64_t)g, (uint64_t)n, &sink, &st);
        if (rc != WHF_OK) { std::fprintf(stderr, "whf_solve: %s\n", whf_strerror(rc)); break; }
        device_ms += st.device_ms;
        tot_tasks += st.tasks;
        tot_groups += st.groups;
        // per bundle: groups and tasks of this call (a call may cross a bundle boundary)
        for (long long k = g; k < g + n; k++) b_groups[(size_t)(k >> 20)]++;
        if ((g >> 20) == ((g + n - 1) >> 20)) b_tasks[(size_t)(g >> 20)] += st.tasks;
        std::string out;
        char tmp[160];
        for (uint64_t i = 0; i < sink.count; i++) {
            const whf_solve_hit &h = hits[i];
            if (!host_check(h.s48, h.cx, h.cz, h.layout)) {
                tot_bad++;
                std::fprintf(stderr, "*** HOST CHECK FAILED: S=%llu X=%llu chunk (%d,%d)\n", (unsigned long long)h.s48,
                             (unsigned long long)h.layout, h.cx, h.cz);
                continue;
            }
            tot_hits++;
            b_hits[h.bundle]++;
            if (hf) {
                std::snprintf(tmp, sizeof tmp, "%llu,%d,%d,%llu,1\n", (unsigned long long)h.s48, h.cx, h.cz,
                              (unsigned long long)h.layout);
                out += tmp;
            }
        }
        if (hf && !out.empty() && std::fwrite(out.data(), 1, out.size(), hf) != out.size()) {
            std::fprintf(stderr, "cannot write %s\n", a.hits_path.c_str());
            rc = WHF_EINVAL;
        }
        if (!a.quiet)
            std::fprintf(stderr, "groups %lld..%lld: %llu hits | %.1f s\n", a.first, g + n, (unsigned long long)tot_hits,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    whf_close(ctx);
    if (hf) std::fclose(hf);
    if (rc != WHF_OK) return 1;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t b = 0; b < bundles.size(); b++)
        if (b_groups[b])
            std::printf("BUNDLE %zu layout %llu regime %d%d tier %d: %llu groups, %llu tasks, %llu hits\n", b,
                        (unsigned long long)bundles[b].layout, bundles[b].regime_x, bundles[b].regime_z, bundles[b].tier,
                        (unsigned long long)b_groups[b], (unsigned long long)b_tasks[b], (unsigned long long)b_hits[b]);
    std::printf("solve: %llu tasks in %llu groups, %llu hits (%.4f per task), host-check failures %llu\n",
                (unsigned long long)tot_tasks, (unsigned long long)tot_groups, (unsigned long long)tot_hits,
                tot_tasks ? (double)tot_hits / (double)tot_tasks : 0.0, (unsigned long long)tot_bad);
    std::printf("solve times: device %.3f s (%.4f us per task), wall %.3f s\n", device_ms / 1e3,
                tot_tasks ? 1e3 * device_ms / (double)tot_tasks : 0.0, wall);
    return tot_bad ? 1 : 0;
}

bool num(const char *s, long long &v) {
    char *e = nullptr;
    v = std::strtoll(s, &e, 10);
    return e && *e == 0 && e != s;
}


// The spots mode: reads FILE, runs whf_dd_terrain on up to 2^20 spots at a time, writes the survivors.
int run_spots(const std::string &path, const whf_dd_request &req, const std::string &version, bool have_p2,
              long long p2_max_y, const std::string &out_path, bool stats) {
    FILE *sf = std::fopen(path.c_str(), "r");
    if (!sf) { std::fprintf(stderr, "cannot read %s\n", path.c_str()); return 1; }
    std::vector<whf_dd_spot> spots;
    char ln[512];
    while (std::fgets(ln, sizeof ln, sf)) {
        long long S;
        int cx, cz;
        if (ln[0] == '#' || std::sscanf(ln, "%lld %d %d", &S, &cx, &cz) != 3) continue;
        whf_dd_spot sp;
        std::memset(&sp, 0, sizeof sp);
        sp.s48 = (uint64_t)S & 0xFFFFFFFFFFFFull;
        sp.cx = cx;
        sp.cz = cz;
        spots.push_back(sp);
    }
    std::fclose(sf);
    int tree = 0;
    if (have_p2 && whf_biome_tree(version.c_str(), &tree) != WHF_OK) {
        std::fprintf(stderr, "unknown version %s\n", version.c_str());
        return 2;
    }
    FILE *out = nullptr;
    if (!out_path.empty()) {
        out = std::fopen(out_path.c_str(), "w");
        if (!out) { std::fprintf(stderr, "cannot write %s\n", out_path.c_str()); return 1; }
        std::fprintf(out, "seed,chunkX,chunkZ,layoutX,airHi,airLo\n");
    }
    whf_ctx *ctx = nullptr;
    whf_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    int rc = whf_open(&cfg, &ctx);
    if (rc != WHF_OK) {
        std::fprintf(stderr, "whf_open: %s\n", whf_strerror(rc));
        if (out) std::fclose(out);
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const size_t kChunk = 1u << 20;
    std::vector<whf_hit> finds(1 << 20);
    uint64_t n_bad = 0, n_surv = 0, n_terrain = 0, starts = 0, expanded = 0;
    double device_ms = 0.0;
    for (size_t s0 = 0; s0 < spots.size() && rc == WHF_OK; s0 += kChunk) {
        const size_t n = std::min(kChunk, spots.size() - s0);
        whf_dd_spot *sp = spots.data() + s0;
        whf_sink fs{finds.data(), finds.size(), 0};
        whf_dd_stats st;
        rc = whf_dd_terrain(ctx, &req, sp, n, &fs, &st);
        if (rc != WHF_OK) { std::fprintf(stderr, "whf_dd_terrain: %s\n", whf_strerror(rc)); break; }
        for (size_t i = 0; i < n; i++)
            if (sp[i].pieces == 0) {
                n_bad++;
                std::fprintf(stderr, "*** not a mineshaft start: S=%" PRIu64 " chunk (%d,%d)\n", sp[i].s48, sp[i].cx,
                             sp[i].cz);
            }
        starts += st.starts;
        expanded += st.expanded_seeds;
        device_ms += st.device_ms;
        n_terrain += fs.count;
        std::vector<whf_hit> keep(finds.begin(), finds.begin() + (std::ptrdiff_t)fs.count);
        if (have_p2 && !keep.empty()) {
            std::vector<whf_hut_query> hq(keep.size());
            std::vector<whf_hut_result> hr(keep.size());
            for (size_t i = 0; i < keep.size(); i++) hq[i] = whf_hut_query{keep[i].seed, keep[i].cx, keep[i].cz};
            rc = whf_hut_eval(ctx, hq.data(), hq.size(), tree, req.vflags, (int)p2_max_y, hr.data());
            if (rc != WHF_OK) { std::fprintf(stderr, "whf_hut_eval: %s\n", whf_strerror(rc)); break; }
            size_t k = 0;
            for (size_t i = 0; i < keep.size(); i++)
                if (hr[i].ok) keep[k++] = keep[i];
            keep.resize(k);
        }
        n_surv += keep.size();
        if (out)
            for (const whf_hit &h : keep) {
                uint64_t X = 0;   // the spot's carver seed
                for (size_t i = 0; i < n; i++)
                    if (sp[i].s48 == ((uint64_t)h.seed & 0xFFFFFFFFFFFFull) && sp[i].cx == h.cx && sp[i].cz == h.cz) {
                        X = sp[i].carver;
                        break;
                    }
                std::fprintf(out, "%" PRId64 ",%d,%d,%" PRIu64 ",0,0\n", h.seed, h.cx, h.cz, X);
            }
    }
    whf_close(ctx);
    if (out) std::fclose(out);
    if (rc != WHF_OK) return 1;
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("spots: %zu spots (%" PRIu64 " not a start), %" PRIu64 " terrain survivors -> %s; terrain %.3f s "
                "(%.1f ms per spot)\n", spots.size() - (size_t)n_bad, n_bad, n_surv,
                out_path.empty() ? "-" : out_path.c_str(), t, spots.empty() ? 0.0 : 1e3 * t / (double)spots.size());
    if (stats)
        std::printf("{\"spots\": %zu, \"not_start\": %" PRIu64 ", \"starts\": %" PRIu64 ", \"expanded_seeds\": %" PRIu64
                    ", \"phase1\": %" PRIu64 ", \"survivors\": %" PRIu64 ", \"device_ms\": %.3f, \"wall_s\": %.3f}\n",
                    spots.size(), n_bad, starts, expanded, n_terrain, n_surv, device_ms, t);
    return n_bad ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
    unsigned long long s_first = 0, count = 0;
    long long rx0 = 0, rx1 = 0, rz0 = 0, rz1 = 0, floor_max = -55, max_h = -50;
    bool expand = false, stats = false, have_s = false, have_n = false, have_rect[4] = {false, false, false, false};
    long long p2_max_y = 0;
    bool have_p2 = false;
    std::string version = "1.19.x~26.1", hits_path, out_path, spots_path;
    bool solve_only = false, solve5 = false;
    SolveArgs sa;
    long long fused_blk = 64, fused_gpb = 32, fused_slots = 2, fused_threads = 2, task_batch = 1;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto next = [&](const char *&v) -> bool {
            if (i + 1 >= argc) return false;
            v = argv[++i];
            return true;
        };
        const char *v = nullptr;
        long long x = 0;
        if (a == "--help" || a == "-h") { usage(); return 0; }
        else if (a == "--expand") expand = true;
        else if (a == "--stats") stats = true;
        else if (a == "--s-first") { if (!next(v) || !num(v, x) || x < 0) { usage(); return 2; } s_first = (unsigned long long)x; have_s = true; }
        else if (a == "--count") { if (!next(v) || !num(v, x) || x < 0) { usage(); return 2; } count = (unsigned long long)x; have_n = true; }
        else if (a == "--min-region-x") { if (!next(v) || !num(v, rx0)) { usage(); return 2; } have_rect[0] = true; }
        else if (a == "--max-region-x") { if (!next(v) || !num(v, rx1)) { usage(); return 2; } have_rect[1] = true; }
        else if (a == "--min-region-z") { if (!next(v) || !num(v, rz0)) { usage(); return 2; } have_rect[2] = true; }
        else if (a == "--max-region-z") { if (!next(v) || !num(v, rz1)) { usage(); return 2; } have_rect[3] = true; }
        else if (a == "--floor") { if (!next(v) || !num(v, floor_max)) { usage(); return 2; } }
        else if (a == "--max-height") { if (!next(v) || !num(v, max_h)) { usage(); return 2; } }
        else if (a == "--version") { if (!next(v)) { usage(); return 2; } version = v; }
        else if (a == "--dump-hits") { if (!next(v)) { usage(); return 2; } hits_path = v; }
        else if (a == "--out") { if (!next(v)) { usage(); return 2; } out_path = v; }
        else if (a == "--spots") { if (!next(v)) { usage(); return 2; } spots_path = v; }
        else if (a == "--phase2-max-y") { if (!next(v) || !num(v, p2_max_y)) { usage(); return 2; } have_p2 = true; }
        else if (a == "--quiet") sa.quiet = true;
        else if (a == "--ero-split") {}
        else if (a == "--catalog") { if (!next(v)) { usage(); return 2; } sa.catalog = v; }
        else if (a == "--solve-only") solve_only = true;
        else if (a == "--solve5") { solve_only = true; solve5 = true; }
        else if (a == "--plan") { if (!next(v)) { usage(); return 2; } sa.plan = v; }
        else if (a == "--first-index") { if (!next(v) || !num(v, sa.first)) { usage(); return 2; } }
        else if (a == "--regime") {
            if (!next(v) || std::strlen(v) != 2 || v[0] < '0' || v[0] > '2' || v[1] < '0' || v[1] > '2') {
                std::fprintf(stderr, "--regime XZ: two digits 0 (exact), 1 (mid), 2 (far)\n");
                return 2;
            }
            sa.regime_x = v[0] - '0';
            sa.regime_z = v[1] - '0';
        }
        else if (a == "--fused-tier") { if (!next(v) || !num(v, x) || x < 1 || x > 3) { usage(); return 2; } sa.tier = (int)x; }
        else if (a == "--fused-blk") { if (!next(v) || !num(v, fused_blk)) { usage(); return 2; } }
        else if (a == "--fused-gpb") { if (!next(v) || !num(v, fused_gpb)) { usage(); return 2; } }
        else if (a == "--fused-slots") { if (!next(v) || !num(v, fused_slots)) { usage(); return 2; } }
        else if (a == "--fused-threads") { if (!next(v) || !num(v, fused_threads)) { usage(); return 2; } }
        else if (a == "--task-batch") { if (!next(v) || !num(v, task_batch)) { usage(); return 2; } }
        else if (a == "--fused-grid") { if (!next(v) || !num(v, x)) { usage(); return 2; } }
        else if (a == "--batch" || a == "--ero-t1" || a == "--ero-c0" || a == "--ero-t3" || a == "--ero-c1") {
            if (!next(v)) { usage(); return 2; }
        }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    if (solve_only) {
        if (!solve5) { std::fprintf(stderr, "only the --solve5 solver is built here\n"); return 2; }
        if (sa.catalog.empty()) { usage(); return 2; }
        if (fused_blk != 64 && fused_blk != 128 && fused_blk != 256 && fused_blk != 512) {
            std::fprintf(stderr, "--fused-blk must be 64, 128, 256 or 512\n");
            return 2;
        }
        if (fused_gpb < 1 || fused_gpb > 4096 || fused_slots < 1 || fused_slots > 8 || fused_threads < 1 ||
            fused_threads > 32) {
            std::fprintf(stderr, "--fused-gpb 1..4096, --fused-slots 1..8, --fused-threads 1..32\n");
            return 2;
        }
        if (task_batch < 1 || task_batch > (1 << 19)) { std::fprintf(stderr, "bad --task-batch (1 .. 524288)\n"); return 2; }
        sa.count = have_n ? (long long)count : 0;
        sa.hits_path = hits_path;
        return run_solve(sa);
    }
    if (!spots_path.empty()) {
        whf_dd_request req;
        std::memset(&req, 0, sizeof req);
        req.expand = 1;
        req.max_height = (int)max_h;
        if (whf_version_flags(version.c_str(), &req.vflags) != WHF_OK) {
            std::fprintf(stderr, "unknown version %s\n", version.c_str());
            usage();
            return 2;
        }
        return run_spots(spots_path, req, version, have_p2, p2_max_y, out_path, stats);
    }
    if (!have_s || !have_n || !(have_rect[0] && have_rect[1] && have_rect[2] && have_rect[3]) || rx1 <= rx0 ||
        rz1 <= rz0) {
        usage();
        return 2;
    }
    whf_dd_request req;
    std::memset(&req, 0, sizeof req);
    req.s_first = s_first;
    req.count = count;
    req.rx0 = (int)rx0; req.rx1 = (int)rx1; req.rz0 = (int)rz0; req.rz1 = (int)rz1;
    req.floor_max = (int)floor_max;
    req.expand = expand ? 1 : 0;
    req.max_height = (int)max_h;
    if (whf_version_flags(version.c_str(), &req.vflags) != WHF_OK) {
        std::fprintf(stderr, "unknown version %s\n", version.c_str());
        usage();
        return 2;
    }
    // the expansion's swamp check uses the version's own tree (a rule-set alias takes its default)
    int tree = 0;
    if (whf_biome_tree(version.c_str(), &tree) == WHF_OK) req.biome_tree = tree + 1;
    whf_ctx *ctx = nullptr;
    whf_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    int rc = whf_open(&cfg, &ctx);
    if (rc != WHF_OK) {
        std::fprintf(stderr, "whf_open: %s\n", whf_strerror(rc));
        return 1;
    }
    std::vector<whf_dd_spot> spots(1 << 20);
    std::vector<whf_hit> finds(1 << 20);
    whf_dd_sink ss{spots.data(), spots.size(), 0};
    whf_sink fs{finds.data(), finds.size(), 0};
    whf_dd_stats st;
    rc = whf_deepdive(ctx, &req, &ss, &fs, &st);
    whf_close(ctx);
    if (rc != WHF_OK) {
        std::fprintf(stderr, "whf_deepdive: %s\n", whf_strerror(rc));
        return 1;
    }
    if (!hits_path.empty()) {
        FILE *f = std::fopen(hits_path.c_str(), "w");
        if (!f) { std::fprintf(stderr, "cannot write %s\n", hits_path.c_str()); return 1; }
        for (uint64_t i = 0; i < ss.count; i++)
            std::fprintf(f, "%" PRIu64 ",%d,%d,%" PRIu64 ",1\n", spots[i].s48, spots[i].cx, spots[i].cz, spots[i].carver);
        std::fclose(f);
    }
    if (!out_path.empty()) {
        FILE *f = std::fopen(out_path.c_str(), "w");
        if (!f) { std::fprintf(stderr, "cannot write %s\n", out_path.c_str()); return 1; }
        std::fprintf(f, "seed,chunkX,chunkZ,layoutX,roomFloor\n");
        for (uint64_t i = 0; i < fs.count; i++) {
            uint64_t X = 0;   // the find's chunk carver seed: the spot's
            for (uint64_t k = 0; k < ss.count; k++)
                if (spots[k].s48 == ((uint64_t)finds[i].seed & 0xFFFFFFFFFFFFull) && spots[k].cx == finds[i].cx &&
                    spots[k].cz == finds[i].cz) { X = spots[k].carver; break; }
            std::fprintf(f, "%" PRId64 ",%d,%d,%" PRIu64 ",%d\n", finds[i].seed, finds[i].cx, finds[i].cz, X,
                         finds[i].floor_y);
        }
        std::fclose(f);
    }
    std::fprintf(stderr, "deepdive: %llu structure seeds, %llu starts, %llu spots, %llu world seeds expanded, %llu finds, "
                 "%.1f ms\n", (unsigned long long)st.structure_seeds, (unsigned long long)st.starts,
                 (unsigned long long)st.spots, (unsigned long long)st.expanded_seeds, (unsigned long long)fs.count,
                 st.device_ms);
    if (stats)
        std::printf("{\"structure_seeds\": %llu, \"tests\": %llu, \"starts\": %llu, \"walks\": [%llu, %llu, %llu], "
                    "\"spots\": %llu, \"expanded_seeds\": %llu, \"finds\": %llu, \"device_ms\": %.3f}\n",
                    (unsigned long long)st.structure_seeds, (unsigned long long)st.tests,
                    (unsigned long long)st.starts, (unsigned long long)st.walks[0], (unsigned long long)st.walks[1],
                    (unsigned long long)st.walks[2], (unsigned long long)st.spots,
                    (unsigned long long)st.expanded_seeds, (unsigned long long)fs.count, st.device_ms);
    return 0;
}
