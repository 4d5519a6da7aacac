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
// Exit status 0 on success, 2 for a bad command line, 1 for a device or output error.
#include <cinttypes>
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
        "  default -50; version V, default 1.19.x~26.1)\n");
}

bool num(const char *s, long long &v) {
    char *e = nullptr;
    v = std::strtoll(s, &e, 10);
    return e && *e == 0 && e != s;
}

}  // namespace

int main(int argc, char **argv) {
    unsigned long long s_first = 0, count = 0;
    long long rx0 = 0, rx1 = 0, rz0 = 0, rz1 = 0, floor_max = -55, max_h = -50;
    bool expand = false, stats = false, have_s = false, have_n = false, have_rect[4] = {false, false, false, false};
    std::string version = "1.19.x~26.1", hits_path, out_path;
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
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
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
