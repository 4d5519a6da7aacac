// phase1_scan — drop-in replacement of the reference's GPU Seed Check engine, on libwhf_hip (MI355X).
//
// Same command line, --worker line protocol, progress text and candidate CSV as the reference engine
// (engine/gpu/low54/phase1_scan.cu:561-601 usage, :641-759 parsing, :1328-1339 CSV, :1348-1387 worker), so
// hunt.py's GpuScanWorker (hunt.py:1276-1332) drives it unchanged. Differences, all additive:
//   * candidates are exactly the reference CPU scanner's phase-1 survivors (lysh_phase1_check_ex), not a
//     superset; --max-height below -50 is clamped to -50 as lysh does (main.c:135-137);
//   * --version V selects the game's phase-1 rules by lysh's version names (default 1.19.x~26.1, hunt.py's
//     default; --v26-2 still selects the 26.2 ridge rule);
//   * --splitmix S N [OFFSET] scans N seeds of the sequential SplitMix64 stream seeded with S;
//   * --stats prints the funnel and erosion tier counters as one JSON line;
//   * --fp32-gate runs the erosion stage's corner blend in FP32 with FP64 re-evaluation near every threshold;
//   * in --worker mode, a request line with --id K is served asynchronously by the library's serving loop
//     (whf_server_open / whf_submit / whf_poll): the worker reads the next line at once, and prints
//     ===SCAN_DONE id=K rc=N=== when that request completes (after writing its --out file), in completion order.
//     Lines without --id keep the reference's one-at-a-time answer, so hunt.py's sequential use is unchanged.
// Options that only tuned the reference's own kernels (--block, --climate-gate, --no-climate-gate,
// --no-density-prefilter) are accepted and have no effect on the answer. Options of searches this engine
// does not run (--self-test, --quad-bases, --require-all-regions, --max-hut-dist) are errors.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cinttypes>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "whf.h"

namespace {

struct Params {
    std::vector<int64_t> seeds;
    int seed_mode = WHF_SEEDS_LIST;
    uint64_t gen_base = 0, gen_count = 0, gen_offset = 0;
    int minRX = -128, maxRX = 128, minRZ = -128, maxRZ = 128;
    int maxHeight = -40;
    int vflags = 0;
    bool have_version = false, v262 = false, stats = false, fp32 = false;
    std::string out, version;
    bool have_p2 = false, exactly = false;
    int p2_max_y = 0;
    std::string id;   // --worker: serve this request asynchronously and tag its completion line with the id
};

bool read_seed_list(const char *path, std::vector<int64_t> &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::string buf;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, n);
    std::fclose(f);
    const char *p = buf.c_str(), *end = p + buf.size();
    while (p < end) {
        while (p < end && !(std::isdigit((unsigned char)*p) || *p == '-')) p++;
        if (p >= end) break;
        char *next = nullptr;
        unsigned long long v = std::strtoull(p, &next, 10);
        if (next == p) { p++; continue; }
        out.push_back((int64_t)v);
        p = next;
    }
    return true;
}

void usage() {
    std::printf(
        "usage: phase1_scan --seed <n> [--seed <n> ...] [--seed-list <file>] [options]\n"
        "       phase1_scan --seed-gen <base> <count> [options]\n"
        "       phase1_scan --splitmix <stream> <count> [<offset>] [options]\n"
        "       phase1_scan --worker\n"
        "  --seed <n>            world seed (repeatable)\n"
        "  --seed-list <file>    more seeds, whitespace/comma separated\n"
        "  --seed-gen <base> <count>   seeds splitmix64(base + i) generated on the GPU\n"
        "  --splitmix <s> <count> [<offset>]  seeds of the sequential SplitMix64 stream seeded with s\n"
        "  --min-region-x <n>    default -128\n"
        "  --max-region-x <n>    default  128   (exclusive)\n"
        "  --min-region-z <n>    default -128\n"
        "  --max-region-z <n>    default  128   (exclusive)\n"
        "  --max-height <n>      phase-1 hut height, default -40 (below -50 counts as -50)\n"
        "  --phase2-max-y <y>    also run the hut evaluation (lysh hut) and keep the huts it accepts with floor <= y;\n"
        "                        gpuEst is then the hut's floor\n"
        "  --or-lower | --exactly  with --phase2-max-y: floor <= y (default) or floor == y\n"
        "  --version <v>         game version rules (lysh names; default 1.19.x~26.1)\n"
        "  --v26-2               26.2 ridge rule\n"
        "  --out <file>          write candidates as CSV (seed,chunkX,chunkZ,blockX,blockZ,gpuEst)\n"
        "  --stats               print the funnel counters as JSON\n"
        "  --fp32-gate           FP32-gated erosion stage with FP64 re-evaluation near thresholds (same answers)\n"
        "  --worker              read one request per stdin line; ===WORKER_READY===, ===SCAN_DONE rc=N===\n"
        "  --id <k>              (a --worker line) serve the request asynchronously beside the following lines;\n"
        "                        its answer is ===SCAN_DONE id=<k> rc=N===, in completion order\n");
}

// Parses argv[1..] or one --worker line (phase1_scan.cu:641-759). Returns false on a bad request.
bool parse(const std::vector<std::string> &t, Params &p) {
    bool ok = true;
    size_t i = 0;
    auto next = [&](const char *what) -> std::string {
        if (i + 1 >= t.size()) {
            std::fprintf(stderr, "missing value for %s\n", what);
            ok = false;
            return std::string();
        }
        return t[++i];
    };
    bool gen = false;
    for (; i < t.size() && ok; i++) {
        const std::string &a = t[i];
        if (a == "--seed") { std::string v = next("--seed"); if (ok) p.seeds.push_back((int64_t)std::strtoull(v.c_str(), nullptr, 10)); }
        else if (a == "--seed-list") {
            std::string v = next("--seed-list");
            if (ok && !read_seed_list(v.c_str(), p.seeds)) { std::fprintf(stderr, "cannot open seed list %s\n", v.c_str()); ok = false; }
        } else if (a == "--seed-gen" || a == "--splitmix") {
            std::string b = next(a.c_str());
            std::string c = ok ? next(a.c_str()) : std::string();
            if (!ok) break;
            gen = true;
            p.seed_mode = a == "--seed-gen" ? WHF_SEEDS_COUNTER : WHF_SEEDS_SPLITMIX;
            p.gen_base = std::strtoull(b.c_str(), nullptr, 10);
            p.gen_count = std::strtoull(c.c_str(), nullptr, 10);
            if (a == "--splitmix" && i + 1 < t.size() && t[i + 1].rfind("--", 0) != 0)
                p.gen_offset = std::strtoull(t[++i].c_str(), nullptr, 10);
        }
        else if (a == "--min-region-x") { std::string v = next(a.c_str()); p.minRX = std::atoi(v.c_str()); }
        else if (a == "--max-region-x") { std::string v = next(a.c_str()); p.maxRX = std::atoi(v.c_str()); }
        else if (a == "--min-region-z") { std::string v = next(a.c_str()); p.minRZ = std::atoi(v.c_str()); }
        else if (a == "--max-region-z") { std::string v = next(a.c_str()); p.maxRZ = std::atoi(v.c_str()); }
        else if (a == "--max-height") { std::string v = next(a.c_str()); p.maxHeight = std::atoi(v.c_str()); }
        else if (a == "--version") {
            std::string v = next(a.c_str());
            if (ok && whf_version_flags(v.c_str(), &p.vflags) != WHF_OK) { std::fprintf(stderr, "unknown --version %s\n", v.c_str()); ok = false; }
            p.have_version = true;
            p.version = v;
        }
        else if (a == "--phase2-max-y") { std::string v = next(a.c_str()); p.p2_max_y = std::atoi(v.c_str()); p.have_p2 = true; }
        else if (a == "--exactly") p.exactly = true;
        else if (a == "--or-lower") p.exactly = false;
        else if (a == "--v26-2") p.v262 = true;
        else if (a == "--out") p.out = next("--out");
        else if (a == "--stats") p.stats = true;
        else if (a == "--fp32-gate") p.fp32 = true;
        else if (a == "--no-density-prefilter" || a == "--climate-gate" || a == "--no-climate-gate") {}
        else if (a == "--block") next("--block");
        else if (a == "--worker") {}
        else if (a == "--id") { p.id = next("--id"); }
        else if (a == "--self-test" || a == "--quad-bases" || a == "--quad-base-offset" || a == "--quad-base-limit" ||
                 a == "--require-all-regions" || a == "--max-hut-dist" || a == "--gate-center-x" || a == "--gate-center-z") {
            std::fprintf(stderr, "%s is not supported by this engine\n", a.c_str());
            ok = false;
        }
        else { std::fprintf(stderr, "unknown argument: %s\n", a.c_str()); ok = false; }
    }
    if (!ok) return false;
    if (gen && !p.seeds.empty()) { std::fprintf(stderr, "--seed-gen/--splitmix cannot be combined with --seed/--seed-list\n"); return false; }
    if (gen && p.gen_count == 0) { std::fprintf(stderr, "--seed-gen count must be positive\n"); return false; }
    if (!gen && p.seeds.empty()) { std::fprintf(stderr, "no seeds given\n"); return false; }
    if (p.maxRX <= p.minRX || p.maxRZ <= p.minRZ) { std::fprintf(stderr, "empty region range\n"); return false; }
    if (!p.have_version) whf_version_flags("1.19.x~26.1", &p.vflags);
    if (p.v262) p.vflags |= WHF_V_262;
    return true;
}

whf_request make_request(const Params &p) {
    whf_request r;
    std::memset(&r, 0, sizeof r);
    r.seed_mode = p.seed_mode;
    r.precision = p.fp32 ? WHF_PREC_FP32_GATE : WHF_PREC_FP64;
    if (p.seed_mode == WHF_SEEDS_LIST) {
        r.seeds = p.seeds.data();
        r.n_seeds = p.seeds.size();
    } else {
        r.stream = p.gen_base;
        r.offset = p.gen_offset;
        r.n_seeds = p.gen_count;
    }
    r.rx0 = p.minRX; r.rx1 = p.maxRX; r.rz0 = p.minRZ; r.rz1 = p.maxRZ;
    r.max_height = p.maxHeight < -50 ? -50 : p.maxHeight;
    r.vflags = p.vflags;
    if (p.have_p2) {
        r.phase2 = p.exactly ? WHF_P2_EXACTLY : WHF_P2_OR_LOWER;
        r.phase2_max_y = p.p2_max_y;
        int tree = 0;
        // a version with a tree of its own selects it, a rule-set alias takes its default (eval.c:104-106)
        r.biome_tree = whf_biome_tree(p.version.empty() ? "1.19.x~26.1" : p.version.c_str(), &tree) == WHF_OK ? tree + 1 : 0;
    }
    return r;
}

// The progress line, the --stats line and the --out CSV of a finished request (phase1_scan.cu:1328-1339).
int report(const Params &p, std::vector<whf_hit> &hits, const whf_stats &st, double secs) {
    std::sort(hits.begin(), hits.end(), [](const whf_hit &a, const whf_hit &b) {
        if (a.seed != b.seed) return a.seed < b.seed;
        if (a.cx != b.cx) return a.cx < b.cx;
        return a.cz < b.cz;
    });
    std::printf("kernel: %.3fs, %zu candidates, %.1f M region-seed-checks/s (%" PRIu64 " seeds x %" PRIu64 " regions)\n",
                secs, hits.size(), secs > 0 ? (double)st.checks / secs / 1e6 : 0.0, st.seeds,
                st.seeds ? st.checks / st.seeds : 0);
    if (p.stats) {
        std::printf("{\"checks\": %" PRIu64 ", \"accepted\": %" PRIu64 ", \"reached_ladder\": %" PRIu64 ", \"funnel\": [",
                    st.checks, st.accepted, st.reached_ladder);
        for (int k = 0; k < WHF_KIND_N; k++) std::printf("%s%" PRIu64, k ? ", " : "", st.funnel[k]);
        std::printf("], \"tiers\": [");
        for (int k = 0; k < 7; k++) std::printf("%s%" PRIu64, k ? ", " : "", st.tiers[k]);
        std::printf("], \"samplers_built\": [%" PRIu64 ", %" PRIu64 ", %" PRIu64 "], \"device_ms\": %.3f, \"phase2\": [%" PRIu64
                    ", %" PRIu64 ", %" PRIu64 ", %" PRIu64 ", %" PRIu64 "]}\n",
                    st.samplers_built[0], st.samplers_built[1], st.samplers_built[2], st.device_ms, st.phase2[0],
                    st.phase2[1], st.phase2[2], st.phase2[3], st.phase2[4]);
    }
    if (!p.out.empty()) {
        FILE *f = std::fopen(p.out.c_str(), "w");
        if (!f) { std::fprintf(stderr, "cannot write %s\n", p.out.c_str()); return 1; }
        std::fprintf(f, "seed,chunkX,chunkZ,blockX,blockZ,gpuEst\n");
        // gpuEst: the hut's floor with --phase2-max-y; phase 1 alone computes no height, and the column stays for
        // readers that expect six fields
        for (const whf_hit &h : hits) {
            if (h.flags & WHF_HIT_PHASE2)
                std::fprintf(f, "%lld,%d,%d,%d,%d,%.3f\n", (long long)h.seed, h.cx, h.cz, h.cx * 16, h.cz * 16, (double)h.floor_y);
            else
                std::fprintf(f, "%lld,%d,%d,%d,%d,nan\n", (long long)h.seed, h.cx, h.cz, h.cx * 16, h.cz * 16);
        }
        std::fclose(f);
        std::printf("wrote %zu candidates to %s\n", hits.size(), p.out.c_str());
    }
    return 0;
}

int run(whf_ctx *ctx, const Params &p) {
    const whf_request r = make_request(p);
    std::vector<whf_hit> hits(1 << 16);
    whf_sink sink{hits.data(), hits.size(), 0};
    whf_stats st;
    int rc = whf_scan(ctx, &r, &sink, &st);
    if (rc == WHF_EOVERFLOW) {
        hits.resize(sink.count);
        sink = whf_sink{hits.data(), hits.size(), 0};
        rc = whf_scan(ctx, &r, &sink, &st);
    }
    if (rc != WHF_OK) {
        std::fprintf(stderr, "scan failed: %s\n", whf_strerror(rc));
        return 1;
    }
    hits.resize(std::min<uint64_t>(sink.count, hits.size()));
    return report(p, hits, st, st.device_ms / 1000.0);
}

// --worker requests with --id: submitted to the serving loop, answered by a poller thread in completion order.
// Output lines of the two threads go through one lock, each request's lines together.
struct Async {
    struct Pending {
        Params p;
        whf_request r;
        std::vector<whf_hit> hits;
        whf_sink sink;
    };
    whf_server *srv = nullptr;
    std::mutex mu;                                   // pending and stdout
    std::map<uint64_t, std::unique_ptr<Pending>> pending;
    std::thread poller;
    std::atomic<bool> closing{false};

    // Opens the server on first use; false if the device cannot.
    bool ensure() {
        if (srv) return true;
        whf_server_config cfg;
        std::memset(&cfg, 0, sizeof cfg);
        const int rc = whf_server_open(&cfg, &srv);
        if (rc != WHF_OK) {
            std::fprintf(stderr, "cannot open the serving loop: %s\n", whf_strerror(rc));
            srv = nullptr;
            return false;
        }
        poller = std::thread(&Async::poll_loop, this);
        return true;
    }

    // mu held
    int submit_locked(std::unique_ptr<Pending> q) {
        q->sink = whf_sink{q->hits.data(), q->hits.size(), 0};
        uint64_t sid = 0;
        const int rc = whf_submit(srv, &q->r, &q->sink, &sid);
        if (rc == WHF_OK) pending[sid] = std::move(q);
        return rc;
    }

    void submit(const Params &p) {
        std::lock_guard<std::mutex> lk(mu);
        if (!ensure()) {
            std::printf("===SCAN_DONE id=%s rc=1===\n", p.id.c_str());
            return;
        }
        auto q = std::make_unique<Pending>();
        q->p = p;
        q->r = make_request(q->p);   // seeds point into q->p
        q->hits.resize(1 << 16);
        const int rc = submit_locked(std::move(q));
        if (rc != WHF_OK) {
            std::fprintf(stderr, "request %s refused: %s\n", p.id.c_str(), whf_strerror(rc));
            std::printf("===SCAN_DONE id=%s rc=1===\n", p.id.c_str());
        }
    }

    void poll_loop() {
        whf_event ev[16];
        for (;;) {
            const int n = whf_poll(srv, ev, 16, 50);
            if (n < 0) return;
            for (int k = 0; k < n; k++) {
                std::lock_guard<std::mutex> lk(mu);
                auto it = pending.find(ev[k].id);
                if (it == pending.end()) continue;
                std::unique_ptr<Pending> q = std::move(it->second);
                pending.erase(it);
                if (ev[k].status == WHF_EOVERFLOW && q->sink.count > q->hits.size()) {
                    // more survivors than the first buffer: the same request again with room for all of them
                    const std::string id = q->p.id;
                    q->hits.resize(q->sink.count);
                    if (submit_locked(std::move(q)) == WHF_OK) continue;
                    std::printf("===SCAN_DONE id=%s rc=1===\n", id.c_str());
                    std::fflush(stdout);
                    continue;
                }
                int rc = 1;
                if (ev[k].status == WHF_OK) {
                    q->hits.resize(std::min<uint64_t>(q->sink.count, q->hits.size()));
                    rc = report(q->p, q->hits, ev[k].stats, ev[k].done_s - ev[k].start_s);
                } else {
                    std::fprintf(stderr, "request %s failed: %s\n", q->p.id.c_str(), whf_strerror(ev[k].status));
                }
                std::printf("===SCAN_DONE id=%s rc=%d===\n", q->p.id.c_str(), rc);
                std::fflush(stdout);
            }
            if (closing) {
                std::lock_guard<std::mutex> lk(mu);
                if (pending.empty()) return;
            }
        }
    }

    // End of input: every submitted request is answered, then the loop closes.
    void drain() {
        if (!srv) return;
        closing = true;
        poller.join();
        whf_server_close(srv);
        srv = nullptr;
    }
};

std::vector<std::string> tokenize(std::string line) {
    if (line.size() >= 3 && (unsigned char)line[0] == 0xEF && (unsigned char)line[1] == 0xBB &&
        (unsigned char)line[2] == 0xBF)
        line.erase(0, 3);
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

int open_ctx(whf_ctx **ctx) {
    whf_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    int rc = whf_open(&cfg, ctx);
    if (rc != WHF_OK) std::fprintf(stderr, "cannot open the GPU: %s\n", whf_strerror(rc));
    return rc;
}

}  // namespace

int main(int argc, char **argv) {
    std::setvbuf(stdout, nullptr, _IOLBF, 1024);
    std::vector<std::string> tokens(argv + 1, argv + argc);
    bool worker = std::find(tokens.begin(), tokens.end(), "--worker") != tokens.end();
    if (!worker) {
        if (tokens.empty()) { usage(); return 2; }
        for (auto &t : tokens)
            if (t == "--help" || t == "-h") { usage(); return 0; }
        Params p;
        if (!parse(tokens, p)) { usage(); return 2; }
        if (!p.id.empty()) { std::fprintf(stderr, "--id is for --worker request lines\n"); usage(); return 2; }
        whf_ctx *ctx = nullptr;
        if (open_ctx(&ctx) != WHF_OK) return 1;
        int rc = run(ctx, p);
        whf_close(ctx);
        return rc;
    }
    whf_ctx *ctx = nullptr;
    if (open_ctx(&ctx) != WHF_OK) return 1;
    std::printf("===WORKER_READY===\n");
    std::fflush(stdout);
    Async async;
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        Params p;
        if (!parse(tokenize(line), p)) {
            std::lock_guard<std::mutex> lk(async.mu);
            std::printf("===SCAN_ERROR: bad request line===\n");
            std::fflush(stdout);
            continue;
        }
        if (!p.id.empty()) {
            async.submit(p);
            std::fflush(stdout);
            continue;
        }
        // a sequential request runs on the worker's own context; its lines stay together
        std::unique_lock<std::mutex> lk(async.mu, std::defer_lock);
        if (async.srv) lk.lock();
        int rc = run(ctx, p);
        std::printf("===SCAN_DONE rc=%d===\n", rc);
        std::fflush(stdout);
    }
    async.drain();
    whf_close(ctx);
    return 0;
}
