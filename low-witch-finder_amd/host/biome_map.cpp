// biome_map — the map tool's biome ids (engine/tools/mineshaft/biome_map.c) on libwhf_hip (MI355X), with the same
// command line and the same output bytes:
//
//   biome_map --seed S --x0 X --z0 Z --w W --h H [--step B] [--ss K] [--ver V] [--y Y]
//     W*H bytes to stdout, row by row (z outer, x inner). Cell (i, j) covers the blocks [X + iB, X + (i+1)B) x
//     [Z + jB, Z + (j+1)B); each of its K x K sample points (K = 1..4) takes the biome at the middle of the chunk
//     that holds it, at occupied height Y (default 63: the reference's height callback answers Y + 1), and the
//     cell reports the dominant one, a swamp winning at a quarter of the samples (biome_map.c:101-126).
//   biome_map --seed S --stdin [--ver V] [--y Y]
//     one byte per "cx cz" line read from stdin: the biome the hut test sees at that chunk (biome_map.c:42-57): the
//     answer at occupied heights 40, 48, ..., 80 when all six say swamp or none does (the one at 64 when none
//     does), otherwise the answer at the chunk middle's pre-carver surface.
//
// V is a game version name (default 26.3): its biome tree and, for the surface, its terrain rules. All biome checks
// of a band of rows go to the device in one whf_biome call; each distinct chunk is asked once. Unlike the reference,
// coordinates beyond +-2^30 blocks (chunks beyond +-2^26) are refused as a usage error: the world spans +-3e7.
// Exit status 0 on success, 2 for a bad command line, 1 for a device or output error.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "whf.h"

namespace {

constexpr int kSwamp = WHF_BIOME_SWAMP;
constexpr uint64_t kBandSamples = 1u << 22;   // sample points per device call in the map mode

void usage() {
    std::fprintf(stderr, "usage: biome_map --seed S (--x0 X --z0 Z --w W --h H [--step B] [--ss 1..4] | --stdin) "
                         "[--ver 1.18.2|1.19.3|...|26.3] [--y Y]\n");
}

long long floor_div(long long a, long long b) {
    const long long q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

bool num(const char *s, long long &v) {
    char *e = nullptr;
    v = std::strtoll(s, &e, 10);
    return e && *e == 0 && e != s;
}

struct KeyHash {
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull >> 17); }
};

uint64_t chunk_key(long long cx, long long cz) { return ((uint64_t)(uint32_t)cx << 32) | (uint32_t)cz; }

// The winner of each query (chunk) at occupied height occ_y, or WHF_OCC_SURFACE; winners (97 bytes per query, may be
// null) receives the winner at every quart height.
int biome_of(whf_ctx *ctx, int64_t seed, const std::vector<uint64_t> &chunks, int occ_y, int tree, int vflags,
             std::vector<int> &winner, std::vector<uint8_t> *winners) {
    std::vector<whf_biome_query> q(chunks.size());
    for (size_t i = 0; i < chunks.size(); i++) {
        const int cx = (int32_t)(chunks[i] >> 32), cz = (int32_t)(uint32_t)chunks[i];
        q[i] = whf_biome_query{seed, cx * 16, cz * 16, occ_y, 0};
    }
    std::vector<whf_biome_result> r(chunks.size());
    if (winners) winners->assign(chunks.size() * WHF_BIOME_QY_N, 0);
    const int rc = whf_biome(ctx, q.data(), q.size(), tree, vflags, r.data(), winners ? winners->data() : nullptr);
    if (rc != WHF_OK) {
        std::fprintf(stderr, "biome_map: whf_biome: %s\n", whf_strerror(rc));
        return 1;
    }
    winner.resize(chunks.size());
    for (size_t i = 0; i < chunks.size(); i++) winner[i] = r[i].winner & 0xFF;
    return 0;
}

int run_map(whf_ctx *ctx, int64_t seed, long long x0, long long z0, int w, int h, int step, int ss, int y, int tree,
            int vflags) {
    const int n = ss * ss;
    const int rows_per_band = (int)std::max<uint64_t>(1, kBandSamples / ((uint64_t)w * n));
    std::vector<uint8_t> out;
    for (int j0 = 0; j0 < h; j0 += rows_per_band) {
        const int j1 = std::min(h, j0 + rows_per_band);
        // the chunk of every sample point of the band, each distinct chunk asked once
        std::vector<uint32_t> sample_chunk;
        sample_chunk.reserve((size_t)(j1 - j0) * w * n);
        std::unordered_map<uint64_t, uint32_t, KeyHash> index;
        std::vector<uint64_t> chunks;
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < w; i++)
                for (int b = 0; b < ss; b++) {
                    const long long Z = z0 + (long long)j * step + ((2LL * b + 1) * step) / (2LL * ss);
                    for (int a = 0; a < ss; a++) {
                        const long long X = x0 + (long long)i * step + ((2LL * a + 1) * step) / (2LL * ss);
                        const uint64_t k = chunk_key(floor_div(X, 16), floor_div(Z, 16));
                        auto it = index.find(k);
                        if (it == index.end()) {
                            it = index.emplace(k, (uint32_t)chunks.size()).first;
                            chunks.push_back(k);
                        }
                        sample_chunk.push_back(it->second);
                    }
                }
        std::vector<int> win;
        if (biome_of(ctx, seed, chunks, y, tree, vflags | WHF_BIO_OCC_ONLY, win, nullptr)) return 1;
        out.resize((size_t)(j1 - j0) * w);
        size_t s = 0;
        for (int j = j0; j < j1; j++)
            for (int i = 0; i < w; i++) {
                int cnt[256];
                std::memset(cnt, 0, sizeof cnt);
                int best = -1, bestn = 0;
                for (int k = 0; k < n; k++) {
                    const int id = win[sample_chunk[s++]];
                    if (++cnt[id] > bestn) {
                        bestn = cnt[id];
                        best = id;
                    }
                }
                if (cnt[kSwamp] * 4 >= n && cnt[kSwamp] > 0) best = kSwamp;   // a swamp wins at a quarter of the cell
                out[(size_t)(j - j0) * w + i] = (uint8_t)best;
            }
        if (std::fwrite(out.data(), 1, out.size(), stdout) != out.size()) return 1;
    }
    return 0;
}

int run_stdin(whf_ctx *ctx, int64_t seed, int tree, int vflags) {
    std::vector<uint64_t> chunks;
    long long cx, cz;
    while (std::scanf("%lld %lld", &cx, &cz) == 2) {
        if (std::llabs(cx) > (1ll << 26) || std::llabs(cz) > (1ll << 26)) {
            std::fprintf(stderr, "biome_map: chunk %lld %lld is outside the world\n", cx, cz);
            return 2;
        }
        chunks.push_back(chunk_key(cx, cz));
    }
    if (chunks.empty()) return 0;
    static const int ys[6] = {40, 48, 56, 64, 72, 80};
    std::vector<int> win;
    std::vector<uint8_t> at;   // winner per quart height
    if (biome_of(ctx, seed, chunks, 64, tree, vflags, win, &at)) return 1;
    std::vector<uint8_t> out(chunks.size());
    std::vector<uint64_t> edge;   // chunks where the six heights disagree: the real surface decides
    std::vector<size_t> edge_at;
    for (size_t i = 0; i < chunks.size(); i++) {
        int swamps = 0;
        for (int y : ys) swamps += at[i * WHF_BIOME_QY_N + (size_t)((y >> 2) - WHF_BIOME_QY0)] == kSwamp;
        if (swamps == 0 || swamps == 6) {
            out[i] = (uint8_t)(swamps ? kSwamp : win[i]);   // win[i]: the answer at occupied height 64
        } else {
            edge.push_back(chunks[i]);
            edge_at.push_back(i);
        }
    }
    if (!edge.empty()) {
        std::vector<int> ew;
        if (biome_of(ctx, seed, edge, WHF_OCC_SURFACE, tree, vflags | WHF_BIO_OCC_ONLY, ew, nullptr)) return 1;
        for (size_t k = 0; k < edge.size(); k++) out[edge_at[k]] = (uint8_t)ew[k];
    }
    return std::fwrite(out.data(), 1, out.size(), stdout) == out.size() ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    long long seed = 0, x0 = 0, z0 = 0, w = 0, h = 0, step = 16, ss = 1, y = 63;
    int have = 0;
    bool from_stdin = false;
    std::string ver = "26.3";
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const bool more = i + 1 < argc;
        long long v = 0;
        if (!std::strcmp(a, "--stdin")) { from_stdin = true; have |= 30; continue; }
        if (!more) { std::fprintf(stderr, "biome_map: unknown arg %s\n", a); return 2; }
        const char *val = argv[++i];
        if (!std::strcmp(a, "--ver")) { ver = val; continue; }
        if (!num(val, v)) { usage(); return 2; }
        if (!std::strcmp(a, "--seed")) { seed = v; have |= 1; }
        else if (!std::strcmp(a, "--x0")) { x0 = v; have |= 2; }
        else if (!std::strcmp(a, "--z0")) { z0 = v; have |= 4; }
        else if (!std::strcmp(a, "--w")) { w = v; have |= 8; }
        else if (!std::strcmp(a, "--h")) { h = v; have |= 16; }
        else if (!std::strcmp(a, "--step")) step = v;
        else if (!std::strcmp(a, "--ss")) ss = v;
        else if (!std::strcmp(a, "--y")) y = v;
        else { std::fprintf(stderr, "biome_map: unknown arg %s\n", a); return 2; }
    }
    if (from_stdin) w = h = 1;
    // block coordinates stay within +-2^30 (the world is +-3e7), so 16 * chunk fits the queries' 32-bit fields
    const long long lim = 1ll << 30;
    const bool in_world = std::llabs(x0) <= lim && std::llabs(z0) <= lim && std::llabs(x0 + w * step) <= lim &&
                          std::llabs(z0 + h * step) <= lim;
    if (have != 31 || w <= 0 || h <= 0 || w > 4096 || h > 4096 || step < 1 || step > (1 << 24) || ss < 1 || ss > 4 ||
        y < -64 || y > 319 || !in_world) {
        usage();
        return 2;
    }
    int tree = 0, vflags = 0;
    if (whf_biome_tree(ver.c_str(), &tree) != WHF_OK || whf_version_flags(ver.c_str(), &vflags) != WHF_OK) {
        std::fprintf(stderr, "biome_map: --ver must be a game version such as 1.18.2, 1.19.3, 1.21.4 or 26.3\n");
        return 2;
    }
    whf_ctx *ctx = nullptr;
    whf_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    if (whf_open(&cfg, &ctx) != WHF_OK) {
        std::fprintf(stderr, "biome_map: no device\n");
        return 1;
    }
    const int rc = from_stdin ? run_stdin(ctx, (int64_t)seed, tree, vflags)
                              : run_map(ctx, (int64_t)seed, x0, z0, (int)w, (int)h, (int)step, (int)ss, (int)y, tree,
                                        vflags);
    whf_close(ctx);
    std::fflush(stdout);
    return rc;
}
