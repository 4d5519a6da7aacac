// whf_server.cpp — the serving C ABI (whf.h). By default Seed Check requests go to the device-resident loop of
// whf_serve.hip (a device request table, batches packed on the device across requests, per-request counters and
// survivor queues on the device); with WHF_SERVE_SLOTS they go to the host-driven slots below instead. Background Deep
// Dive jobs always run on slots: with the device loop, one slot context of their own takes a Deep Dive chunk whenever
// the loop has no Seed Check work left to claim.
//
// Slot mode: continuous batching of Seed Check requests over several device slots.
//
// Background Deep Dive jobs share the slots: a slot takes a Deep Dive chunk (2^18 structure seeds: sweep, walks and
// the expansion of its spots) only when no Seed Check chunk is waiting, and at most half the slots run Deep Dive
// chunks at once, so interactive requests keep at least half the slots and wait at most one chunk for the rest.
//
// A slot is a whf_ctx (its own HIP stream and working buffers) driven by its own host thread. Requests are cut
// into chunks of seeds (about chunk_checks (seed, region) checks each); an idle slot takes the next chunk of the
// request with the least work left (the oldest on ties), runs it with whf_scan and folds the chunk's hits and
// counters into the request. A request admitted while others run therefore starts when the first slot frees up, and the chunks of
// several requests run at once on the device. Each chunk is a plain whf_scan over a sub-range of the request's
// seeds, and counters add over chunks, so a request's hits and counters are those of whf_scan on it alone. A request
// with lysh's phase 2 runs its chunks as phase 1 only; the slot that completes the last chunk runs the hut verdicts
// of all the request's survivors in one batch, as whf_scan does at the end of a scan.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/whf.h"
#include "whf_serve.h"

// Device bytes one context holds (whf_scan.hip).
__attribute__((visibility("hidden"))) uint64_t whf_ctx_device_bytes(const whf_ctx *c);
namespace whf {
// lysh's phase 2 on a request's phase-1 survivors, in place (whf_scan.hip).
__attribute__((visibility("hidden"))) int phase2_filter(whf_ctx *c, const whf_request *req, whf_hit *hits, uint64_t n,
                                                        uint64_t *kept, uint64_t p2[5], double *ms);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    whf_request req{};
    whf_sink *sink = nullptr;
    uint64_t per_chunk = 1, next = 0, done_seeds = 0;   // seeds
    uint64_t regions = 1;
    int in_flight = 0;
    int status = WHF_OK;
    whf_stats stats{};
    std::vector<whf_hit> hits;
    double submit_s = 0, start_s = -1;
    uint64_t remaining_checks() const { return (req.n_seeds - next) * regions; }
};

struct DdJob {
    uint64_t id = 0;
    whf_dd_request req{};
    whf_dd_sink *spots = nullptr;
    whf_sink *finds = nullptr;
    whf_dd_stats *out = nullptr;
    uint64_t next = 0, done = 0;   // structure seeds
    int in_flight = 0;
    int status = WHF_OK;
    whf_dd_stats stats{};
    std::vector<whf_dd_spot> spot_list;
    std::vector<whf_hit> find_list;
    double submit_s = 0, start_s = -1;
};

constexpr uint64_t kDdChunk = 1ull << 18;   // structure seeds per Deep Dive chunk (about 20 ms of device time)
constexpr uint64_t kDdChunkSpots = 1ull << 16;
constexpr uint64_t kDdChunkFinds = 1ull << 14;

void add_dd_stats(whf_dd_stats &a, const whf_dd_stats &b) {
    a.structure_seeds += b.structure_seeds;
    a.tests += b.tests;
    a.starts += b.starts;
    for (int k = 0; k < 3; k++) a.walks[k] += b.walks[k];
    a.spots += b.spots;
    a.draws += b.draws;
    a.pieces += b.pieces;
    a.expanded_seeds += b.expanded_seeds;
    a.accepted += b.accepted;
    for (int k = 0; k < WHF_KIND_N; k++) a.funnel[k] += b.funnel[k];
    for (int k = 0; k < 8; k++) a.tiers[k] += b.tiers[k];
    for (int k = 0; k < 3; k++) a.samplers_built[k] += b.samplers_built[k];
    a.device_ms += b.device_ms;
    for (int k = 0; k < 3; k++) a.stage_ms[k] += b.stage_ms[k];
    a.biome_checked += b.biome_checked;
    a.biome_rejected += b.biome_rejected;
}

bool spot_less(const whf_dd_spot &x, const whf_dd_spot &y) {
    if (x.s48 != y.s48) return x.s48 < y.s48;
    if (x.rx != y.rx) return x.rx < y.rx;
    return x.rz < y.rz;
}

void add_stats(whf_stats &a, const whf_stats &b) {
    a.seeds += b.seeds;
    a.checks += b.checks;
    a.accepted += b.accepted;
    a.reached_ladder += b.reached_ladder;
    for (int k = 0; k < WHF_KIND_N; k++) a.funnel[k] += b.funnel[k];
    for (int k = 0; k < 8; k++) a.tiers[k] += b.tiers[k];
    for (int k = 0; k < 3; k++) a.samplers_built[k] += b.samplers_built[k];
    a.device_ms += b.device_ms;
    for (int k = 0; k < 8; k++) a.kernel_ms[k] += b.kernel_ms[k];
    a.fp64_rechecks += b.fp64_rechecks;
    for (int k = 0; k < 5; k++) a.phase2[k] += b.phase2[k];
    a.phase2_ms += b.phase2_ms;
}

bool hit_less(const whf_hit &x, const whf_hit &y) {
    if (x.seed != y.seed) return x.seed < y.seed;
    if (x.rx != y.rx) return x.rx < y.rx;
    return x.rz < y.rz;
}

constexpr uint64_t kChunkHits = 1u << 20;

}  // namespace

struct whf_server {
    std::vector<whf_ctx *> ctx;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable work_cv, event_cv;
    std::deque<std::shared_ptr<Job>> jobs;   // admitted requests with chunks left, in arrival order
    std::deque<std::shared_ptr<DdJob>> dd_jobs;   // background Deep Dive jobs with chunks left
    int dd_running = 0;                             // slots running a Deep Dive chunk now
    std::deque<whf_event> events;
    uint64_t chunk_checks = 1ull << 24;
    uint64_t next_id = 1;
    int in_flight_jobs = 0;
    bool stop = false;
    Clock::time_point epoch = Clock::now();
    whf::DqServer *dq = nullptr;   // the device-resident loop (default mode); null with WHF_SERVE_SLOTS
    std::mutex dd_open_mu;         // the device loop's Deep Dive slot is opened by the first background job
    int device = 0;

    // the device loop's completions
    static void post_event(void *p, const whf_event &e) {
        whf_server *s = static_cast<whf_server *>(p);
        std::lock_guard<std::mutex> lk(s->mu);
        s->events.push_back(e);
        s->in_flight_jobs--;
        s->event_cv.notify_all();
    }

    double now() const { return std::chrono::duration<double>(Clock::now() - epoch).count(); }

    void finish(const std::shared_ptr<Job> &j) {   // mu held
        std::sort(j->hits.begin(), j->hits.end(), hit_less);
        int status = j->status;
        if (j->sink) {
            j->sink->count = j->hits.size();
            const uint64_t m = std::min<uint64_t>(j->sink->capacity, j->hits.size());
            if (m && j->sink->hits) std::memcpy(j->sink->hits, j->hits.data(), m * sizeof(whf_hit));
            if (j->hits.size() > j->sink->capacity && status == WHF_OK) status = WHF_EOVERFLOW;
        }
        whf_event e{};
        e.id = j->id;
        e.status = status;
        e.stats = j->stats;
        e.stats.seeds = j->req.n_seeds;
        e.submit_s = j->submit_s;
        e.start_s = j->start_s;
        e.done_s = now();
        events.push_back(e);
        in_flight_jobs--;
        event_cv.notify_all();
    }

    void finish_dd(const std::shared_ptr<DdJob> &j) {   // mu held
        std::sort(j->spot_list.begin(), j->spot_list.end(), spot_less);
        std::sort(j->find_list.begin(), j->find_list.end(), hit_less);
        int status = j->status;
        if (j->spots) {
            j->spots->count = j->spot_list.size();
            const uint64_t m = std::min<uint64_t>(j->spots->capacity, j->spot_list.size());
            if (m && j->spots->spots) std::memcpy(j->spots->spots, j->spot_list.data(), m * sizeof(whf_dd_spot));
            if (j->spot_list.size() > j->spots->capacity && status == WHF_OK) status = WHF_EOVERFLOW;
        }
        if (j->finds) {
            j->finds->count = j->find_list.size();
            const uint64_t m = std::min<uint64_t>(j->finds->capacity, j->find_list.size());
            if (m && j->finds->hits) std::memcpy(j->finds->hits, j->find_list.data(), m * sizeof(whf_hit));
            if (j->find_list.size() > j->finds->capacity && status == WHF_OK) status = WHF_EOVERFLOW;
        }
        if (j->out) *j->out = j->stats;
        whf_event e{};
        e.id = j->id;
        e.status = status;
        e.submit_s = j->submit_s;
        e.start_s = j->start_s;
        e.done_s = now();
        events.push_back(e);
        in_flight_jobs--;
        event_cv.notify_all();
    }

    // Runs one Deep Dive chunk of job j on context c (mu not held).
    void run_dd_chunk(whf_ctx *c, const std::shared_ptr<DdJob> &j, uint64_t s0, uint64_t n,
                      std::vector<whf_dd_spot> &sbuf, std::vector<whf_hit> &fbuf) {
        whf_dd_request sub = j->req;
        sub.s_first = j->req.s_first + s0;
        sub.count = n;
        whf_dd_sink ss{sbuf.data(), sbuf.size(), 0};
        whf_sink fs{fbuf.data(), fbuf.size(), 0};
        whf_dd_stats st{};
        const int rc = whf_deepdive(c, &sub, &ss, &fs, &st);
        std::lock_guard<std::mutex> lk(mu);
        if (rc == WHF_OK || rc == WHF_EOVERFLOW) {
            add_dd_stats(j->stats, st);
            const uint64_t ms = std::min<uint64_t>(ss.count, sbuf.size()), mf = std::min<uint64_t>(fs.count, fbuf.size());
            j->spot_list.insert(j->spot_list.end(), sbuf.begin(), sbuf.begin() + ms);
            j->find_list.insert(j->find_list.end(), fbuf.begin(), fbuf.begin() + mf);
            if (rc == WHF_EOVERFLOW && j->status == WHF_OK) j->status = WHF_EOVERFLOW;
        } else if (j->status == WHF_OK || j->status == WHF_EOVERFLOW) {
            j->status = rc;
        }
        j->done += n;
        j->in_flight--;
        dd_running--;
        if (j->done >= j->req.count && j->in_flight == 0) finish_dd(j);
        work_cv.notify_all();
    }

    void worker(int slot) {
        whf_ctx *c = ctx[slot];
        std::vector<whf_hit> buf(kChunkHits);
        std::vector<whf_dd_spot> dd_spots;
        std::vector<whf_hit> dd_finds;
        const int dd_limit = dq ? 1 : std::max(1, (int)ctx.size() / 2);
        for (;;) {
            std::shared_ptr<Job> j;
            uint64_t s0 = 0, n = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                // wake for interactive work, for a Deep Dive chunk this slot may take, or to exit once stopped with
                // nothing queued; while Deep Dive chunks are queued but the other slots hold every Deep Dive place,
                // a stopping slot sleeps until one of them finishes its chunk (run_dd_chunk notifies)
                // with the device loop, a Deep Dive chunk also waits until the loop has no Seed Check work to claim,
                // which it does not signal here: such a slot looks again every millisecond
                auto ready = [&] {
                    return !jobs.empty() ||
                           (!dd_jobs.empty() && dd_running < dd_limit && !(dq && whf::dq_busy(dq))) ||
                           (stop && dd_jobs.empty());
                };
                if (dq) {
                    while (!ready()) work_cv.wait_for(lk, std::chrono::milliseconds(1));
                } else {
                    work_cv.wait(lk, ready);
                }
                if (jobs.empty()) {
                    if (!dd_jobs.empty() && dd_running < dd_limit) {
                        // no interactive work waiting: one chunk of the oldest background Deep Dive
                        std::shared_ptr<DdJob> d = dd_jobs.front();
                        const uint64_t d0 = d->next, dn = std::min(kDdChunk, d->req.count - d->next);
                        d->next += dn;
                        d->in_flight++;
                        dd_running++;
                        if (d->start_s < 0) d->start_s = now();
                        if (d->next >= d->req.count) dd_jobs.pop_front();
                        lk.unlock();
                        if (dd_spots.empty()) {
                            dd_spots.resize(kDdChunkSpots);
                            dd_finds.resize(kDdChunkFinds);
                        }
                        run_dd_chunk(c, d, d0, dn, dd_spots, dd_finds);
                        continue;
                    }
                    return;   // stopped with nothing queued (the wait's only other way out)
                }
                // shortest remaining work first (ties: the older request), so small requests overtake big ones
                auto best = jobs.begin();
                for (auto it = jobs.begin(); it != jobs.end(); ++it)
                    if ((*it)->remaining_checks() < (*best)->remaining_checks()) best = it;
                j = *best;
                s0 = j->next;
                n = std::min(j->per_chunk, j->req.n_seeds - s0);
                j->next += n;
                j->in_flight++;
                if (j->start_s < 0) j->start_s = now();
                if (j->next >= j->req.n_seeds) jobs.erase(best);
            }
            whf_request sub = j->req;
            sub.n_seeds = n;
            sub.phase2 = WHF_P2_NONE;
            if (sub.seed_mode == WHF_SEEDS_LIST) sub.seeds = j->req.seeds + s0;
            else sub.offset = j->req.offset + s0;
            whf_sink sink{buf.data(), kChunkHits, 0};
            whf_stats st{};
            const int rc = n ? whf_scan(c, &sub, &sink, &st) : WHF_OK;
            bool verdicts = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (rc == WHF_OK || rc == WHF_EOVERFLOW) {
                    add_stats(j->stats, st);
                    const uint64_t m = std::min<uint64_t>(sink.count, kChunkHits);
                    j->hits.insert(j->hits.end(), buf.begin(), buf.begin() + m);
                    if (rc == WHF_EOVERFLOW && j->status == WHF_OK) j->status = WHF_EOVERFLOW;
                } else if (j->status == WHF_OK || j->status == WHF_EOVERFLOW) {
                    j->status = rc;
                }
                j->done_seeds += n;
                j->in_flight--;
                if (j->done_seeds >= j->req.n_seeds && j->in_flight == 0) {
                    if (j->req.phase2 != WHF_P2_NONE && j->status == WHF_OK) {
                        verdicts = true;
                    } else {
                        // a phase-1 overflow leaves no complete survivor set to judge: as whf_scan, report
                        // WHF_EOVERFLOW with an empty sink rather than unjudged survivors
                        if (j->req.phase2 != WHF_P2_NONE) j->hits.clear();
                        finish(j);
                    }
                }
                work_cv.notify_all();
            }
            if (verdicts) {
                // no chunk of j is left or in flight, so this thread alone touches its hits now
                uint64_t kept = 0, p2[5] = {0, 0, 0, 0, 0};
                double ms = 0.0;
                const int rc2 = whf::phase2_filter(c, &j->req, j->hits.data(), j->hits.size(), &kept, p2, &ms);
                std::lock_guard<std::mutex> lk(mu);
                if (rc2 == WHF_OK) {
                    j->hits.resize(kept);
                    for (int k = 0; k < 5; k++) j->stats.phase2[k] += p2[k];
                    j->stats.phase2_ms += ms;
                } else {
                    j->status = rc2;
                }
                finish(j);
                work_cv.notify_all();
            }
        }
    }
};

extern "C" {

int whf_server_open(const whf_server_config *cfg, whf_server **out) {
    if (!out) return WHF_EINVAL;
    *out = nullptr;
    int slots = cfg && cfg->slots > 0 ? cfg->slots : 4;
    if (slots > 16) return WHF_EINVAL;
    whf_server *s = new (std::nothrow) whf_server();
    if (!s) return WHF_ENOMEM;
    if (cfg && cfg->chunk_checks) s->chunk_checks = cfg->chunk_checks;
    const uint64_t budget = cfg ? cfg->hbm_budget : 0;
    s->device = cfg ? cfg->device : 0;
    if (!(cfg && (cfg->flags & WHF_SERVE_SLOTS))) {
        whf::DqHooks hooks{s, &whf_server::post_event};
        const int rc = whf::dq_open(s->device, cfg ? cfg->slots : 0, cfg ? cfg->chunk_checks : 0, budget, hooks,
                                    s->epoch, &s->dq);
        if (rc) {
            delete s;
            return rc;
        }
        *out = s;
        return WHF_OK;
    }
    // one stream per slot: the slots' contexts already overlap on the device, which has few hardware queues to share
    whf_config cc{cfg ? cfg->device : 0, WHF_CFG_ONE_STREAM, 0};
    for (int i = 0; i < slots; i++) {
        whf_ctx *c = nullptr;
        int rc = whf_open(&cc, &c);
        if (rc) {
            for (whf_ctx *x : s->ctx) whf_close(x);
            delete s;
            return rc;
        }
        s->ctx.push_back(c);
        if (i == 0 && !(cfg && cfg->chunk_checks)) {
            // the device's tuned chunk size (whf_tuning.chunk_checks)
            whf_tuning t;
            if (whf_get_tuning(c, &t) == WHF_OK && t.chunk_checks) s->chunk_checks = t.chunk_checks;
        }
        // admission by memory: stop adding slots once the next one would exceed the budget
        if (budget && (uint64_t)(i + 2) * whf_ctx_device_bytes(c) > budget) break;
    }
    for (int i = 0; i < (int)s->ctx.size(); i++) s->threads.emplace_back(&whf_server::worker, s, i);
    *out = s;
    return WHF_OK;
}

void whf_server_close(whf_server *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->work_cv.notify_all();
    // Deep Dive slots first: they may wait for the device loop to go idle
    for (auto &t : s->threads) t.join();
    if (s->dq) whf::dq_close(s->dq);
    for (whf_ctx *c : s->ctx) whf_close(c);
    delete s;
}

int whf_submit(whf_server *s, const whf_request *req, whf_sink *sink, uint64_t *id) {
    if (!s || !req) return WHF_EINVAL;
    if (req->rx1 <= req->rx0 || req->rz1 <= req->rz0) return WHF_EINVAL;
    if (req->seed_mode == WHF_SEEDS_LIST && req->n_seeds && !req->seeds) return WHF_EINVAL;
    if (sink && sink->capacity && !sink->hits) return WHF_EINVAL;
    if (s->dq) {
        if (int rc = whf::dq_check(req)) return rc;
        uint64_t rid;
        double t;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop) return WHF_EINVAL;
            rid = s->next_id++;
            t = s->now();
            s->in_flight_jobs++;
        }
        if (id) *id = rid;
        if (req->n_seeds == 0) {
            if (sink) sink->count = 0;
            whf_event e{};
            e.id = rid;
            e.status = WHF_OK;
            e.submit_s = e.start_s = e.done_s = t;
            whf_server::post_event(s, e);
        } else {
            whf::dq_submit(s->dq, rid, req, sink, t);
        }
        return WHF_OK;
    }
    auto j = std::make_shared<Job>();
    j->req = *req;
    j->sink = sink;
    const uint64_t R = (uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return WHF_EINVAL;
    j->id = s->next_id++;
    j->regions = R;
    j->per_chunk = std::max<uint64_t>(1, s->chunk_checks / R);
    j->submit_s = s->now();
    if (id) *id = j->id;
    s->in_flight_jobs++;
    if (req->n_seeds == 0) {
        s->finish(j);
    } else {
        s->jobs.push_back(j);
        s->work_cv.notify_all();
    }
    return WHF_OK;
}

int whf_submit_deepdive(whf_server *s, const whf_dd_request *req, whf_dd_sink *spots, whf_sink *finds,
                        whf_dd_stats *stats, uint64_t *id) {
    if (!s || !req) return WHF_EINVAL;
    if (req->rx1 <= req->rx0 || req->rz1 <= req->rz0 || req->count > (1ull << 48)) return WHF_EINVAL;
    if (spots && spots->capacity && !spots->spots) return WHF_EINVAL;
    if (finds && finds->capacity && !finds->hits) return WHF_EINVAL;
    if (s->dq) {
        // the device loop's Deep Dive slot: one context and its thread, opened by the first background job
        std::lock_guard<std::mutex> olk(s->dd_open_mu);
        bool need = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop) return WHF_EINVAL;
            need = s->ctx.empty();
        }
        if (need) {
            whf_config cc{s->device, WHF_CFG_ONE_STREAM, 0};
            whf_ctx *c = nullptr;
            if (int rc = whf_open(&cc, &c)) return rc;
            std::lock_guard<std::mutex> lk(s->mu);
            s->ctx.push_back(c);
            s->threads.emplace_back(&whf_server::worker, s, 0);
        }
    }
    auto j = std::make_shared<DdJob>();
    j->req = *req;
    j->spots = spots;
    j->finds = finds;
    j->out = stats;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return WHF_EINVAL;
    j->id = s->next_id++;
    j->submit_s = s->now();
    if (id) *id = j->id;
    s->in_flight_jobs++;
    if (req->count == 0) {
        s->finish_dd(j);
    } else {
        s->dd_jobs.push_back(j);
        s->work_cv.notify_all();
    }
    return WHF_OK;
}

int whf_poll(whf_server *s, whf_event *ev, int max_events, int timeout_ms) {
    if (!s || (max_events > 0 && !ev) || max_events < 0) return WHF_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->events.empty(); };
    if (timeout_ms < 0) s->event_cv.wait(lk, ready);
    else s->event_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    int n = 0;
    while (n < max_events && !s->events.empty()) {
        ev[n++] = s->events.front();
        s->events.pop_front();
    }
    return n;
}

}  // extern "C"
