// whf_serve.h — the device-resident serving loop (whf_serve.hip) as whf_server.cpp drives it.
#pragma once

#include <chrono>
#include <cstdint>

#include "../../include/whf.h"

namespace whf {

struct DqServer;

// How the loop hands completions back: post(ctx, event) for every request, once, after its sink is filled.
struct DqHooks {
    void *ctx;
    void (*post)(void *ctx, const whf_event &e);
};

// Opens the loop on a device: its context, lanes, request table, pools and threads. lanes: streams that run whole
// batches side by side (3 or 4; fewer give 3); item_cap: (seed, region) checks per batch (0: the most a batch holds);
// hbm_budget: device bytes the server may hold (0: no limit; the seed-list pool is cut to fit, WHF_ENOMEM below the
// fixed part). Times of the events are seconds since epoch.
int dq_open(int device, int lanes, uint64_t item_cap, uint64_t hbm_budget, const DqHooks &hooks,
            std::chrono::steady_clock::time_point epoch, DqServer **out);
// WHF_OK when the loop can serve the request (at most 2^16 regions, fewer than 2^32 seeds, a seed list that fits the
// seed pool), else WHF_EINVAL.
int dq_check(const whf_request *req);
// Queues a checked request with at least one seed; it is admitted when a table entry and pool space are free.
void dq_submit(DqServer *s, uint64_t id, const whf_request *req, whf_sink *sink, double submit_s);
// Whether Seed Check work is queued or left unclaimed on the device (background work waits meanwhile).
bool dq_busy(DqServer *s);
// Finishes every queued and admitted request, then releases the loop.
void dq_close(DqServer *s);

}  // namespace whf
