/*
 * server.cpp -- the device-resident step server of arc_server.h.
 *
 * One arc_sim holds the worlds of every admitted client, sized once by the
 * HBM budget.  Clients get contiguous world ranges (first fit).  Requests
 * queue per client; a dispatcher thread forms each batch from the oldest
 * pending request of every client, copies host targets into the pinned
 * staging rows of their worlds and uploads those rows (targets a client
 * already holds in device memory are read where they lie: the batch table
 * carries each request's row address), and replays the captured
 * batch graph (step + fall rules + auto-reset of the batch's worlds) without
 * waiting for it; a completer thread waits for each batch's event in launch
 * order and stamps its requests done.  Two batches may be in flight, each
 * with its own targets buffer, so uploads overlap the running step.
 *
 * When a batch starts: at once when every open client has a request pending;
 * while a batch runs, shortly before it is expected to end (its expected
 * length is the running mean of the recent batches' device time), so that
 * the requests arriving meanwhile are stepped together next and the device
 * does not idle between batches; with none running, when requests stop
 * arriving for a short gap or at the latest a window after the oldest.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <limits.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/arc_server.h"
#include "sim_internal.h"

namespace {

using clock_type = std::chrono::steady_clock;

struct Request {
    long long ticket;
    int client;
    int start = 0, count = 0;     /* the client's worlds, copied when the batch is formed */
    std::vector<float> targets;   /* host targets (arcs_submit), or */
    const float *dev = nullptr;   /* the client's device targets (arcs_submit_device) */
    clock_type::time_point submitted;
};

struct Client {
    int start = 0, count = 0;
    bool open = false;
    std::deque<Request> queue;
};

constexpr int DEPTH = 2; /* batches in flight on the stream (double-buffered targets) */

/* a launched batch, until its completion event has fired */
struct Flight {
    std::vector<Request> reqs;
    int worlds = 0;
    int buf = 0;
    bool ok = true;
    clock_type::time_point launched;
};

} /* namespace */

struct arc_server {
    arc_sim *sim = nullptr;
    int capacity = 0, used = 0;
    float *d_targets[DEPTH] = {}; /* [capacity][12] each */
    float *h_targets[DEPTH] = {}; /* pinned staging, same layout */
    hipEvent_t ev[DEPTH] = {};     /* batch done (sim stream) */
    hipEvent_t copied[DEPTH] = {}; /* its targets uploaded (copy stream) */
    hipStream_t copy = nullptr;    /* uploads overlap the running step */
    std::vector<std::vector<float>> pool; /* request buffers for reuse (under mu) */
    std::deque<Flight> flights;     /* launched, in stream order (under mu) */
    std::condition_variable landed; /* a flight was launched or completed */
    std::thread completer;
    long long launched = 0;
    /* admitted clients by id; ids are issued once and never reused, so a
     * caller's id cannot come to name another client after a close */
    std::map<int, Client> clients;
    int next_client = 0;
    std::vector<std::pair<int, int>> free_ranges; /* (start, count), sorted by start */

    std::mutex mu;                 /* queues, clients, tickets, stats */
    std::condition_variable work;  /* dispatcher: requests pending or stopping */
    std::condition_variable done;  /* waiters: a batch completed */
    std::mutex sim_mu;             /* the sim and its stream */
    std::thread dispatcher;
    bool stop = false;
    /* with no batch in flight a batch waits at most this long after its first
     * request for the other clients' requests, and no longer than `gap` after
     * the last request that arrived; it starts at once when every open client
     * has one */
    std::chrono::microseconds window{ 500 };
    std::chrono::microseconds gap{ 100 };
    /* while a batch runs, the next is launched this long before the running
     * one is expected to end (the time to form and launch a batch) */
    std::chrono::microseconds lead{ 60 };
    clock_type::time_point last_arrival = clock_type::time_point::min();
    clock_type::time_point last_done = clock_type::time_point::min();
    double batch_us = 0.0; /* running mean of the batches' device time */
    long long next_ticket = 1;
    std::map<long long, double> finished; /* ticket -> latency (ms), until waited */
    std::map<long long, int> inflight;    /* ticket -> client, queued or running */

    /* statistics since the last reset */
    std::vector<double> lat;
    long long batches = 0, world_steps = 0, requests = 0;
    clock_type::time_point since = clock_type::now();
};

static void release_range(arc_server *sv, int start, int count)
{
    auto &f = sv->free_ranges;
    f.emplace_back(start, count);
    std::sort(f.begin(), f.end());
    std::vector<std::pair<int, int>> m;
    for (auto &r : f) {
        if (!m.empty() && m.back().first + m.back().second == r.first)
            m.back().second += r.second;
        else
            m.push_back(r);
    }
    f.swap(m);
}

/* Forms batches and launches them; up to DEPTH batches are in flight, each
 * with its own targets buffer, so the next batch is uploaded and queued
 * while the current one runs. */
static void dispatch_loop(arc_server *sv)
{
    std::vector<int> starts, counts;
    std::vector<const float *> rows;
    for (;;) {
        Flight f;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            auto pending = [&](bool all) {
                int open = 0, ready = 0;
                for (auto &kv : sv->clients) {
                    open += kv.second.open || !kv.second.queue.empty();
                    ready += !kv.second.queue.empty();
                }
                return all ? ready > 0 && ready == open : ready > 0;
            };
            sv->work.wait(lk, [&] { return sv->stop || pending(false); });
            /* gather requests until the batch's start (re-evaluated whenever a
             * request arrives or a batch ends) */
            while (!sv->stop && !pending(true)) {
                clock_type::time_point when;
                if (sv->flights.empty()) {
                    clock_type::time_point first = clock_type::time_point::max();
                    for (auto &kv : sv->clients)
                        if (!kv.second.queue.empty())
                            first = std::min(first, kv.second.queue.front().submitted);
                    when = std::min(first + sv->window, sv->last_arrival + sv->gap);
                } else {
                    /* expected end of the last flight: each starts when it was
                     * launched or when the one before it ends */
                    const auto len = std::chrono::microseconds((long long)sv->batch_us);
                    clock_type::time_point end = sv->last_done;
                    for (const Flight &fl : sv->flights)
                        end = std::max(fl.launched, end) + len;
                    when = end - sv->lead;
                }
                if (clock_type::now() >= when)
                    break;
                const size_t nf = sv->flights.size();
                sv->work.wait_until(lk, when, [&] { return sv->stop || pending(true) || sv->flights.size() != nf; });
            }
            if (sv->stop && !pending(false))
                break;
            /* a free targets buffer: at most DEPTH flights */
            sv->landed.wait(lk, [&] { return (int)sv->flights.size() < DEPTH; });
            for (auto &kv : sv->clients)
                if (!kv.second.queue.empty()) {
                    Client &c = kv.second;
                    f.reqs.push_back(std::move(c.queue.front()));
                    c.queue.pop_front();
                    /* the range as it is now: a client's range is fixed while it has
                     * requests queued or in flight */
                    f.reqs.back().start = c.start;
                    f.reqs.back().count = c.count;
                }
            f.buf = (int)(sv->launched % DEPTH);
        }
        starts.clear();
        counts.clear();
        rows.clear();
        {
            std::lock_guard<std::mutex> sl(sv->sim_mu);
            arc_sim *s = sv->sim;
            float *h = sv->h_targets[f.buf], *d = sv->d_targets[f.buf];
            bool uploads = false;
            for (auto &r : f.reqs) {
                const size_t off = (size_t)r.start * NJ, bytes = (size_t)r.count * NJ * sizeof(float);
                if (r.dev) {
                    rows.push_back(r.dev); /* read where the client holds it */
                } else {
                    memcpy(h + off, r.targets.data(), bytes);
                    f.ok = f.ok && hipMemcpyAsync(d + off, h + off, bytes, hipMemcpyHostToDevice, sv->copy)
                                       == hipSuccess;
                    rows.push_back(d + off);
                    uploads = true;
                }
                starts.push_back(r.start);
                counts.push_back(r.count);
                f.worlds += r.count;
            }
            if (uploads)
                f.ok = f.ok && hipEventRecord(sv->copied[f.buf], sv->copy) == hipSuccess
                       && hipStreamWaitEvent(s->stream, sv->copied[f.buf], 0) == hipSuccess;
            /* rows: the staging buffers, or device targets checked at submission */
            f.ok = f.ok
                   && step_batch_rows_checked(s, starts.data(), counts.data(), (int)f.reqs.size(), rows.data(), 1, 1)
                          == 0;
            f.ok = f.ok && hipEventRecord(sv->ev[f.buf], s->stream) == hipSuccess;
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            for (auto &r : f.reqs)
                if (!r.dev)
                    sv->pool.push_back(std::move(r.targets)); /* copied into the staging */
            f.launched = clock_type::now();
            sv->flights.push_back(std::move(f));
            sv->launched++;
        }
        sv->landed.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->launched = -1; /* no more flights */
    }
    sv->landed.notify_all();
}

/* Waits for the flights in launch order and stamps their requests done. */
static void complete_loop(arc_server *sv)
{
    for (;;) {
        int buf;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->landed.wait(lk, [&] { return !sv->flights.empty() || sv->launched < 0; });
            if (sv->flights.empty())
                return;
            buf = sv->flights.front().buf;
        }
        /* the pinned staging of this buffer is reused only after the event */
        bool ok = hipEventSynchronize(sv->ev[buf]) == hipSuccess;
        const auto t = clock_type::now();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            Flight &f = sv->flights.front();
            ok = ok && f.ok;
            for (auto &r : f.reqs) {
                double ms = std::chrono::duration<double, std::milli>(t - r.submitted).count();
                sv->finished[r.ticket] = ok ? ms : -1.0;
                sv->inflight.erase(r.ticket);
                sv->lat.push_back(ms);
            }
            sv->batches++;
            sv->requests += (long long)f.reqs.size();
            sv->world_steps += f.worlds;
            /* the batch's device time: from its launch, or from the end of the
             * batch before it when it was queued behind that one */
            const double us = std::chrono::duration<double, std::micro>(t - std::max(f.launched, sv->last_done)).count();
            sv->batch_us = sv->batch_us > 0.0 ? 0.8 * sv->batch_us + 0.2 * us : us;
            sv->last_done = t;
            sv->flights.pop_front();
        }
        if (!ok)
            fprintf(stderr, "libarcsim: step server: batch failed (%s)\n", arcx_error_message(sv->sim));
        sv->done.notify_all();
        sv->landed.notify_all();
        sv->work.notify_all();
    }
}

extern "C" {

arc_server *arcs_create(double hbm_budget_bytes, double dt, int substeps, int grid, double cell)
{
    /* bytes per world of this configuration, from a one-world sim */
    arc_sim *probe = arc_sim_create(1, dt, substeps, grid, cell);
    if (!probe)
        return nullptr;
    const double per = arcx_bytes_per_world(probe) + NJ * sizeof(float) * 2;
    arc_sim_destroy(probe);
    const double cap = floor(hbm_budget_bytes / per);
    if (!(cap >= 1.0)) {
        fprintf(stderr, "libarcsim: arcs_create: the budget holds no world (%.0f bytes per world)\n", per);
        return nullptr;
    }
    arc_server *sv = new (std::nothrow) arc_server();
    if (!sv)
        return nullptr;
    sv->capacity = (int)std::min(cap, 16777216.0);
    sv->sim = arc_sim_create(sv->capacity, dt, substeps, grid, cell);
    if (!sv->sim) {
        delete sv;
        return nullptr;
    }
    const size_t nt = (size_t)sv->capacity * NJ;
    for (int k = 0; k < DEPTH; ++k)
        if (hipMalloc((void **)&sv->d_targets[k], nt * sizeof(float)) != hipSuccess
            || hipHostMalloc((void **)&sv->h_targets[k], nt * sizeof(float), hipHostMallocDefault) != hipSuccess
            || hipEventCreateWithFlags(&sv->ev[k], hipEventDisableTiming) != hipSuccess
            || hipEventCreateWithFlags(&sv->copied[k], hipEventDisableTiming) != hipSuccess
            || (!sv->copy && hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking) != hipSuccess)) {
            fprintf(stderr, "libarcsim: arcs_create: staging allocation failed\n");
            arcs_destroy(sv);
            return nullptr;
        }
    sv->free_ranges.emplace_back(0, sv->capacity);
    sv->dispatcher = std::thread(dispatch_loop, sv);
    sv->completer = std::thread(complete_loop, sv);
    return sv;
}

void arcs_destroy(arc_server *sv)
{
    if (!sv)
        return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->work.notify_all();
    if (sv->dispatcher.joinable())
        sv->dispatcher.join();
    else {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->launched = -1;
    }
    sv->landed.notify_all();
    if (sv->completer.joinable())
        sv->completer.join();
    for (int k = 0; k < DEPTH; ++k) {
        if (sv->d_targets[k])
            (void)hipFree(sv->d_targets[k]);
        if (sv->h_targets[k])
            (void)hipHostFree(sv->h_targets[k]);
        if (sv->ev[k])
            (void)hipEventDestroy(sv->ev[k]);
        if (sv->copied[k])
            (void)hipEventDestroy(sv->copied[k]);
    }
    if (sv->copy)
        (void)hipStreamDestroy(sv->copy);
    arc_sim_destroy(sv->sim);
    delete sv;
}

int arcs_capacity(const arc_server *sv) { return sv->capacity; }

int arcs_used(const arc_server *cs)
{
    arc_server *sv = const_cast<arc_server *>(cs);
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->used;
}

int arcs_open(arc_server *sv, int n_worlds, unsigned int seed0, double difficulty)
{
    if (n_worlds <= 0 || difficulty > MAX_DIFFICULTY) /* beyond the device terrain's range */
        return -1;
    int id = -1, start = -1;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->used + n_worlds > sv->capacity || sv->next_client == INT_MAX)
            return -1;
        for (auto &r : sv->free_ranges)
            if (r.second >= n_worlds) {
                start = r.first;
                r.first += n_worlds;
                r.second -= n_worlds;
                break;
            }
        if (start < 0)
            return -1;
        sv->free_ranges.erase(std::remove_if(sv->free_ranges.begin(), sv->free_ranges.end(),
                                             [](const std::pair<int, int> &r) { return r.second == 0; }),
                              sv->free_ranges.end());
        id = sv->next_client++;
        Client &c = sv->clients[id];
        c.start = start;
        c.count = n_worlds;
        c.open = true;
        sv->used += n_worlds;
    }
    /* reset the client's worlds and clear their task state */
    std::vector<int> ids(n_worlds);
    std::vector<unsigned> seeds(n_worlds);
    std::vector<double> diff(n_worlds);
    std::vector<unsigned> u;
    if (difficulty < 0.0) {
        u.resize(n_worlds);
        arcx_rng_stream(seed0, n_worlds, u.data());
    }
    for (int k = 0; k < n_worlds; ++k) {
        ids[k] = start + k;
        seeds[k] = seed0 + (unsigned)k;
        diff[k] = difficulty < 0.0 ? (double)(u[k] >> 8) * (1.0 / 16777216.0) : difficulty;
    }
    {
        std::lock_guard<std::mutex> sl(sv->sim_mu);
        arc_sim *s = sv->sim;
        (void)launch_fill_int(s->ts.torso_steps + start, 0, n_worlds, s->stream);
        (void)launch_fill_int(s->ts.cause + start, 0, n_worlds, s->stream);
        (void)launch_fill_int((int *)s->ts.episode + start, 0, n_worlds, s->stream);
        arc_sim_reset(s, ids.data(), n_worlds, seeds.data(), diff.data());
        if (arcx_last_error(s)) {
            fprintf(stderr, "libarcsim: arcs_open: %s\n", arcx_error_message(s));
            arcx_clear_error(s);
        }
    }
    return id;
}

int arcs_close(arc_server *sv, int client)
{
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->clients.find(client);
    if (it == sv->clients.end() || !it->second.open)
        return -1;
    Client &c = it->second; /* map nodes stay put while other clients come and go */
    c.open = false;         /* no new submissions */
    sv->done.wait(lk, [&] {
        if (!c.queue.empty())
            return false;
        for (auto &kv : sv->inflight)
            if (kv.second == client)
                return false;
        return true;
    });
    release_range(sv, c.start, c.count);
    sv->used -= c.count;
    sv->clients.erase(it);
    return 0;
}

int arcs_client_range(const arc_server *cs, int client, int *start, int *count)
{
    arc_server *sv = const_cast<arc_server *>(cs);
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->clients.find(client);
    if (it == sv->clients.end() || !it->second.open)
        return -1;
    *start = it->second.start;
    *count = it->second.count;
    return 0;
}

long long arcs_submit(arc_server *sv, int client, const float *targets)
{
    const auto t = clock_type::now();
    long long ticket;
    Request r;
    size_t nf;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->clients.find(client);
        if (it == sv->clients.end() || !it->second.open || !targets)
            return -1;
        nf = (size_t)it->second.count * NJ;
        if (!sv->pool.empty()) {
            r.targets = std::move(sv->pool.back());
            sv->pool.pop_back();
        }
    }
    /* the copy runs outside the lock, into a recycled buffer */
    r.targets.assign(targets, targets + nf);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->clients.find(client); /* closed meanwhile? */
        if (it == sv->clients.end() || !it->second.open || (size_t)it->second.count * NJ != nf) {
            sv->pool.push_back(std::move(r.targets));
            return -1;
        }
        Client &c = it->second;
        r.ticket = ticket = sv->next_ticket++;
        r.client = client;
        r.submitted = t;
        sv->last_arrival = std::max(sv->last_arrival, t);
        sv->inflight[ticket] = client;
        c.queue.push_back(std::move(r));
    }
    sv->work.notify_one();
    return ticket;
}

long long arcs_submit_device(arc_server *sv, int client, const float *targets_dev)
{
    const auto t = clock_type::now();
    long long ticket;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->clients.find(client);
        if (it == sv->clients.end() || !it->second.open || !targets_dev)
            return -1;
        /* the step kernels will read the client's n x 12 floats here: they must
         * lie in memory of the server's device (a host address would fault
         * the launch) */
        if (!device_memory(sv->sim, targets_dev, (size_t)it->second.count * NJ * sizeof(float)))
            return -1;
        Request r;
        r.dev = targets_dev;
        r.ticket = ticket = sv->next_ticket++;
        r.client = client;
        r.submitted = t;
        sv->last_arrival = std::max(sv->last_arrival, t);
        sv->inflight[ticket] = client;
        it->second.queue.push_back(std::move(r));
    }
    sv->work.notify_one();
    return ticket;
}

double arcs_wait(arc_server *sv, long long ticket)
{
    std::unique_lock<std::mutex> lk(sv->mu);
    if (!sv->finished.count(ticket) && !sv->inflight.count(ticket))
        return -1.0;
    sv->done.wait(lk, [&] { return sv->finished.count(ticket) != 0; });
    double ms = sv->finished[ticket];
    sv->finished.erase(ticket);
    return ms;
}

int arcs_stats(const arc_server *cs, double *out)
{
    arc_server *sv = const_cast<arc_server *>(cs);
    std::lock_guard<std::mutex> lk(sv->mu);
    std::vector<double> v = sv->lat;
    auto pct = [&](double p) {
        if (v.empty())
            return 0.0;
        size_t k = (size_t)floor(p * (double)(v.size() - 1) + 0.5);
        std::nth_element(v.begin(), v.begin() + k, v.end());
        return v[k];
    };
    const double el = std::chrono::duration<double>(clock_type::now() - sv->since).count();
    out[0] = (double)sv->requests;
    out[1] = (double)sv->batches;
    out[2] = (double)sv->world_steps;
    out[3] = sv->batches ? (double)sv->world_steps / (double)sv->batches : 0.0;
    out[4] = pct(0.50);
    out[5] = pct(0.99);
    out[6] = v.empty() ? 0.0 : *std::max_element(v.begin(), v.end());
    out[7] = el;
    out[8] = el > 0.0 ? (double)sv->world_steps / el : 0.0;
    out[9] = sv->batches ? (double)sv->requests / (double)sv->batches : 0.0;
    return 0;
}

int arcs_autotune(arc_server *sv, int steps, int *out5, double *ms)
{
    /* no batch is launched meanwhile; the tuner's copies follow the batches
     * already on the stream, and it restores the worlds it times */
    std::lock_guard<std::mutex> sl(sv->sim_mu);
    int rc = arcx_autotune(sv->sim, steps, out5, ms);
    if (rc) {
        fprintf(stderr, "libarcsim: arcs_autotune: %s\n", arcx_error_message(sv->sim));
        arcx_clear_error(sv->sim);
    }
    return rc;
}

int arcs_set_window(arc_server *sv, double microseconds)
{
    if (!(microseconds >= 0.0))
        return -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->window = std::chrono::microseconds((long long)microseconds);
    return 0;
}

void arcs_reset_stats(arc_server *sv)
{
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->lat.clear();
    sv->batches = sv->world_steps = sv->requests = 0;
    sv->since = clock_type::now();
}

int arcs_observe(arc_server *sv, int client, double *root)
{
    int start, count;
    if (arcs_client_range(sv, client, &start, &count))
        return -1;
    std::lock_guard<std::mutex> sl(sv->sim_mu);
    arc_sim *s = sv->sim;
    std::vector<double> all((size_t)s->n * 13);
    arc_sim_observe(s, all.data(), nullptr, nullptr, nullptr, nullptr);
    if (arcx_last_error(s))
        return -1;
    memcpy(root, all.data() + (size_t)start * 13, (size_t)count * 13 * sizeof(double));
    return 0;
}

} /* extern "C" */
