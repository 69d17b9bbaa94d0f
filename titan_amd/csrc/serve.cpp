// Serving layer: continuous batching of rollout requests over the C ABI.
//
// Requests are bodies with a horizon.  The host loop calls titan_serve_pump;
// each pump admits queued requests in arrival order while they fit the
// running-set limits, steps all running requests together through
// titan_step_bodies (one launch per step covers every body; each body stops
// at its own horizon), and retires the finished ones.  Layouts are built
// on the host by background worker threads from the moment a request is
// submitted (in submission order, under the schedule current at its
// submission), so that admission places bodies already laid out while the
// device runs the quantum just enqueued (it waits for a layout only when
// nothing runs, and passes over requests whose layouts are not built yet);
// placing a body on the device does not wait for the launches in flight.  The engine keeps every
// body's arithmetic independent of the others, so a request's result is the
// same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_internal.h"
#include "titan_mi355x.h"

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

// One array of a body to copy into a request's block.
struct CopyJob {
    void *dst;
    const void *src;
    size_t bytes;
};

// Copy the jobs' bytes, split into `parts` equal spans of their
// concatenation: the caller's thread copies the first, one thread each the
// others (a span runs on the caller's thread when no thread can be made).
// The destination pages are first touched by these copies, so that page
// faults are taken by several threads at once.
void copy_spans(const std::vector<CopyJob> &jobs, size_t total, unsigned parts) {
    auto span = [&jobs](size_t lo, size_t hi) {
        size_t base = 0;
        for (const CopyJob &j : jobs) {
            const size_t a = std::max(lo, base), b = std::min(hi, base + j.bytes);
            if (a < b) std::memcpy(static_cast<char *>(j.dst) + (a - base), static_cast<const char *>(j.src) + (a - base), b - a);
            base += j.bytes;
            if (base >= hi) break;
        }
    };
    std::vector<std::thread> th;
    for (unsigned p = 1; p < parts; ++p) {
        const size_t lo = total * p / parts, hi = total * (p + 1) / parts;
        try {
            th.emplace_back(span, lo, hi);
        } catch (const std::exception &) {
            span(lo, hi);
        }
    }
    span(0, total / parts);
    for (auto &t : th) t.join();
}

// A request with its own copy of the body's arrays.
struct Request {
    int64_t ticket = 0;
    int32_t state = TITAN_REQ_QUEUED;
    int64_t n = 0, s = 0, horizon = 0, done = 0;
    int32_t body = -1;
    double t_submit = 0, t_done = 0;
    uint64_t est_bytes = 0;
    double *out_pos = nullptr, *out_vel = nullptr;
    titan_body_desc desc{};
    titan_mi355x::Prepared *prep = nullptr;  // laid out on the host, not yet placed
    int prep_rc = TITAN_OK;
    std::string prep_err;
    // layout by a worker: 0 none yet, 1 in progress, 2 done (prep / prep_rc
    // set); guarded by the server's prep_mu
    int prep_state = 0;
    titan_schedule sched{};  // the context's schedule when submitted

    ~Request() { titan_mi355x::free_prepared(prep); }
    // the body's arrays, copied at submission into one block (64-byte
    // aligned pieces) that desc points into
    std::unique_ptr<unsigned char[]> block;

    // Point `field` at the next piece of the block (with `jobs`, after the
    // block is allocated) for `cnt` elements of src; a null or empty array
    // stays null.
    template <class T>
    void piece(const T *src, int64_t cnt, const T *&field, size_t &off, std::vector<CopyJob> *jobs) {
        field = nullptr;
        if (!src || cnt <= 0) return;
        const size_t bytes = (size_t)cnt * sizeof(T);
        if (jobs) {
            T *dst = reinterpret_cast<T *>(block.get() + off);
            jobs->push_back({dst, src, bytes});
            field = dst;
        }
        off += (bytes + 63) & ~size_t(63);
    }
    size_t pieces(const titan_body_desc &d, std::vector<CopyJob> *jobs) {
        const int64_t nl = d.lc_ptr ? d.lc_ptr[n] : 0;
        size_t off = 0;
        piece(d.pos, 3 * n, desc.pos, off, jobs);
        piece(d.vel, 3 * n, desc.vel, off, jobs);
        piece(d.m, n, desc.m, off, jobs);
        piece(d.T, n, desc.T, off, jobs);
        piece(d.ext, 3 * n, desc.ext, off, jobs);
        piece(d.fixed, n, desc.fixed, off, jobs);
        piece(d.drag, n, desc.drag, off, jobs);
        piece(d.left, s, desc.left, off, jobs);
        piece(d.right, s, desc.right, off, jobs);
        piece(d.k, s, desc.k, off, jobs);
        piece(d.rest, s, desc.rest, off, jobs);
        piece(d.damping, s, desc.damping, off, jobs);
        piece(d.type, s, desc.type, off, jobs);
        piece(d.omega, s, desc.omega, off, jobs);
        piece(d.act_a, s, desc.act_a, off, jobs);
        piece(d.act_phi, s, desc.act_phi, off, jobs);
        piece(d.lc_ptr, d.lc_ptr ? n + 1 : 0, desc.lc_ptr, off, jobs);
        piece(d.lc_type, nl, desc.lc_type, off, jobs);
        piece(d.lc_vec, 3 * nl, desc.lc_vec, off, jobs);
        piece(d.lc_num, nl, desc.lc_num, off, jobs);
        return off;
    }

    // Copy the body's arrays; a body of 4 MB or more is copied by one
    // thread per 2 MB, up to four: copying into fresh pages runs at the rate
    // the kernel maps them, about 10 GB/s on one thread, and the caller's
    // thread is the one that also pumps the server.
    void copy_from(const titan_body_desc &d) {
        n = d.n_masses;
        s = d.n_springs;
        desc = d;
        const size_t size = pieces(d, nullptr);
        block.reset(new unsigned char[std::max<size_t>(size, 1)]);
        std::vector<CopyJob> jobs;
        pieces(d, &jobs);
        size_t total = 0;
        for (const CopyJob &j : jobs) total += j.bytes;
        const size_t part = size_t(2) << 20;
        copy_spans(jobs, total, (unsigned)std::min<size_t>(4, std::max<size_t>(1, total / part)));
    }

    void release_arrays() {
        titan_mi355x::free_prepared(prep);
        prep = nullptr;
        block.reset();
    }
};

// Device bytes a body needs, estimated from its size: state, side arrays and
// incidences with their SELL padding and tile halos, rounded up.
uint64_t estimate_bytes(int64_t n, int64_t s) {
    return (uint64_t)(256 * n + 96 * s + (1 << 16));
}

unsigned host_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hw ? hw : 1u));
}

// A request whose last steps are enqueued and whose result nobody reads: it
// completes when the step stream passes its event.
struct Finishing {
    Request *r;
    hipEvent_t ev;
};

struct Server {
    // background layouts: requests waiting for one, in submission order, and
    // the worker threads taking them (started with the first submission,
    // stopped when the context is destroyed)
    std::mutex prep_mu;
    std::condition_variable prep_cv;
    std::deque<Request *> todo;
    std::vector<std::thread> workers;
    bool stop = false;

    int32_t max_running = 4096;
    uint64_t hbm_bytes = 0;
    int64_t quantum = 256;
    // TITAN_SERVE_ADMIT=wait (read when the server is made): admission waits
    // for the layout of every request it picks and places every request that
    // fits, so that which requests a pump admits depends neither on the
    // workers' timing nor on the device's (tests that count pumps use it);
    // by default it admits only layouts already built, and only until the
    // device has run the older of its two quanta
    bool admit_waits = false;
    // layout worker threads (TITAN_SERVE_WORKERS, read when the server is
    // made; default: one per host thread but four, which are left to the
    // caller's thread and the three that help it copy a submitted body)
    unsigned n_workers = 1;
    int64_t next_ticket = 1;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request *> queue;
    std::vector<Request *> running;
    std::deque<Finishing> finishing;
    std::vector<hipEvent_t> spare;  // events for reuse
    // events recorded after the two latest quanta (q_prev, then q_last): a
    // pump enqueues its quantum only once q_prev has passed, so that at
    // most two quanta are in the step stream and the next is enqueued
    // while the last one runs
    hipEvent_t q_prev = nullptr, q_last = nullptr;
    uint64_t running_bytes = 0;
    // statistics
    int64_t submitted = 0, completed = 0, cancelled = 0, failed = 0;
    std::vector<double> latencies;
    double mass_steps = 0, first_submit = -1, last_done = -1;
    // device idle time between quanta: timed events recorded on the step
    // stream after each quantum (last_end) and before the next one; a pair
    // (previous end, next begin) is read once the begin has passed
    std::vector<hipEvent_t> tspare;
    hipEvent_t last_end = nullptr;
    std::deque<std::pair<hipEvent_t, hipEvent_t>> gaps;
    double idle_s = 0;

    void worker() {
        for (;;) {
            Request *r;
            {
                std::unique_lock<std::mutex> lk(prep_mu);
                prep_cv.wait(lk, [&] { return stop || !todo.empty(); });
                if (stop) return;
                r = todo.front();
                todo.pop_front();
                r->prep_state = 1;
            }
            int rc = TITAN_OK;
            std::string err;
            titan_mi355x::Prepared *p = titan_mi355x::prepare_body_sched(r->sched, &r->desc, &rc, &err);
            {
                std::lock_guard<std::mutex> lk(prep_mu);
                r->prep = p;
                r->prep_rc = rc;
                r->prep_err.swap(err);
                r->prep_state = 2;
            }
            prep_cv.notify_all();
        }
    }
    void start_workers(unsigned n) {
        for (unsigned t = 0; t < n; ++t) {
            try {
                workers.emplace_back([this] { worker(); });
            } catch (const std::exception &) {  // no more threads: admission lays out the rest itself
                break;
            }
        }
    }
    void stop_workers() {
        {
            std::lock_guard<std::mutex> lk(prep_mu);
            stop = true;
        }
        prep_cv.notify_all();
        for (auto &t : workers) t.join();
        workers.clear();
    }
    ~Server() { stop_workers(); }  // a context never destroyed: its workers end with the process
    // Take r out of the workers' hands: unclaimed, it leaves the to-do list;
    // claimed, wait for its layout.  Afterwards no worker touches r.
    void settle(Request *r) {
        std::unique_lock<std::mutex> lk(prep_mu);
        if (r->prep_state == 0) {
            auto q = std::find(todo.begin(), todo.end(), r);
            if (q != todo.end()) todo.erase(q);
            return;
        }
        prep_cv.wait(lk, [&] { return r->prep_state == 2; });
    }

    // Whether r's layout is built (with no worker thread, admission lays it
    // out itself, so it counts as ready).
    bool layout_ready(Request *r) {
        std::lock_guard<std::mutex> lk(prep_mu);
        return r->prep_state == 2 || workers.empty();
    }

    // Wait for r's layout (the workers take the to-do list in order, and r
    // is among its oldest); with no worker thread, lay it out here.
    void wait_layout(Request *r) {
        std::unique_lock<std::mutex> lk(prep_mu);
        if (workers.empty() && r->prep_state == 0) {
            auto q = std::find(todo.begin(), todo.end(), r);
            if (q != todo.end()) todo.erase(q);
            r->prep_state = 1;
            lk.unlock();
            int rc = TITAN_OK;
            std::string err;
            titan_mi355x::Prepared *p = titan_mi355x::prepare_body_sched(r->sched, &r->desc, &rc, &err);
            lk.lock();
            r->prep = p;
            r->prep_rc = rc;
            r->prep_err.swap(err);
            r->prep_state = 2;
            return;
        }
        prep_cv.wait(lk, [&] { return r->prep_state == 2; });
    }

    // called when the context is destroyed (not from a destructor, which
    // may run after the HIP runtime's own teardown at process exit)
    void release_events() {
        for (auto &f : finishing) (void)hipEventDestroy(f.ev);
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
        for (auto &g : gaps) {
            (void)hipEventDestroy(g.first);
            (void)hipEventDestroy(g.second);
        }
        if (last_end) (void)hipEventDestroy(last_end);
        if (q_prev) (void)hipEventDestroy(q_prev);
        if (q_last) (void)hipEventDestroy(q_last);
        q_prev = q_last = nullptr;
        for (hipEvent_t e : tspare) (void)hipEventDestroy(e);
        finishing.clear();
        spare.clear();
        gaps.clear();
        last_end = nullptr;
        tspare.clear();
    }

    // A timed event from the spare list, or a new one.
    int take_timed(hipEvent_t *ev) {
        if (!tspare.empty()) {
            *ev = tspare.back();
            tspare.pop_back();
            return TITAN_OK;
        }
        return hipEventCreate(ev) == hipSuccess ? TITAN_OK : TITAN_E_HIP;
    }

    // Add the gaps whose second event the device has passed (in order; never
    // waits).
    void read_gaps() {
        while (!gaps.empty() && hipEventQuery(gaps.front().second) == hipSuccess) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, gaps.front().first, gaps.front().second) == hipSuccess && ms > 0.0f)
                idle_s += 1e-3 * ms;
            tspare.push_back(gaps.front().first);
            tspare.push_back(gaps.front().second);
            gaps.pop_front();
        }
    }
};

std::mutex g_mu;
std::map<const titan_ctx *, std::unique_ptr<Server>> g_servers;

Server &server_of(titan_ctx *ctx) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto &p = g_servers[ctx];
    if (!p) {
        p.reset(new Server);
        const char *e = std::getenv("TITAN_SERVE_ADMIT");
        p->admit_waits = e && std::strcmp(e, "wait") == 0;
        const char *w = std::getenv("TITAN_SERVE_WORKERS");
        const long nw = w ? std::strtol(w, nullptr, 10) : 0;
        p->n_workers = nw > 0 ? (unsigned)std::min(nw, 64L) : (host_threads() > 4 ? host_threads() - 4 : 1u);
    }
    return *p;
}

void finish(Server &sv, Request *r, int32_t state, double t) {
    r->state = state;
    r->t_done = t;
    r->release_arrays();
    if (state == TITAN_REQ_DONE) {
        sv.completed++;
        sv.latencies.push_back(t - r->t_submit);
        sv.mass_steps += (double)r->n * (double)r->horizon;
        sv.last_done = t;
    } else if (state == TITAN_REQ_CANCELLED) {
        sv.cancelled++;
    } else {
        sv.failed++;
    }
}

// An event from the spare list, or a new one.
int take_event(Server &sv, hipEvent_t *ev) {
    if (!sv.spare.empty()) {
        *ev = sv.spare.back();
        sv.spare.pop_back();
        return TITAN_OK;
    }
    return hipEventCreateWithFlags(ev, hipEventDisableTiming) == hipSuccess ? TITAN_OK : TITAN_E_HIP;
}

// Retire a request whose steps are all enqueued.  With a result to read,
// the download waits for the steps, and the request completes then.
// Without one, the body is released at once (its slots are reused only
// after the steps in flight) and the request completes when the host sees
// the step stream pass an event recorded after its last step, so that its
// completion time is never earlier than the device's.
int retire(titan_ctx *ctx, Server &sv, Request *r) {
    if (r->out_pos || r->out_vel) {
        const int rc = titan_body_download(ctx, r->body, r->out_pos, r->out_vel, nullptr, nullptr);
        if (rc != TITAN_OK) return rc;
    }
    hipEvent_t ev = nullptr;
    if (!(r->out_pos || r->out_vel)) {
        const int rt = take_event(sv, &ev);
        if (rt != TITAN_OK) return rt;
        const int rc = titan_mi355x::record_step_event(ctx, ev);
        if (rc != TITAN_OK) {
            sv.spare.push_back(ev);
            return rc;
        }
    }
    const int rc = titan_body_destroy(ctx, r->body);
    if (rc != TITAN_OK) {
        if (ev) sv.spare.push_back(ev);
        return rc;
    }
    r->body = -1;
    sv.running_bytes -= r->est_bytes;
    if (ev) sv.finishing.push_back({r, ev});
    else finish(sv, r, TITAN_REQ_DONE, now_s());
    return TITAN_OK;
}

// Complete the finishing requests whose events the step stream has passed
// (all of them, waiting, when `wait`); they pass in the order retired.
int observe(Server &sv, bool wait) {
    while (!sv.finishing.empty()) {
        Finishing f = sv.finishing.front();
        const hipError_t e = wait ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
        if (e == hipErrorNotReady) return TITAN_OK;
        if (e != hipSuccess) return TITAN_E_HIP;
        sv.finishing.pop_front();
        sv.spare.push_back(f.ev);
        finish(sv, f.r, TITAN_REQ_DONE, now_s());
    }
    return TITAN_OK;
}

}  // namespace

// Called by titan_ctx_destroy (internal to the library).
extern "C" __attribute__((visibility("hidden"))) void titan_internal_serve_forget(const titan_ctx *ctx) {
    std::unique_ptr<Server> sv;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_servers.find(ctx);
        if (it == g_servers.end()) return;
        sv = std::move(it->second);
        g_servers.erase(it);
    }
    sv->stop_workers();  // they read requests' arrays, which go with the server
    sv->release_events();
}

static int serve_config(titan_ctx *ctx, int32_t max_running, uint64_t hbm_bytes, int64_t quantum) {
    if (!ctx || max_running < 1 || quantum < 1) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    sv.max_running = max_running;
    sv.hbm_bytes = hbm_bytes;
    sv.quantum = quantum;
    return TITAN_OK;
}

static int serve_submit(titan_ctx *ctx, const titan_request *req, int64_t *ticket) {
    if (!ctx || !req || req->horizon < 0 || req->body.n_masses <= 0 || !req->body.pos || !req->body.m)
        return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    std::unique_ptr<Request> r(new Request);
    try {
        r->copy_from(req->body);
    } catch (const std::bad_alloc &) {
        return TITAN_E_NOMEM;
    }
    r->ticket = sv.next_ticket++;
    r->horizon = req->horizon;
    r->out_pos = req->out_pos;
    r->out_vel = req->out_vel;
    r->t_submit = now_s();
    r->est_bytes = estimate_bytes(r->n, r->s);
    if (titan_get_schedule(ctx, &r->sched) != TITAN_OK) return TITAN_E_INVALID;
    if (sv.first_submit < 0) sv.first_submit = r->t_submit;
    sv.submitted++;
    Request *rp = r.get();
    sv.queue.push_back(rp);
    if (ticket) *ticket = r->ticket;
    sv.reqs[r->ticket] = std::move(r);
    {
        std::lock_guard<std::mutex> lk(sv.prep_mu);
        sv.todo.push_back(rp);
    }
    if (sv.workers.empty()) sv.start_workers(sv.n_workers);
    sv.prep_cv.notify_one();
    return TITAN_OK;
}

static int serve_pump(titan_ctx *ctx, int32_t *n_running, int32_t *n_queued) {
    if (!ctx) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    std::vector<Request *> fresh;  // admitted during this pump's quantum: they join the next one
    // On every exit, the error returns included, the running set becomes the
    // requests that hold a device body: those admitted by this pump join it
    // and those retired (body released) leave it, so that a failed pump
    // loses no request and the next one steps no released body.  `fresh` is
    // emptied whenever its requests join the running set.
    struct Reconcile {
        Server &sv;
        std::vector<Request *> &fresh;
        ~Reconcile() {
            std::vector<Request *> keep;
            keep.reserve(sv.running.size() + fresh.size());
            for (const std::vector<Request *> *list : {&sv.running, &fresh})
                for (Request *r : *list)
                    if (r->state == TITAN_REQ_RUNNING && r->body >= 0) keep.push_back(r);
            sv.running.swap(keep);
            fresh.clear();
        }
    } reconcile{sv, fresh};
    // admit, in arrival order, while the running set has room: up to `cap`
    // requests, in batches of queued requests whose layouts the workers have
    // built, placed one by one.  A request whose layout is still being built
    // is passed over (a later pump admits it), so that the pump never waits
    // for the host while the device has work and a large layout holds up no
    // other request; with nothing running or admitted the pump waits for the
    // oldest request's.  A request that does not fit the limits stops the
    // admission, so that none overtakes it there.  `until`: stop (between
    // placements, after the first) once the step stream has passed this
    // event, so that the device does not wait for admission (placing a body
    // copies its arrays to the device: milliseconds for a large one)
    auto admit = [&](size_t cap, hipEvent_t until) -> int {
        const unsigned nt = host_threads();
        size_t placed = 0;
        while (!sv.queue.empty() && cap > 0) {
            std::vector<Request *> pick;
            uint64_t bytes = sv.running_bytes;
            size_t n_run = sv.running.size() + fresh.size();
            for (Request *r : sv.queue) {
                if (pick.size() >= std::min<size_t>(cap, nt)) break;
                if ((int64_t)(n_run + pick.size()) >= sv.max_running) break;
                const bool busy = n_run + pick.size() > 0;
                if (sv.hbm_bytes && busy && bytes + r->est_bytes > sv.hbm_bytes) break;
                if (busy && !sv.admit_waits && !sv.layout_ready(r)) continue;
                pick.push_back(r);
                bytes += r->est_bytes;
            }
            if (pick.empty()) break;
            for (Request *r : pick) sv.wait_layout(r);
            for (Request *r : pick) {
                if (until && !sv.admit_waits && placed > 0 && hipEventQuery(until) == hipSuccess) return TITAN_OK;
                ++placed;
                const bool busy = !sv.running.empty() || !fresh.empty();
                int32_t id = -1;
                int rc = r->prep ? titan_mi355x::commit_body_device(ctx, r->prep, &id) : r->prep_rc;
                if (rc == TITAN_E_NOMEM && busy) return TITAN_OK;  // wait for memory to free up
                sv.queue.erase(std::find(sv.queue.begin(), sv.queue.end(), r));
                --cap;
                if (rc != TITAN_OK) {  // the body is invalid or can never fit: it fails alone
                    finish(sv, r, TITAN_REQ_FAILED, now_s());
                    continue;
                }
                titan_mi355x::free_prepared(r->prep);
                r->prep = nullptr;
                r->body = id;
                r->state = TITAN_REQ_RUNNING;
                sv.running_bytes += r->est_bytes;
                if (r->horizon == 0) {
                    const int rr = retire(ctx, sv, r);
                    if (rr != TITAN_OK) return rr;
                    continue;
                }
                fresh.push_back(r);
            }
        }
        return TITAN_OK;
    };
    int rc = observe(sv, false);
    if (rc != TITAN_OK) return rc;
    if (sv.running.empty()) {  // nothing on the device: a first batch starts it
        rc = admit(host_threads(), nullptr);
        if (rc != TITAN_OK) return rc;
        sv.running.swap(fresh);  // fresh is now the empty old running set
    }
    // at most two quanta in the step stream: wait for the one before the
    // last (its successor keeps the device busy meanwhile)
    if (!sv.running.empty() && sv.q_prev) {
        if (hipEventSynchronize(sv.q_prev) != hipSuccess) return TITAN_E_HIP;
        sv.spare.push_back(sv.q_prev);
        sv.q_prev = nullptr;
    }
    // step every running request together, each up to the quantum and its
    // horizon; the call returns once the launches are enqueued
    if (!sv.running.empty()) {
        std::vector<int32_t> ids;
        std::vector<int64_t> steps;
        for (Request *r : sv.running) {
            ids.push_back(r->body);
            steps.push_back(std::min(sv.quantum, r->horizon - r->done));
        }
        sv.read_gaps();
        hipEvent_t tb = nullptr, te = nullptr;  // the step stream's idle time since the last quantum
        if (sv.last_end && sv.take_timed(&tb) == TITAN_OK) {
            if (titan_mi355x::record_step_event(ctx, tb) == TITAN_OK) {
                sv.gaps.push_back({sv.last_end, tb});
                sv.last_end = nullptr;
            } else {
                sv.tspare.push_back(tb);
            }
        }
        rc = titan_step_bodies(ctx, ids.data(), steps.data(), (int32_t)ids.size());
        if (rc != TITAN_OK) return rc;
        for (size_t q = 0; q < ids.size(); ++q) sv.running[q]->done += steps[q];
        if (sv.take_timed(&te) == TITAN_OK) {
            if (titan_mi355x::record_step_event(ctx, te) == TITAN_OK) {
                if (sv.last_end) sv.tspare.push_back(sv.last_end);
                sv.last_end = te;
            } else {
                sv.tspare.push_back(te);
            }
        }
    }
    // admit more while the device works (their layouts are built on the host
    // meanwhile), until it has run the quantum before the one just enqueued,
    // so that the next pump enqueues its quantum while this one runs
    if (!sv.running.empty()) {
        hipEvent_t stepped = nullptr;
        rc = take_event(sv, &stepped);
        if (rc == TITAN_OK) rc = titan_mi355x::record_step_event(ctx, stepped);
        if (rc != TITAN_OK) {
            if (stepped) sv.spare.push_back(stepped);
            return rc;
        }
        if (sv.q_prev) sv.spare.push_back(sv.q_prev);
        sv.q_prev = sv.q_last;
        sv.q_last = stepped;
    }
    rc = admit(SIZE_MAX, sv.running.empty() ? nullptr : (sv.q_prev ? sv.q_prev : sv.q_last));
    if (rc != TITAN_OK) return rc;
    // retire the finished ones (their downloads wait for the quantum)
    std::vector<Request *> still;
    for (Request *r : sv.running) {
        if (r->done < r->horizon) {
            still.push_back(r);
            continue;
        }
        rc = retire(ctx, sv, r);
        if (rc != TITAN_OK) return rc;
    }
    still.insert(still.end(), fresh.begin(), fresh.end());
    fresh.clear();
    sv.running.swap(still);
    // nothing left to step: wait for the last finishing requests
    rc = observe(sv, sv.running.empty() && sv.queue.empty());
    if (rc != TITAN_OK) return rc;
    if (n_running) *n_running = (int32_t)(sv.running.size() + sv.finishing.size());
    if (n_queued) *n_queued = (int32_t)sv.queue.size();
    return TITAN_OK;
}

static int serve_poll(titan_ctx *ctx, int64_t ticket, int32_t *state, int64_t *steps_done) {
    if (!ctx) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    auto it = sv.reqs.find(ticket);
    if (it == sv.reqs.end()) return TITAN_E_NOTFOUND;
    if (state) *state = it->second->state;
    if (steps_done) *steps_done = it->second->done;
    return TITAN_OK;
}

static int serve_cancel(titan_ctx *ctx, int64_t ticket) {
    if (!ctx) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    auto it = sv.reqs.find(ticket);
    if (it == sv.reqs.end()) return TITAN_E_NOTFOUND;
    Request *r = it->second.get();
    if (r->state == TITAN_REQ_QUEUED) {
        sv.settle(r);  // no worker may still read its arrays
        auto q = std::find(sv.queue.begin(), sv.queue.end(), r);
        if (q != sv.queue.end()) sv.queue.erase(q);
    } else if (r->state == TITAN_REQ_RUNNING && r->body < 0) {  // finishing: its body is released
        auto f = std::find_if(sv.finishing.begin(), sv.finishing.end(),
                              [&](const Finishing &x) { return x.r == r; });
        if (f != sv.finishing.end()) {
            sv.spare.push_back(f->ev);
            sv.finishing.erase(f);
        }
    } else if (r->state == TITAN_REQ_RUNNING) {
        const int rc = titan_body_destroy(ctx, r->body);
        if (rc != TITAN_OK) return rc;
        r->body = -1;
        sv.running_bytes -= r->est_bytes;
        auto q = std::find(sv.running.begin(), sv.running.end(), r);
        if (q != sv.running.end()) sv.running.erase(q);
    } else {
        return TITAN_OK;  // already finished
    }
    finish(sv, r, TITAN_REQ_CANCELLED, now_s());
    return TITAN_OK;
}

static int serve_stats(titan_ctx *ctx, titan_serve_report *out) {
    if (!ctx || !out) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    std::memset(out, 0, sizeof(*out));
    out->submitted = sv.submitted;
    out->completed = sv.completed;
    out->cancelled = sv.cancelled;
    out->failed = sv.failed;
    out->running = (int64_t)(sv.running.size() + sv.finishing.size());
    out->queued = (int64_t)sv.queue.size();
    out->mass_steps = sv.mass_steps;
    if (!sv.latencies.empty()) {
        std::vector<double> l = sv.latencies;
        std::sort(l.begin(), l.end());
        // nearest-rank percentiles
        auto pct = [&](double p) {
            const size_t r = (size_t)std::ceil(p / 100.0 * (double)l.size());
            return l[std::min(l.size() - 1, r == 0 ? 0 : r - 1)];
        };
        out->p50_latency_s = pct(50);
        out->p99_latency_s = pct(99);
        out->max_latency_s = l.back();
    }
    if (sv.last_done > sv.first_submit && sv.first_submit >= 0)
        out->mass_steps_per_s = sv.mass_steps / (sv.last_done - sv.first_submit);
    sv.read_gaps();
    out->device_idle_s = sv.idle_s;
    return TITAN_OK;
}

static int serve_reset_stats(titan_ctx *ctx) {
    if (!ctx) return TITAN_E_INVALID;
    Server &sv = server_of(ctx);
    for (auto it = sv.reqs.begin(); it != sv.reqs.end();) {
        const int32_t st = it->second->state;
        if (st == TITAN_REQ_DONE || st == TITAN_REQ_CANCELLED || st == TITAN_REQ_FAILED) it = sv.reqs.erase(it);
        else ++it;
    }
    sv.submitted = (int64_t)sv.reqs.size();
    sv.completed = sv.cancelled = sv.failed = 0;
    sv.latencies.clear();
    sv.mass_steps = 0;
    sv.first_submit = sv.reqs.empty() ? -1 : now_s();
    sv.last_done = -1;
    sv.read_gaps();
    sv.idle_s = 0;
    return TITAN_OK;
}

// The C ABI: no C++ exception leaves the library (a host allocation that
// fails is TITAN_E_NOMEM).
namespace {
template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return TITAN_E_NOMEM;
    } catch (...) {
        return TITAN_E_INVALID;
    }
}
}  // namespace

int titan_serve_config(titan_ctx *ctx, int32_t max_running, uint64_t hbm_bytes, int64_t quantum) {
    return guarded([&] { return serve_config(ctx, max_running, hbm_bytes, quantum); });
}

int titan_serve_submit(titan_ctx *ctx, const titan_request *req, int64_t *ticket) {
    return guarded([&] { return serve_submit(ctx, req, ticket); });
}

int titan_serve_pump(titan_ctx *ctx, int32_t *n_running, int32_t *n_queued) {
    return guarded([&] { return serve_pump(ctx, n_running, n_queued); });
}

int titan_serve_poll(titan_ctx *ctx, int64_t ticket, int32_t *state, int64_t *steps_done) {
    return guarded([&] { return serve_poll(ctx, ticket, state, steps_done); });
}

int titan_serve_cancel(titan_ctx *ctx, int64_t ticket) {
    return guarded([&] { return serve_cancel(ctx, ticket); });
}

int titan_serve_stats(titan_ctx *ctx, titan_serve_report *out) {
    return guarded([&] { return serve_stats(ctx, out); });
}

int titan_serve_reset_stats(titan_ctx *ctx) {
    return guarded([&] { return serve_reset_stats(ctx); });
}
