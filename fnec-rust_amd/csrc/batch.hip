// Batched solves and the serving session (nec_worker serving mode).
//
// fnec_hip_solve_batch runs many independent Hallén problems (frequency-sweep
// points, geometries of uneven N) through one set of batched launches: every
// kernel walks a work list of (problem, tile) items, so problems of different
// sizes share each launch. Far-field grids of the batch run afterwards in one
// batched RP launch that reads the currents in place on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fnec_hip.h"
#include "engine.h"

using namespace fnec;

namespace {

double now_us()
{
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

// Owned copy of a request (the session keeps it past submit).
struct OwnedRequest {
    std::vector<fnec_seg> segs;
    std::vector<fnec_c64> rhs;
    std::vector<double> cosv;
    std::vector<size_t> wf, wl, ja, jb;
    std::vector<double> js, theta, phi;
    std::vector<fnec_stamp> stamps;
    fnec_request req;
    double t_submit = 0.0;
    size_t footprint = 0;

    // Every array is reassigned (or cleared when the request has none): an
    // object is reused for later requests.
    void own(const fnec_request &r)
    {
        segs.assign(r.segs, r.segs + r.n);
        rhs.assign(r.rhs, r.rhs + r.n);
        cosv.assign(r.cos_vec, r.cos_vec + r.n);
        wf.clear();
        wl.clear();
        ja.clear();
        jb.clear();
        js.clear();
        theta.clear();
        phi.clear();
        if (r.n_wires) {
            wf.assign(r.wire_first, r.wire_first + r.n_wires);
            wl.assign(r.wire_last, r.wire_last + r.n_wires);
        }
        if (r.n_junc) {
            ja.assign(r.j_a, r.j_a + r.n_junc);
            jb.assign(r.j_b, r.j_b + r.n_junc);
            js.assign(r.j_sign, r.j_sign + r.n_junc);
        }
        if (r.rp_npts && r.rp_theta && r.rp_phi) {
            theta.assign(r.rp_theta, r.rp_theta + r.rp_npts);
            phi.assign(r.rp_phi, r.rp_phi + r.rp_npts);
        }
        stamps.clear();
        if (r.n_stamps && r.stamps)
            stamps.assign(r.stamps, r.stamps + r.n_stamps);
        req = r;
        req.segs = segs.data();
        req.rhs = rhs.data();
        req.cos_vec = cosv.data();
        req.wire_first = wf.empty() ? nullptr : wf.data();
        req.wire_last = wl.empty() ? nullptr : wl.data();
        req.j_a = ja.empty() ? nullptr : ja.data();
        req.j_b = jb.empty() ? nullptr : jb.data();
        req.j_sign = js.empty() ? nullptr : js.data();
        req.rp_theta = theta.empty() ? nullptr : theta.data();
        req.rp_phi = phi.empty() ? nullptr : phi.data();
        req.rp_npts = theta.size();
        req.stamps = stamps.empty() ? nullptr : stamps.data();
        req.n_stamps = stamps.size();
    }
};

// Run one batch of requests; fills results (x / rp copied when the caller
// gave buffers, otherwise kept internally in `xs` / `rps`).
int run_batch(const fnec_request *const *reqs, size_t count, int precision, fnec_result *results,
              std::vector<std::vector<fnec_c64>> *xs, std::vector<std::vector<fnec_rp_row>> *rps)
{
    // the whole batch runs on one lane: the far-field helper thread binds it too
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    void *lane = lane_scope.lane();
    // problem descriptions keep their capacity across this thread's batches
    static thread_local std::vector<HostProblem> hps;
    if (hps.size() < count)
        hps.resize(count);
    std::vector<int> pre(count, FNEC_OK);
    {
        HostTimer ht(kHostPrep);
        for (size_t i = 0; i < count; ++i) {
            const fnec_request &r = *reqs[i];
            results[i].tag = r.tag;
            results[i].status = FNEC_OK;
            if (!results[i].x)
                results[i].x = r.x_out;
            if (!results[i].rp)
                results[i].rp = r.rp_out;
            pre[i] = problem_from_hallen(hps[i], r.segs, r.n, r.rhs, r.cos_vec, r.wire_first, r.wire_last,
                                         r.n_wires, r.j_a, r.j_b, r.j_sign, r.n_junc, r.freq_hz, &r.ground);
            if (!pre[i] && r.feed_index >= r.n) {
                set_error("feed index outside the segment list");
                pre[i] = FNEC_INVALID_ARG;
            }
            if (!pre[i])
                pre[i] = problem_add_stamps(hps[i], r.stamps, r.n_stamps, r.n);
        }
    }
    std::vector<const HostProblem *> ok;
    std::vector<size_t> idx;
    for (size_t i = 0; i < count; ++i) {
        if (pre[i])
            results[i].status = pre[i];
        else {
            ok.push_back(&hps[i]);
            idx.push_back(i);
        }
    }
    if (ok.empty())
        return FNEC_OK;
    std::vector<std::vector<fnec_c64>> xloc(ok.size());
    std::vector<fnec_c64 *> xo(ok.size());
    for (size_t q = 0; q < ok.size(); ++q) {
        xloc[q].resize(ok[q]->s);
        xo[q] = xloc[q].data();
    }
    // far-field jobs read the currents in place on the device; their inputs
    // are prepared and uploaded while the device works on the solve
    std::vector<RpHost> jobs;
    std::vector<size_t> job_q;
    std::vector<std::pair<size_t, int>> job_err; // (q, status) of grids that could not be set up
    RpStaged staged;
    int rp_st = FNEC_OK;
    std::string rp_err;
    // set up on a host thread of its own while this thread queues the solve
    std::thread rp_thread;
    struct Join {
        std::thread &t;
        ~Join()
        {
            if (t.joinable())
                t.join();
        }
    } rp_join{rp_thread};
    std::vector<double *> rp_dev_x;
    auto rp_setup = [&]() {
        LaneScope bind(lane);
        HostTimer ht(kHostRp);
        const std::vector<double *> &dev_x = rp_dev_x;
        for (size_t q = 0; q < ok.size(); ++q) {
            const fnec_request &r = *reqs[idx[q]];
            if (!(r.rp_npts && r.rp_theta && r.rp_phi))
                continue;
            RpHost h;
            h.cur_dev = dev_x[q];
            const int e = rp_host_from(h, r.segs, nullptr, r.n, r.freq_hz, &r.ground, r.rp_theta, r.rp_phi,
                                       r.rp_npts);
            if (e) {
                job_err.push_back({q, e});
                continue;
            }
            jobs.push_back(std::move(h));
            job_q.push_back(q);
        }
        if (!jobs.empty())
            rp_st = rp_stage(jobs, staged);
        if (rp_st)
            rp_err = last_error(); // the message is per thread
    };
    bool any_rp = false;
    for (size_t q = 0; q < ok.size(); ++q) {
        const fnec_request &r = *reqs[idx[q]];
        any_rp = any_rp || (r.rp_npts && r.rp_theta && r.rp_phi);
    }
    auto overlap = [&](const std::vector<double *> &dev_x) {
        if (!any_rp)
            return;
        rp_dev_x = dev_x;
        rp_thread = std::thread(rp_setup);
    };
    // a failed call leaves every admitted request with its status, not OK
    auto fail_admitted = [&](int code) {
        for (size_t q = 0; q < ok.size(); ++q)
            results[idx[q]].status = code;
        return code;
    };
    std::vector<fnec_solve_stats> stats;
    std::vector<int> status;
    std::vector<double *> dev_x;
    int st = solve_problems(ok, xo, stats, status, &dev_x, precision, overlap);
    if (rp_thread.joinable())
        rp_thread.join();
    if (st)
        return fail_admitted(st);
    if (rp_st) {
        set_error(rp_err);
        return fail_admitted(rp_st);
    }
    {
        HostTimer ht(kHostPost);
        for (size_t q = 0; q < ok.size(); ++q) {
            const size_t i = idx[q];
            const fnec_request &r = *reqs[i];
            fnec_result &res = results[i];
            res.status = status[q];
            res.stats = stats[q];
            if (status[q])
                continue;
            const fnec_c64 cur = xloc[q][r.feed_index];
            res.feed_current = cur;
            // Z = V / I when |I| > 1e-60, else V (nec_worker/src/solve.rs:254-260)
            if (std::hypot(cur.re, cur.im) > 1e-60) {
                const double ns = cur.re * cur.re + cur.im * cur.im;
                res.z_in.re = (r.v_feed.re * cur.re + r.v_feed.im * cur.im) / ns;
                res.z_in.im = (r.v_feed.im * cur.re - r.v_feed.re * cur.im) / ns;
            } else {
                res.z_in = r.v_feed;
            }
            if (res.x)
                std::memcpy(res.x, xloc[q].data(), sizeof(fnec_c64) * xloc[q].size());
            if (xs)
                (*xs)[i] = std::move(xloc[q]);
        }
        for (const auto &qe : job_err)
            if (!status[qe.first])
                results[idx[qe.first]].status = qe.second;
    }
    if (!jobs.empty()) {
        // rows are read back only into the caller's buffers (or the internal
        // copies a caller asked for); a failed solve's pattern is dropped
        std::vector<fnec_rp_row *> outs(jobs.size(), nullptr);
        if (rps)
            for (size_t j = 0; j < jobs.size(); ++j)
                if (!status[job_q[j]])
                    (*rps)[idx[job_q[j]]].resize(jobs[j].npts);
        for (size_t j = 0; j < jobs.size(); ++j) {
            const size_t i = idx[job_q[j]];
            if (status[job_q[j]])
                continue;
            outs[j] = results[i].rp ? results[i].rp : (rps ? (*rps)[i].data() : nullptr);
        }
        std::vector<double> totals;
        st = rp_finish(staged, outs, totals);
        if (st)
            return fail_admitted(st);
        for (size_t j = 0; j < jobs.size(); ++j)
            if (!status[job_q[j]])
                results[idx[job_q[j]]].rp_total = totals[j];
    }
    return FNEC_OK;
}

} // namespace

// Serving session. Worker threads (one per engine lane) take admitted
// requests off the queue as batches and run them; a batch runs on its own
// lane, so two batches' kernels overlap on the device while new requests are
// submitted. Requests are admitted in arrival order while the footprint of
// every batch in flight stays within the budget.
struct fnec_session {
    size_t budget = 0;
    int precision = FNEC_PREC_FP64;
    int max_batch = 64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<OwnedRequest *> queue;
    std::deque<fnec_result> done;
    // finished requests, reused by later submits: their arrays keep their
    // capacity, so a submit copies into memory that is already mapped
    std::vector<OwnedRequest *> spare;
    size_t in_flight_bytes = 0; // footprint of the batches being solved
    size_t running = 0;         // requests in those batches
    uint64_t batches = 0;       // batches formed so far (the next batch id)
    double last_submit = 0.0;
    bool stop = false;
    std::string err; // message of the last failed batch
    std::vector<std::thread> workers;
};

namespace {

// Batch formation: a worker that finds admissible requests waits until no
// request has been submitted for kSettleUs (a burst of submits ends up in one
// batch), the batch is full, or the oldest request has waited kLingerUs.
constexpr double kSettleUs = 500.0;
constexpr double kLingerUs = 20000.0;

void session_worker(fnec_session *s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            return s->stop ||
                   (!s->queue.empty() && s->in_flight_bytes + s->queue.front()->footprint <= s->budget);
        });
        if (s->stop)
            return;
        while (!s->stop && !s->queue.empty() && (int)s->queue.size() < s->max_batch) {
            const double now = now_us();
            const double quiet = now - s->last_submit, age = now - s->queue.front()->t_submit;
            if (quiet >= kSettleUs || age >= kLingerUs)
                break;
            const double w = std::max(1.0, std::min(kSettleUs - quiet, kLingerUs - age));
            s->cv_work.wait_for(lk, std::chrono::microseconds((long)w));
        }
        if (s->stop)
            return;
        std::vector<OwnedRequest *> batch;
        size_t bytes = 0;
        // size classes (tunable batch_split_n): while the queue holds less
        // than a full batch (latency regime), the batch takes only requests
        // of the oldest request's class, so that small solves do not wait for
        // a large one's factorisation (FIFO within the class); with a full
        // batch queued (throughput regime) it is formed FIFO across classes
        const long split = (int)s->queue.size() < s->max_batch ? get_tunable("batch_split_n") : 0;
        const bool big = split > 0 && s->queue.front()->req.n > (size_t)split;
        for (auto it = s->queue.begin(); it != s->queue.end() && (int)batch.size() < s->max_batch;) {
            OwnedRequest *o = *it;
            if (split > 0 && (o->req.n > (size_t)split) != big) {
                ++it;
                continue;
            }
            if (s->in_flight_bytes + bytes + o->footprint > s->budget)
                break;
            bytes += o->footprint;
            batch.push_back(o);
            it = s->queue.erase(it);
        }
        if (batch.empty())
            continue; // another worker took the requests, or they do not fit yet
        s->in_flight_bytes += bytes;
        s->running += batch.size();
        const uint64_t id = s->batches++;
        lk.unlock();

        const double t_start = now_us();
        std::vector<const fnec_request *> rp(batch.size());
        std::vector<fnec_result> res(batch.size());
        std::memset(res.data(), 0, sizeof(fnec_result) * res.size());
        for (size_t i = 0; i < batch.size(); ++i)
            rp[i] = &batch[i]->req;
        int st;
        {
            std::lock_guard<std::mutex> g(global_mutex());
            st = ensure_ready();
        }
        if (!st)
            st = run_batch(rp.data(), rp.size(), s->precision, res.data(), nullptr, nullptr);
        std::string msg = st ? std::string(last_error()) : std::string();
        for (size_t i = 0; i < batch.size() && !st; ++i)
            if (res[i].status && msg.empty())
                msg = last_error();
        const double t_done = now_us();

        lk.lock();
        if (!msg.empty())
            s->err = msg;
        for (size_t i = 0; i < batch.size(); ++i) {
            res[i].tag = batch[i]->req.tag; // also when the batch never ran
            if (st)
                res[i].status = st;
            res[i].t_submit_us = batch[i]->t_submit;
            res[i].t_start_us = t_start;
            res[i].t_done_us = t_done;
            res[i].batch = id;
            s->done.push_back(res[i]);
            s->spare.push_back(batch[i]);
        }
        s->in_flight_bytes -= bytes;
        s->running -= batch.size();
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

} // namespace

extern "C" {

size_t fnec_hip_request_footprint(size_t n, size_t n_wires, size_t n_junc, size_t rp_npts, int precision)
{
    const size_t w = n_wires ? n_wires : 1;
    const size_t s = n + w;
    const size_t ncons = 2 * w + n_junc;
    size_t b = problem_bytes((int)n, (int)s, (int)ncons, (int)w + 1, precision);
    b += rp_npts * (8 + 8 + 32 + 32) + 8 * ((181 * 360 + 255) / 256) + 1024;
    return b;
}

int fnec_hip_solve_batch(const fnec_request *reqs, size_t count, int precision, fnec_result *results)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (count && (!reqs || !results)) {
        set_error("null request or result array");
        return FNEC_INVALID_ARG;
    }
    if (precision != FNEC_PREC_FP64 && precision != FNEC_PREC_MIXED && precision != FNEC_PREC_FP64_MFMA) {
        set_error("unknown precision");
        return FNEC_INVALID_ARG;
    }
    std::vector<const fnec_request *> rp(count);
    for (size_t i = 0; i < count; ++i)
        rp[i] = &reqs[i];
    return run_batch(rp.data(), count, precision, results, nullptr, nullptr);
}

fnec_session *fnec_hip_session_create(size_t budget_bytes, int precision, int max_batch)
{
    if (precision != FNEC_PREC_FP64 && precision != FNEC_PREC_MIXED && precision != FNEC_PREC_FP64_MFMA) {
        set_error("unknown precision");
        return nullptr;
    }
    fnec_session *s = new fnec_session;
    s->budget = budget_bytes;
    s->precision = precision;
    s->max_batch = max_batch > 0 ? max_batch : 64;
    const long lanes = std::max(1L, get_tunable("lanes"));
    for (long i = 0; i < lanes; ++i)
        s->workers.emplace_back(session_worker, s);
    return s;
}

int fnec_hip_session_submit(fnec_session *s, const fnec_request *req)
{
    if (!s || !req || !req->segs || !req->rhs || !req->cos_vec || req->n == 0 || (req->n_stamps && !req->stamps)) {
        set_error("invalid request");
        return FNEC_INVALID_ARG;
    }
    const size_t fp = fnec_hip_request_footprint(req->n, req->n_wires, req->n_junc, req->rp_npts,
                                                 s->precision);
    if (fp > s->budget) {
        set_error("request footprint exceeds the session budget");
        return FNEC_RESOURCE_EXHAUSTED;
    }
    OwnedRequest *o = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->spare.empty()) {
            o = s->spare.back();
            s->spare.pop_back();
        }
    }
    if (!o)
        o = new OwnedRequest;
    o->own(*req);
    o->footprint = fp;
    std::lock_guard<std::mutex> lk(s->mu);
    o->t_submit = s->last_submit = now_us();
    s->queue.push_back(o);
    s->cv_work.notify_one();
    return FNEC_OK;
}

int fnec_hip_session_pending(fnec_session *s)
{
    if (!s)
        return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    return (int)(s->queue.size() + s->running + s->done.size());
}

int fnec_hip_session_poll(fnec_session *s, fnec_result *out, int max, int wait_us)
{
    if (!s || (max > 0 && !out)) {
        set_error("invalid poll");
        return -FNEC_INVALID_ARG;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_us > 0 && (s->running || !s->queue.empty()))
        s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !s->done.empty(); });
    int k = 0;
    bool failed = false;
    while (k < max && !s->done.empty()) {
        out[k] = s->done.front();
        failed = failed || out[k].status != FNEC_OK;
        ++k;
        s->done.pop_front();
    }
    if (failed)
        set_error(s->err);
    return k;
}

void fnec_hip_session_destroy(fnec_session *s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers)
        t.join(); // a batch already on the device finishes first
    for (auto *o : s->queue)
        delete o;
    for (auto *o : s->spare)
        delete o;
    delete s;
}

} // extern "C"
