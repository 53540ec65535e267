// serve_load.cpp - load generator for the serving benchmark (bench.py), built
// as libgdp_serveload.so next to libGpuProp.so.  Not part of the library.
//
// Drives a running server (gdpServerStart) through its C API the way a
// multi-threaded branch-and-bound solver would: `threads` submitting threads
// issue requests at given arrival times (request k on thread k mod threads,
// each waiting until its arrival time), and one collecting thread takes
// completed tickets and their results.  Requests cycle through a pool of
// node vectors.  Returns each request's completion time, seconds from the
// start, so the caller computes latencies against the arrival times, and
// (submitted_s, may be null) the time its submit call returned.
//
// Request forms (mode):
//   0  full vectors copied by the server (gdpServerSubmit / gdpServerResult)
//   1  zero-copy, result into separate pinned arrays (gdpServerSubmitPinned)
//   2  zero-copy in place: the node vector is written into a free pinned
//      pair, which the result replaces (only changed columns come back)
//   3  changes to the server's base vector (gdpServerSubmitDelta /
//      gdpServerResultChanges)
// Pinned modes hold at most `pinned_pairs` requests in flight: a submitter
// waits for a free pair (back-pressure from the client's own buffers).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gdp_ext.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pool {
    int n = 0, size = 0;
    const double *const *lb = nullptr;
    const double *const *ub = nullptr;
    const int *const *cols = nullptr;
    const int *ncols = nullptr;
};

}  // namespace

extern "C" int gdpServeLoad(gdp_server s, int mode, int n_requests, const double *arrival_s, int pool_size,
                            int n_vars, const double *const *pool_lb, const double *const *pool_ub,
                            const int *const *pool_cols, const int *pool_ncols, int threads, int pinned_pairs,
                            double *done_s, double *submitted_s) {
    if (!s || n_requests < 1 || pool_size < 1 || threads < 1 || mode < 0 || mode > 3) return 0;
    const size_t n = static_cast<size_t>(n_vars);
    Pool P{n_vars, pool_size, pool_lb, pool_ub, pool_cols, pool_ncols};
    // pinned buffers of the zero-copy modes: inputs (mode 1: one pair per pool
    // vector) and a ring of output / in-place pairs
    std::vector<double *> pin_in, pin_out;
    auto host_alloc = [&](std::vector<double *> &v, int pairs) {
        for (int k = 0; k < 2 * pairs; k++) {
            void *p = nullptr;
            if (gdpHostAlloc(&p, n * sizeof(double)) != GDP_OKAY) return false;
            v.push_back(static_cast<double *>(p));
        }
        return true;
    };
    bool ok = true;
    if (mode == 1) ok = host_alloc(pin_in, pool_size);
    if (ok && (mode == 1 || mode == 2)) ok = host_alloc(pin_out, pinned_pairs);
    if (ok && mode == 1)
        for (int k = 0; k < pool_size; k++) {
            std::memcpy(pin_in[2 * k], P.lb[k], n * sizeof(double));
            std::memcpy(pin_in[2 * k + 1], P.ub[k], n * sizeof(double));
        }
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int> free_pairs;
    for (int q = pinned_pairs - 1; q >= 0 && (mode == 1 || mode == 2); q--) free_pairs.push_back(q);
    std::unordered_map<long long, int> req_of;      // ticket -> request
    std::unordered_map<long long, int> pair_of;     // ticket -> output pair
    std::unordered_map<long long, double> early;    // completed before its submitter recorded it
    std::atomic<int> done_count{0};
    std::atomic<bool> failed{!ok};
    const Clock::time_point t0 = Clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };

    auto record_done = [&](long long t, double when) {
        // with mu held
        auto it = req_of.find(t);
        if (it == req_of.end()) {
            early[t] = when;
            return;
        }
        done_s[it->second] = when;
        req_of.erase(it);
        auto pq = pair_of.find(t);
        if (pq != pair_of.end()) {
            free_pairs.push_back(pq->second);
            pair_of.erase(pq);
            cv.notify_all();
        }
        done_count++;
    };

    auto submitter = [&](int h) {
        for (int k = h; k < n_requests && !failed; k += threads) {
            const double due = arrival_s[k];
            for (double t = now_s(); t < due; t = now_s()) {
                const double wait = due - t;
                if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
            }
            const int v = k % P.size;
            int pair = -1;
            if (mode == 1 || mode == 2) {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !free_pairs.empty() || failed; });
                if (failed) return;
                pair = free_pairs.back();
                free_pairs.pop_back();
            }
            long long t = 0;
            GDP_RETCODE rc;
            if (mode == 0) {
                rc = gdpServerSubmit(s, P.lb[v], P.ub[v], P.cols[v], P.ncols[v], MAX_NUM_ROUNDS, &t);
            } else if (mode == 1) {
                rc = gdpServerSubmitPinned(s, pin_in[2 * v], pin_in[2 * v + 1], pin_out[2 * pair],
                                           pin_out[2 * pair + 1], P.cols[v], P.ncols[v], MAX_NUM_ROUNDS, &t);
            } else if (mode == 2) {
                std::memcpy(pin_out[2 * pair], P.lb[v], n * sizeof(double));
                std::memcpy(pin_out[2 * pair + 1], P.ub[v], n * sizeof(double));
                rc = gdpServerSubmitPinned(s, pin_out[2 * pair], pin_out[2 * pair + 1], pin_out[2 * pair],
                                           pin_out[2 * pair + 1], P.cols[v], P.ncols[v], MAX_NUM_ROUNDS, &t);
            } else {
                // the changed columns' new bounds
                std::vector<double> l(P.ncols[v]), u(P.ncols[v]);
                for (int q = 0; q < P.ncols[v]; q++) {
                    l[q] = P.lb[v][P.cols[v][q]];
                    u[q] = P.ub[v][P.cols[v][q]];
                }
                rc = gdpServerSubmitDelta(s, P.ncols[v], P.cols[v], l.data(), u.data(), MAX_NUM_ROUNDS, &t);
            }
            if (submitted_s) submitted_s[k] = now_s();
            std::lock_guard<std::mutex> lk(mu);
            if (rc != GDP_OKAY) {
                failed = true;
                cv.notify_all();
                return;
            }
            req_of[t] = k;
            if (pair >= 0) pair_of[t] = pair;
            auto e = early.find(t);
            if (e != early.end()) {
                const double when = e->second;
                early.erase(e);
                record_done(t, when);
            }
        }
    };

    auto collector = [&] {
        std::vector<long long> tickets(4096);
        std::vector<double> lb(n), ub(n), cl(n), cu(n);
        std::vector<int> cc(n);
        while (done_count < n_requests && !failed) {
            int cnt = 0;
            if (gdpServerCompleted(s, tickets.data(), static_cast<int>(tickets.size()), &cnt) != GDP_OKAY) {
                failed = true;
                break;
            }
            if (cnt == 0) {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                continue;
            }
            for (int q = 0; q < cnt; q++) {
                gdp_run_info info;
                int ready = 0, count = 0;
                GDP_RETCODE rc;
                if (mode == 0)
                    rc = gdpServerResult(s, tickets[q], lb.data(), ub.data(), &info, &ready);
                else if (mode == 3)
                    rc = gdpServerResultChanges(s, tickets[q], cc.data(), cl.data(), cu.data(), n_vars, &count, &info,
                                                &ready);
                else
                    rc = gdpServerResult(s, tickets[q], nullptr, nullptr, &info, &ready);
                if (rc != GDP_OKAY || !ready) {
                    failed = true;
                    break;
                }
                const double when = now_s();
                std::lock_guard<std::mutex> lk(mu);
                record_done(tickets[q], when);
            }
        }
        cv.notify_all();
    };

    std::vector<std::thread> th;
    for (int h = 0; h < threads; h++) th.emplace_back(submitter, h);
    std::thread col(collector);
    for (auto &t : th) t.join();
    col.join();
    for (double *p : pin_in) gdpHostFree(p);
    for (double *p : pin_out) gdpHostFree(p);
    return failed ? 0 : 1;
}
