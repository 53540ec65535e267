// gdp_host.cpp - host propagators exported by libGpuProp.so.
//
//   seq_propagate      cpu_seq: Gauss-Seidel rounds over marked rows
//                      (reference src/propagators/sequential_propagator.h:13-123,
//                      :207-296).
//   seq_dis_propagate  cpu_seq_dis: activities of every row computed at the
//                      start of a round, then the Gauss-Seidel sweep
//                      (sequential_propagator.h:125-204).
//   omp_propagate      cpu_omp: marked rows in parallel against a round-start
//                      copy of the bounds, bound writes under per-variable locks
//                      re-checked against the live bound
//                      (src/propagators/OMP_propagator.h:10-233).  The locks
//                      live on the heap (the reference's stack array overflows
//                      for large n).
//
// All three stop at the end of a round after which some variable has
// lb > ub + GDP_EPS, and share the arithmetic of gdp_arith.h with the GPU.
#include "gdp_host.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "gdp_arith.h"

namespace gdp {

// computeActivities (src/propagation_methods.cuh:161-221): CSR order.
RowActs row_activities(int i, const int *col, const int *row_ptr, const double *val, const double *lb,
                       const double *ub) {
    RowActs r{0.0, 0.0, 0.0, 0, 0};
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        const int j = col[k];
        const Contribution c = contribution(val[k], lb[j], ub[j]);
        r.delta = eps_max_step(r.delta, c.delta);
        r.min_inf += c.min_inf;
        r.max_inf += c.max_inf;
        if (!c.min_inf) r.minact += c.min_part;
        if (!c.max_inf) r.maxact += c.max_part;
    }
    return r;
}

namespace {

struct Csc {
    std::vector<int> colptr, rowidx;
};

// CSR -> CSC structure by counting sort (stands in for csr_to_csc,
// src/kernels/util_kernels.cuh:92-152; only the row sets per column matter).
Csc build_csc(int m, int n, const int *col, const int *row_ptr) {
    Csc c;
    c.colptr.assign(n + 1, 0);
    c.rowidx.resize(row_ptr[m]);
    for (int k = 0; k < row_ptr[m]; k++) c.colptr[col[k] + 1]++;
    for (int j = 0; j < n; j++) c.colptr[j + 1] += c.colptr[j];
    std::vector<int> fill(c.colptr.begin(), c.colptr.end() - 1);
    for (int i = 0; i < m; i++)
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) c.rowidx[fill[col[k]]++] = i;
    return c;
}

void mark_column(int j, const Csc &c, int *marked) {
    for (int p = c.colptr[j]; p < c.colptr[j + 1]; p++) marked[c.rowidx[p]] = 1;
}

bool crossed_any(int n, const double *lb, const double *ub, const unsigned char *only) {
    for (int j = 0; j < n; j++)
        if ((!only || only[j]) && eps_gt(lb[j], ub[j])) return true;
    return false;
}

// One Gauss-Seidel sweep (sequential_propagator.h:13-123).  With `acts`
// given, the row activities of the round are taken from there.
bool seq_round(int m, const int *col, const int *row_ptr, const double *val, const double *lhs,
               const double *rhs, double *lb, double *ub, const int *vartype, const Csc &csc,
               std::vector<int> &marked, std::vector<int> &next, std::vector<unsigned char> &chg,
               const std::vector<RowActs> *acts) {
    bool change = false;
    std::fill(next.begin(), next.end(), 0);
    for (int i = 0; i < m; i++) {
        if (!marked[i]) continue;
        const RowActs a = acts ? (*acts)[i] : row_activities(i, col, row_ptr, val, lb, ub);
        if (!row_gate(a.minact, a.maxact, a.min_inf, a.max_inf, lhs[i], rhs[i], a.delta)) continue;
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            const int j = col[k];
            double nl = lb[j], nu = ub[j];
            const int t = tighten_variable(val[k], lhs[i], rhs[i], a.minact, a.maxact, a.min_inf, a.max_inf,
                                           vartype[j] == kVarContinuous, lb[j], ub[j], nl, nu);
            if (t & 1) lb[j] = nl;
            if (t & 2) ub[j] = nu;
            if (t) {
                change = true;
                chg[j] = 1;
                mark_column(j, csc, next.data());
            }
        }
    }
    marked.swap(next);
    return change;
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

}  // namespace

HostResult seq_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds, const RoundObserver *observer) {
    HostResult res;
    if (crossed_any(n, lb, ub, nullptr)) { res.infeasible = 1; return res; }
    const Csc csc = build_csc(m, n, col, row_ptr);
    std::vector<int> marked(m, 1), next(m, 0);
    std::vector<unsigned char> chg(n, 0);
    bool change = true;
    int round;
    const Clock::time_point t0 = Clock::now();
    for (round = 1; round <= max_rounds && change; round++) {
        std::fill(chg.begin(), chg.end(), 0);
        change = seq_round(m, col, row_ptr, val, lhs, rhs, lb, ub, vartype, csc, marked, next, chg, nullptr);
        const bool go_on = !observer || (*observer)(round, lb, ub);
        if (change && crossed_any(n, lb, ub, chg.data())) {
            res.infeasible = 1;
            res.rounds = round;
            res.loop_ms = ms_since(t0);
            return res;
        }
        if (!go_on) {
            res.rounds = round;
            res.capped = change ? 1 : 0;
            res.loop_ms = ms_since(t0);
            return res;
        }
    }
    res.loop_ms = ms_since(t0);
    res.rounds = round - 1;
    res.capped = change ? 1 : 0;
    return res;
}

HostResult seq_dis_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                             const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                             int max_rounds) {
    HostResult res;
    if (crossed_any(n, lb, ub, nullptr)) { res.infeasible = 1; return res; }
    const Csc csc = build_csc(m, n, col, row_ptr);
    std::vector<int> marked(m, 1), next(m, 0);
    std::vector<unsigned char> chg(n, 0);
    std::vector<RowActs> acts(m);
    bool change = true;
    int round;
    const Clock::time_point t0 = Clock::now();
    for (round = 1; round <= max_rounds && change; round++) {
        for (int i = 0; i < m; i++) acts[i] = row_activities(i, col, row_ptr, val, lb, ub);
        std::fill(chg.begin(), chg.end(), 0);
        change = seq_round(m, col, row_ptr, val, lhs, rhs, lb, ub, vartype, csc, marked, next, chg, &acts);
        if (change && crossed_any(n, lb, ub, chg.data())) {
            res.infeasible = 1;
            res.rounds = round;
            res.loop_ms = ms_since(t0);
            return res;
        }
    }
    res.loop_ms = ms_since(t0);
    res.rounds = round - 1;
    res.capped = change ? 1 : 0;
    return res;
}

int omp_default_threads() {
    const char *e = std::getenv("GDP_OMP_THREADS");
    if (e) {
        const int t = std::atoi(e);
        if (t > 0) return t;
    }
    return 8;  // SHARED_MEM_THREADS of the reference (src/params.h:13)
}

HostResult omp_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds, int threads) {
    HostResult res;
    if (crossed_any(n, lb, ub, nullptr)) { res.infeasible = 1; return res; }
    if (threads <= 0) threads = omp_default_threads();
    const Csc csc = build_csc(m, n, col, row_ptr);
    std::vector<int> marked(m, 1), next(m, 0), list;
    std::vector<unsigned char> chg(n, 0);
    std::vector<double> lb_snap(n), ub_snap(n);
    std::unique_ptr<omp_lock_t[]> locks(new omp_lock_t[n]);
    for (int j = 0; j < n; j++) omp_init_lock(&locks[j]);
    list.reserve(m);

    // cpu_omp runs while round < MAX_NUM_ROUNDS (OMP_propagator.h:195), one
    // round fewer than cpu_seq's cap; the cap passed in is honoured as given.
    bool change = true;
    int round;
    const Clock::time_point t0 = Clock::now();
    for (round = 1; round <= max_rounds && change; round++) {
        std::memcpy(lb_snap.data(), lb, n * sizeof(double));
        std::memcpy(ub_snap.data(), ub, n * sizeof(double));
        std::fill(next.begin(), next.end(), 0);
        std::fill(chg.begin(), chg.end(), 0);
        list.clear();
        for (int i = 0; i < m; i++)
            if (marked[i]) list.push_back(i);
        int any = 0;
        const int cnt = static_cast<int>(list.size());
#pragma omp parallel for schedule(static) num_threads(threads) reduction(| : any)
        for (int q = 0; q < cnt; q++) {
            const int i = list[q];
            const RowActs a = row_activities(i, col, row_ptr, val, lb_snap.data(), ub_snap.data());
            if (!row_gate(a.minact, a.maxact, a.min_inf, a.max_inf, lhs[i], rhs[i], a.delta)) continue;
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
                const int j = col[k];
                double nl = lb_snap[j], nu = ub_snap[j];
                const int t = tighten_variable(val[k], lhs[i], rhs[i], a.minact, a.maxact, a.min_inf,
                                               a.max_inf, vartype[j] == kVarContinuous, lb_snap[j], ub_snap[j],
                                               nl, nu);
                if (!t) continue;
                omp_set_lock(&locks[j]);
                if ((t & 1) && eps_gt(nl, lb[j])) lb[j] = nl;
                if ((t & 2) && eps_lt(nu, ub[j])) ub[j] = nu;
                omp_unset_lock(&locks[j]);
                any = 1;
                __atomic_store_n(&chg[j], static_cast<unsigned char>(1), __ATOMIC_RELAXED);
                for (int p = csc.colptr[j]; p < csc.colptr[j + 1]; p++)
                    __atomic_store_n(&next[csc.rowidx[p]], 1, __ATOMIC_RELAXED);
            }
        }
        change = any != 0;
        marked.swap(next);
        if (change && crossed_any(n, lb, ub, chg.data())) {
            res.infeasible = 1;
            res.rounds = round;
            break;
        }
    }
    res.loop_ms = ms_since(t0);
    if (!res.infeasible) {
        res.rounds = round - 1;
        res.capped = change ? 1 : 0;
    }
    for (int j = 0; j < n; j++) omp_destroy_lock(&locks[j]);
    return res;
}

}  // namespace gdp
