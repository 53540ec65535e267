// gdp_interface.cpp - the C ABI of libGpuProp.so: the eleven entry points of
// the reference's include/interface.h plus the extension API of gdp_ext.h.
//
// Error behaviour follows src/interface.cu:7-16 (an exception becomes
// GDP_ERROR and a message on stderr), except that device errors never
// terminate the process (the reference exit()s, src/cuda_def.cuh:8-15).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/gdp_ext.h"
#include "../../include/interface.h"
#include "gdp_engine.h"
#include "gdp_host.h"

using gdp::HostResult;
using gdp::RunResult;

struct gdp_problem_s {
    gdp::Problem *p;
};

namespace {

bool empty_problem(int m, int n, int nnz) { return m == 0 || n == 0 || nnz == 0; }

GDP_RETCODE fail(const char *where, const std::exception &e) {
    std::string msg = std::string(where) + ": " + e.what();
    gdp::set_last_error(msg);
    std::fprintf(stderr, "Error happened. Exception:\n%s\n", msg.c_str());
    return GDP_ERROR;
}

bool sizes_ok(int m, int n, int nnz) { return m >= 0 && n >= 0 && nnz >= 0; }

void fill_info(gdp_run_info *info, int rounds, int infeasible, int capped, double ms) {
    if (!info) return;
    info->rounds = rounds;
    info->infeasible = infeasible;
    info->capped = capped;
    info->kernel_ms = ms;
}

// Runs the FP64 GPU path on one call's arrays (create, propagate, destroy).
GDP_RETCODE gpu_double(int m, int n, int nnz, const int *col, const int *rp, const double *val,
                       const double *lhs, const double *rhs, double *lb, double *ub, const int *vt,
                       int sync_type, gdp_run_info *info) {
    if (!sizes_ok(m, n, nnz)) {
        gdp::set_last_error("negative size");
        return GDP_ERROR;
    }
    if (sync_type < CPU_LOOP || sync_type > MEGAKERNEL) {
        gdp::set_last_error("Unknown sync type");
        std::fprintf(stderr, "Error happened. Exception:\nUnknown sync type\n");
        return GDP_ERROR;
    }
    if (empty_problem(m, n, nnz)) return GDP_OKAY;
    gdp::Problem *p = nullptr;
    try {
        p = gdp::problem_create(m, n, nnz, col, rp, val, lhs, rhs, vt, lb, ub, nullptr);
        // the results are copied into lb / ub after the propagation succeeded
        const RunResult r = gdp::problem_propagate(p, lb, ub, sync_type, MAX_NUM_ROUNDS);
        fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        gdp::problem_destroy(p);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        if (p) gdp::problem_destroy(p);
        return fail("gpu_atomic", e);
    }
}

// Float entry points compute in FP64; a bound that moved is rounded outward
// to float (lower bounds down, upper bounds up) so that it stays valid.
float outward_lower(double v) {
    float f = static_cast<float>(v);
    if (static_cast<double>(f) > v) f = std::nextafter(f, -INFINITY);
    return f;
}
float outward_upper(double v) {
    float f = static_cast<float>(v);
    if (static_cast<double>(f) < v) f = std::nextafter(f, INFINITY);
    return f;
}

template <typename Fn>
GDP_RETCODE via_double(int m, int n, int nnz, const float *val, const float *lhs, const float *rhs, float *lb,
                       float *ub, Fn &&run) {
    if (!sizes_ok(m, n, nnz)) return GDP_ERROR;
    if (empty_problem(m, n, nnz)) return GDP_OKAY;
    std::vector<double> v(val, val + nnz), l(lhs, lhs + m), r(rhs, rhs + m), lo(lb, lb + n), hi(ub, ub + n);
    const GDP_RETCODE rc = run(v.data(), l.data(), r.data(), lo.data(), hi.data());
    if (rc != GDP_OKAY) return rc;
    for (int j = 0; j < n; j++) {
        if (lo[j] != static_cast<double>(lb[j])) lb[j] = outward_lower(lo[j]);
        if (hi[j] != static_cast<double>(ub[j])) ub[j] = outward_upper(hi[j]);
    }
    return GDP_OKAY;
}

GDP_RETCODE host_double(int which, int m, int n, int nnz, const int *col, const int *rp, const double *val,
                        const double *lhs, const double *rhs, double *lb, double *ub, const int *vt,
                        int max_rounds, int threads, gdp_run_info *info) {
    if (!sizes_ok(m, n, nnz)) return GDP_ERROR;
    if (empty_problem(m, n, nnz)) return GDP_OKAY;
    try {
        HostResult r;
        if (which == 0)
            r = gdp::seq_propagate(m, n, col, rp, val, lhs, rhs, lb, ub, vt, max_rounds);
        else if (which == 1)
            r = gdp::seq_dis_propagate(m, n, col, rp, val, lhs, rhs, lb, ub, vt, max_rounds);
        else
            r = gdp::omp_propagate(m, n, col, rp, val, lhs, rhs, lb, ub, vt, max_rounds, threads);
        fill_info(info, r.rounds, r.infeasible, r.capped, r.loop_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("host propagator", e);
    }
}

// The progress-measure runs (progressPropagators.cuh:11-131 for cpu_seq,
// :134-272 for gpu_atomic): the weakest bounds of the input
// (initMeasureData, progress_function.h:19-53), the cpu_seq fixpoint as the
// limit, the largest attainable measure, then the rounds of the propagation
// with the measure of each (measureAndPrintProgress, :152-195).  The
// reference's closing run without the measure (timing only, its result
// discarded) is not repeated.
GDP_RETCODE with_measure(bool gpu, int m, int n, int nnz, const int *col, const int *rp, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vt,
                         int max_rounds, bool print, double *p_fin, double *p_inf, int *n_rounds,
                         gdp_run_info *info) {
    if (n_rounds) *n_rounds = 0;
    if (!sizes_ok(m, n, nnz)) {
        gdp::set_last_error("negative size");
        return GDP_ERROR;
    }
    if (max_rounds < 1 || max_rounds > MAX_NUM_ROUNDS) {
        gdp::set_last_error("max_rounds out of range");
        return GDP_ERROR;
    }
    if (empty_problem(m, n, nnz)) {
        if (print) std::printf("propagation of 0 size problem. Nothing to propagate.\n");
        return GDP_OKAY;
    }
    gdp::Problem *p = nullptr;
    try {
        if (print)
            std::printf(gpu ? "\n=== gpu_atomic execution with measure of progress ===\n"
                            : "\n=== cpu_seq execution with measure of progress ===\n");
        const std::vector<double> lb_in(lb, lb + n), ub_in(ub, ub + n);
        std::vector<double> wl(lb_in), wu(ub_in), ll(lb_in), lu(ub_in);
        if (print) std::printf("\n====   Running the weakest bounds procedure  ====");
        gdp::weakest_bounds(m, n, col, rp, val, lhs, rhs, vt, wl.data(), wu.data());
        if (print) std::printf("====   end weakest bounds procedure  ====");
        if (print) std::printf("\n\n==== propagator run for limit poitns ====");
        gdp::seq_propagate(m, n, col, rp, val, lhs, rhs, ll.data(), lu.data(), vt, MAX_NUM_ROUNDS);
        if (print) std::printf("==== end limit points ====\n");
        const int bad = gdp::measure_defect(n, wu.data(), lu.data());
        if (bad >= 0) {
            if (print) std::printf("varidx: %d\n", bad);
            std::fflush(stdout);
            gdp::set_last_error("variable " + std::to_string(bad) +
                                " has a finite limit upper bound but an infinite weakest upper bound");
            return GDP_ERROR;
        }
        const gdp::Measure mx = gdp::measure_max(n, lb_in.data(), ub_in.data(), wl.data(), wu.data(), ll.data(),
                                                 lu.data());
        std::vector<double> fin;
        std::vector<int> inf;
        int rounds = 0, infeasible = 0, capped = 0;
        double ms = 0.0;
        if (gpu) {
            p = gdp::problem_create(m, n, nnz, col, rp, val, lhs, rhs, vt, lb, ub, nullptr);
            std::vector<double> l(lb_in), u(ub_in);
            const RunResult r = gdp::problem_propagate_measured(p, l.data(), u.data(), max_rounds, wl.data(),
                                                                wu.data(), ll.data(), lu.data(), fin, inf);
            gdp::problem_destroy(p);
            p = nullptr;
            std::memcpy(lb, l.data(), n * sizeof(double));
            std::memcpy(ub, u.data(), n * sizeof(double));
            rounds = r.rounds;
            infeasible = r.infeasible;
            capped = r.capped;
            ms = r.kernel_ms;
        } else {
            std::vector<double> pl(lb_in), pu(ub_in);
            const gdp::RoundObserver obs = [&](int, const double *l, const double *u) {
                const gdp::Measure x = gdp::measure_round(n, l, u, wl.data(), wu.data(), ll.data(), lu.data(),
                                                          pl.data(), pu.data());
                fin.push_back(x.fin);
                inf.push_back(x.inf);
                pl.assign(l, l + n);
                pu.assign(u, u + n);
                return true;
            };
            const HostResult r = gdp::seq_propagate(m, n, col, rp, val, lhs, rhs, lb, ub, vt, max_rounds, &obs);
            rounds = r.rounds;
            infeasible = r.infeasible;
            capped = r.capped;
        }
        int inf_total = 0;
        for (size_t k = 0; k < fin.size(); k++) {
            const double pf = gdp::measure_normalize(fin[k], mx.fin);
            const double pi = gdp::measure_normalize(static_cast<double>(inf[k] + inf_total),
                                                     static_cast<double>(mx.inf));
            inf_total += inf[k];
            if (print) std::printf("round=%d, P_fin=%.10f, P_inf=%.2f\n", static_cast<int>(k) + 1, pf, pi);
            if (p_fin) p_fin[k] = pf;
            if (p_inf) p_inf[k] = pi;
        }
        if (print) std::fflush(stdout);
        if (n_rounds) *n_rounds = static_cast<int>(fin.size());
        fill_info(info, rounds, infeasible, capped, ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        if (p) gdp::problem_destroy(p);
        return fail("propagation with measure", e);
    }
}

}  // namespace

namespace {

gdp::Tuning tuning_from(const gdp_tuning *tuning) {
    gdp::Tuning T;
    if (tuning) {
        if (tuning->item_nnz > 0) T.item_nnz = tuning->item_nnz;
        if (tuning->mega_blocks_per_cu > 0) T.mega_blocks_per_cu = tuning->mega_blocks_per_cu;
        if (tuning->mega_threads > 0) T.mega_threads = tuning->mega_threads;
        if (tuning->graph_rounds > 0) T.graph_rounds = tuning->graph_rounds;
        if (tuning->max_split_segs >= 0) T.max_split_segs = tuning->max_split_segs;
        if (tuning->compact_bounds >= 0) T.compact_bounds = tuning->compact_bounds;
        if (tuning->row_marking >= 0) T.row_marking = tuning->row_marking;
        if (tuning->mark_limit >= 0) T.mark_limit = tuning->mark_limit;
        for (int b = 0; b < 10; b++)
            if (tuning->group_log[b] >= 0) T.group_log[b] = tuning->group_log[b];
    }
    return T;
}

}  // namespace

namespace {

void tuning_to(const gdp::Tuning &T, gdp_tuning *t) {
    t->item_nnz = T.item_nnz;
    t->mega_blocks_per_cu = T.mega_blocks_per_cu;
    t->mega_threads = T.mega_threads;
    t->graph_rounds = T.graph_rounds;
    t->max_split_segs = T.max_split_segs;
    t->compact_bounds = T.compact_bounds;
    t->row_marking = T.row_marking;
    t->mark_limit = T.mark_limit;
    for (int b = 0; b < 10; b++) t->group_log[b] = T.group_log[b];
}

GDP_RETCODE autotune_views(const std::vector<gdp::InstanceView> &v, const double *lbs, const double *ubs,
                           const int sync_type, gdp_tuning *tuning, double *best_ms, const char *who) {
    try {
        const gdp::Tuning start = tuning_from(tuning);
        std::vector<gdp::AutotuneTrial> trials;
        const gdp::Tuning best = gdp::autotune(v, lbs, ubs, sync_type, &start, &trials);
        double ms = 0.0;
        for (const auto &t : trials)
            if (std::memcmp(&t.tuning, &best, sizeof(best)) == 0) ms = t.kernel_ms;
        tuning_to(best, tuning);
        if (best_ms) *best_ms = ms;
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail(who, e);
    }
}

}  // namespace

namespace {
GDP_RETCODE hip_rc(hipError_t e, const char *what) {
    if (e == hipSuccess) return GDP_OKAY;
    gdp::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return GDP_ERROR;
}
}  // namespace

extern "C" {

// ---- reference entry points (include/interface.h) ----

GDP_RETCODE propagateConstraintsGPUReductionDouble(const int n_cons, const int n_vars, const int nnz,
                                                   const int *csr_col_indices, const int *csr_row_ptrs,
                                                   const double *csr_vals, const double *lhss,
                                                   const double *rhss, double *lbs, double *ubs,
                                                   const int *vartypes) {
    return gpu_double(n_cons, n_vars, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                      vartypes, CPU_LOOP, nullptr);
}

GDP_RETCODE propagateConstraintsGPUReductionFloat(const int n_cons, const int n_vars, const int nnz,
                                                  const int *csr_col_indices, const int *csr_row_ptrs,
                                                  const float *csr_vals, const float *lhss, const float *rhss,
                                                  float *lbs, float *ubs, const int *vartypes) {
    return via_double(n_cons, n_vars, nnz, csr_vals, lhss, rhss, lbs, ubs,
                      [&](double *v, double *l, double *r, double *lo, double *hi) {
                          return gpu_double(n_cons, n_vars, nnz, csr_col_indices, csr_row_ptrs, v, l, r, lo, hi,
                                            vartypes, CPU_LOOP, nullptr);
                      });
}

GDP_RETCODE propagateConstraintsGPUAtomicDouble(const int n_cons, const int n_vars, const int nnz,
                                                const int *csr_col_indices, const int *csr_row_ptrs,
                                                const double *csr_vals, const double *lhss, const double *rhss,
                                                double *lbs, double *ubs, const int *vartypes,
                                                const int sync_type) {
    return gpu_double(n_cons, n_vars, nnz, csr_col_indices, csr_row_ptrs, csr_vals, lhss, rhss, lbs, ubs,
                      vartypes, sync_type, nullptr);
}

GDP_RETCODE propagateConstraintsGPUAtomicFloat(const int n_cons, const int n_vars, const int nnz,
                                               const int *csr_col_indices, const int *csr_row_ptrs,
                                               const float *csr_vals, const float *lhss, const float *rhss,
                                               float *lbs, float *ubs, const int *vartypes, const int sync_type) {
    if (sync_type < CPU_LOOP || sync_type > MEGAKERNEL) {
        gdp::set_last_error("Unknown sync type");
        std::fprintf(stderr, "Error happened. Exception:\nUnknown sync type\n");
        return GDP_ERROR;
    }
    return via_double(n_cons, n_vars, nnz, csr_vals, lhss, rhss, lbs, ubs,
                      [&](double *v, double *l, double *r, double *lo, double *hi) {
                          return gpu_double(n_cons, n_vars, nnz, csr_col_indices, csr_row_ptrs, v, l, r, lo, hi,
                                            vartypes, sync_type, nullptr);
                      });
}

GDP_RETCODE propagateConstraintsSequentialDouble(const int n_cons, const int n_vars, const int nnz,
                                                 const int *col_indices, const int *row_indices,
                                                 const double *vals, const double *lhss, const double *rhss,
                                                 double *lbs, double *ubs, const int *vartypes) {
    return host_double(0, n_cons, n_vars, nnz, col_indices, row_indices, vals, lhss, rhss, lbs, ubs, vartypes,
                       MAX_NUM_ROUNDS, 0, nullptr);
}

GDP_RETCODE propagateConstraintsSequentialFloat(const int n_cons, const int n_vars, const int nnz,
                                                const int *col_indices, const int *row_indices, const float *vals,
                                                const float *lhss, const float *rhss, float *lbs, float *ubs,
                                                const int *vartypes) {
    return via_double(n_cons, n_vars, nnz, vals, lhss, rhss, lbs, ubs,
                      [&](double *v, double *l, double *r, double *lo, double *hi) {
                          return host_double(0, n_cons, n_vars, nnz, col_indices, row_indices, v, l, r, lo, hi,
                                             vartypes, MAX_NUM_ROUNDS, 0, nullptr);
                      });
}

GDP_RETCODE propagateConstraintsFullOMPDouble(const int n_cons, const int n_vars, const int nnz,
                                              const int *col_indices, const int *row_indices, const double *vals,
                                              const double *lhss, const double *rhss, double *lbs, double *ubs,
                                              const int *vartypes) {
    // cpu_omp stops after MAX_NUM_ROUNDS - 1 rounds (OMP_propagator.h:195)
    return host_double(2, n_cons, n_vars, nnz, col_indices, row_indices, vals, lhss, rhss, lbs, ubs, vartypes,
                       MAX_NUM_ROUNDS - 1, 0, nullptr);
}

GDP_RETCODE propagateConstraintsFullOMPFloat(const int n_cons, const int n_vars, const int nnz,
                                             const int *col_indices, const int *row_indices, const float *vals,
                                             const float *lhss, const float *rhss, float *lbs, float *ubs,
                                             const int *vartypes) {
    return via_double(n_cons, n_vars, nnz, vals, lhss, rhss, lbs, ubs,
                      [&](double *v, double *l, double *r, double *lo, double *hi) {
                          return host_double(2, n_cons, n_vars, nnz, col_indices, row_indices, v, l, r, lo, hi,
                                             vartypes, MAX_NUM_ROUNDS - 1, 0, nullptr);
                      });
}

GDP_RETCODE propagateConstraintsSequentialDisjointDouble(const int n_cons, const int n_vars, const int nnz,
                                                         const int *col_indices, const int *row_indices,
                                                         const double *vals, const double *lhss,
                                                         const double *rhss, double *lbs, double *ubs,
                                                         const int *vartypes) {
    return host_double(1, n_cons, n_vars, nnz, col_indices, row_indices, vals, lhss, rhss, lbs, ubs, vartypes,
                       MAX_NUM_ROUNDS, 0, nullptr);
}

GDP_RETCODE sequentialPropagateWithMeasureDouble(const int n_cons, const int n_vars, const int nnz,
                                                 const int *col_indices, const int *row_indices,
                                                 const double *vals, const double *lhss, const double *rhss,
                                                 double *lbs, double *ubs, const int *vartypes) {
    return with_measure(false, n_cons, n_vars, nnz, col_indices, row_indices, vals, lhss, rhss, lbs, ubs, vartypes,
                        MAX_NUM_ROUNDS, true, nullptr, nullptr, nullptr, nullptr);
}

GDP_RETCODE atomicPropagateWithMeasureDouble(const int n_cons, const int n_vars, const int nnz,
                                             const int *col_indices, const int *row_indices, const double *vals,
                                             const double *lhss, const double *rhss, double *lbs, double *ubs,
                                             const int *vartypes) {
    return with_measure(true, n_cons, n_vars, nnz, col_indices, row_indices, vals, lhss, rhss, lbs, ubs, vartypes,
                        MAX_NUM_ROUNDS, true, nullptr, nullptr, nullptr, nullptr);
}

// ---- extension API (gdp_ext.h) ----

GDP_RETCODE gdpPropagateWithMeasure(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                                    const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                                    double *lbs, double *ubs, const int *vartypes, const int method,
                                    const int max_rounds, const int print, double *p_fin, double *p_inf,
                                    int *n_rounds, gdp_run_info *info) {
    if (method != GDP_MEASURE_CPU_SEQ && method != GDP_MEASURE_GPU) {
        gdp::set_last_error("unknown measure method");
        return GDP_ERROR;
    }
    return with_measure(method == GDP_MEASURE_GPU, n_cons, n_vars, nnz, col_indices, row_ptrs, vals, lhss, rhss,
                        lbs, ubs, vartypes, max_rounds, print != 0, p_fin, p_inf, n_rounds, info);
}

GDP_RETCODE gdpWeakestBounds(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                             const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                             double *lbs, double *ubs, const int *vartypes) {
    if (!sizes_ok(n_cons, n_vars, nnz)) return GDP_ERROR;
    if (empty_problem(n_cons, n_vars, nnz)) return GDP_OKAY;
    try {
        gdp::weakest_bounds(n_cons, n_vars, col_indices, row_ptrs, vals, lhss, rhss, vartypes, lbs, ubs);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("weakest bounds", e);
    }
}

void gdpTuningDefaults(gdp_tuning *t) {
    if (!t) return;
    const gdp::Tuning d;
    t->item_nnz = d.item_nnz;
    t->mega_blocks_per_cu = d.mega_blocks_per_cu;
    t->mega_threads = d.mega_threads;
    t->graph_rounds = d.graph_rounds;
    t->max_split_segs = d.max_split_segs;
    t->compact_bounds = d.compact_bounds;
    t->row_marking = d.row_marking;
    t->mark_limit = d.mark_limit;
    for (int b = 0; b < 10; b++) t->group_log[b] = d.group_log[b];
}

GDP_RETCODE gdpProblemCreate(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                             const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                             const int *vartypes, const double *env_lbs, const double *env_ubs,
                             const gdp_tuning *tuning, gdp_problem *out) {
    if (!out) return GDP_ERROR;
    *out = nullptr;
    try {
        const gdp::Tuning T = tuning_from(tuning);
        gdp::Problem *p = gdp::problem_create(n_cons, n_vars, nnz, col_indices, row_ptrs, vals, lhss, rhss,
                                              vartypes, env_lbs, env_ubs, &T);
        *out = new gdp_problem_s{p};
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpProblemCreate", e);
    }
}

GDP_RETCODE gdpProblemDestroy(gdp_problem p) {
    if (!p) return GDP_OKAY;
    try {
        gdp::problem_destroy(p->p);
    } catch (const std::exception &e) {
        return fail("gdpProblemDestroy", e);
    }
    delete p;
    return GDP_OKAY;
}

GDP_RETCODE gdpProblemGetStats(gdp_problem p, gdp_problem_stats *out) {
    if (!p || !out) return GDP_ERROR;
    const gdp::ProblemStats s = gdp::problem_stats(p->p);
    out->n_cons = s.m;
    out->n_vars = s.n;
    out->nnz = s.nnz;
    out->n_cons_nonempty = s.m_nonempty;
    out->n_items = s.n_items;
    out->n_long_rows = s.n_long_rows;
    for (int b = 0; b < 11; b++) {
        out->bin_rows[b] = s.bin_rows[b];
        out->bin_nnz[b] = s.bin_nnz[b];
    }
    out->n16 = s.n16;
    out->n32 = s.n32;
    out->nf = s.nf;
    out->n_units = s.n_units;
    out->unit_state_bytes = s.unit_state_bytes;
    out->mega_threads = s.mega_threads;
    out->mega_grid = s.mega_grid;
    out->run_state_bytes = static_cast<long long>(gdp::problem_run_state_bytes(p->p));
    out->n_binary = s.n_bin;
    out->n_lds_binary = s.n_lds;
    return GDP_OKAY;
}

GDP_RETCODE gdpProblemReserve(gdp_problem p, const int capacity) {
    if (!p) return GDP_ERROR;
    try {
        gdp::problem_reserve(p->p, capacity);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpProblemReserve", e);
    }
}

GDP_RETCODE gdpPropagateNodesDevice(gdp_problem p, const int count, double *const *d_lbs, double *const *d_ubs,
                                    const int *hint_ptr, const int *hints, const int sync_type, const int max_rounds,
                                    gdp_run_info *infos) {
    if (!p || !d_lbs || !d_ubs || count < 1) return GDP_ERROR;
    try {
        std::vector<RunResult> r(count);
        gdp::nodes_propagate_device(p->p, count, d_lbs, d_ubs, hint_ptr, hints, sync_type, max_rounds, r.data());
        if (infos)
            for (int k = 0; k < count; k++) fill_info(infos + k, r[k].rounds, r[k].infeasible, r[k].capped, r[k].kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpPropagateNodesDevice", e);
    }
}

GDP_RETCODE gdpBatchCreate(const int n_inst, const int *n_cons, const int *n_vars, const int *nnz,
                           const int *const *col_indices, const int *const *row_ptrs, const double *const *vals,
                           const double *const *lhss, const double *const *rhss, const int *const *vartypes,
                           const double *const *env_lbs, const double *const *env_ubs, const gdp_tuning *tuning,
                           gdp_problem *out) {
    if (!out || n_inst < 1 || !n_cons || !n_vars || !nnz || !col_indices || !row_ptrs || !vals || !lhss || !rhss ||
        !vartypes)
        return GDP_ERROR;
    *out = nullptr;
    try {
        std::vector<gdp::InstanceView> v(n_inst);
        for (int i = 0; i < n_inst; i++) {
            v[i].m = n_cons[i];
            v[i].n = n_vars[i];
            v[i].nnz = nnz[i];
            v[i].col = col_indices[i];
            v[i].row_ptr = row_ptrs[i];
            v[i].val = vals[i];
            v[i].lhs = lhss[i];
            v[i].rhs = rhss[i];
            v[i].vartype = vartypes[i];
            v[i].env_lb = env_lbs ? env_lbs[i] : nullptr;
            v[i].env_ub = env_ubs ? env_ubs[i] : nullptr;
        }
        const gdp::Tuning T = tuning_from(tuning);
        *out = new gdp_problem_s{gdp::batch_create(v, &T)};
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpBatchCreate", e);
    }
}

GDP_RETCODE gdpBatchPropagateDevice(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                                    const int max_rounds, gdp_run_info *infos) {
    if (!p || !d_lbs || !d_ubs) return GDP_ERROR;
    try {
        const int ni = gdp::problem_stats(p->p).n_units;
        std::vector<RunResult> r(ni);
        gdp::batch_propagate_device(p->p, d_lbs, d_ubs, sync_type, max_rounds, r.data());
        if (infos)
            for (int k = 0; k < ni; k++) fill_info(infos + k, r[k].rounds, r[k].infeasible, r[k].capped, r[k].kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpBatchPropagateDevice", e);
    }
}

GDP_RETCODE gdpBatchPropagate(gdp_problem p, double *lbs, double *ubs, const int sync_type, const int max_rounds,
                              gdp_run_info *infos) {
    if (!p || !lbs || !ubs) return GDP_ERROR;
    double *d = nullptr;
    const size_t n = static_cast<size_t>(gdp::problem_stats(p->p).n);
    if (hipMalloc(&d, 2 * n * sizeof(double)) != hipSuccess) {
        gdp::set_last_error("gdpBatchPropagate: hipMalloc failed");
        return GDP_ERROR;
    }
    GDP_RETCODE rc = GDP_ERROR;
    if (hipMemcpy(d, lbs, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d + n, ubs, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
        rc = gdpBatchPropagateDevice(p, d, d + n, sync_type, max_rounds, infos);
        if (rc == GDP_OKAY && (hipMemcpy(lbs, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                               hipMemcpy(ubs, d + n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)) {
            gdp::set_last_error("gdpBatchPropagate: copy back failed");
            rc = GDP_ERROR;
        }
    } else {
        gdp::set_last_error("gdpBatchPropagate: upload failed");
    }
    (void)hipFree(d);
    return rc;
}

GDP_RETCODE gdpAutotune(const int n_cons, const int n_vars, const int nnz, const int *col_indices, const int *row_ptrs,
                        const double *vals, const double *lhss, const double *rhss, const int *vartypes,
                        const double *lbs, const double *ubs, const int sync_type, gdp_tuning *tuning,
                        double *best_ms) {
    if (!tuning || !lbs || !ubs) return GDP_ERROR;
    gdp::InstanceView v;
    v.m = n_cons;
    v.n = n_vars;
    v.nnz = nnz;
    v.col = col_indices;
    v.row_ptr = row_ptrs;
    v.val = vals;
    v.lhs = lhss;
    v.rhs = rhss;
    v.vartype = vartypes;
    v.env_lb = lbs;
    v.env_ub = ubs;
    return autotune_views({v}, lbs, ubs, sync_type, tuning, best_ms, "gdpAutotune");
}

GDP_RETCODE gdpBatchAutotune(const int n_inst, const int *n_cons, const int *n_vars, const int *nnz,
                             const int *const *col_indices, const int *const *row_ptrs, const double *const *vals,
                             const double *const *lhss, const double *const *rhss, const int *const *vartypes,
                             const double *lbs, const double *ubs, const int sync_type, gdp_tuning *tuning,
                             double *best_ms) {
    if (!tuning || !lbs || !ubs || n_inst < 1 || !n_cons || !n_vars || !nnz) return GDP_ERROR;
    std::vector<gdp::InstanceView> v(n_inst);
    size_t off = 0;
    for (int i = 0; i < n_inst; i++) {
        v[i].m = n_cons[i];
        v[i].n = n_vars[i];
        v[i].nnz = nnz[i];
        v[i].col = col_indices[i];
        v[i].row_ptr = row_ptrs[i];
        v[i].val = vals[i];
        v[i].lhs = lhss[i];
        v[i].rhs = rhss[i];
        v[i].vartype = vartypes[i];
        v[i].env_lb = lbs + off;
        v[i].env_ub = ubs + off;
        off += static_cast<size_t>(n_vars[i]);
    }
    return autotune_views(v, lbs, ubs, sync_type, tuning, best_ms, "gdpBatchAutotune");
}

struct gdp_server_s {
    gdp::Server *s;
};

GDP_RETCODE gdpServerCreate(gdp_problem p, const long long hbm_budget, const int max_slots, gdp_server *out) {
    if (!p || !out) return GDP_ERROR;
    *out = nullptr;
    try {
        *out = new gdp_server_s{gdp::server_create(p->p, hbm_budget, max_slots)};
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerCreate", e);
    }
}

GDP_RETCODE gdpServerDestroy(gdp_server s) {
    if (!s) return GDP_OKAY;
    gdp::server_destroy(s->s);
    delete s;
    return GDP_OKAY;
}

GDP_RETCODE gdpServerSlots(gdp_server s, int *slots) {
    if (!s || !slots) return GDP_ERROR;
    *slots = gdp::server_slots(s->s);
    return GDP_OKAY;
}

GDP_RETCODE gdpServerSubmit(gdp_server s, const double *lbs, const double *ubs, const int *hints, const int n_hints,
                            const int max_rounds, long long *ticket) {
    if (!s || !lbs || !ubs || !ticket) return GDP_ERROR;
    try {
        *ticket = gdp::server_submit(s->s, lbs, ubs, hints, n_hints, max_rounds);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerSubmit", e);
    }
}

GDP_RETCODE gdpServerSubmitPinned(gdp_server s, const double *lbs, const double *ubs, double *out_lbs,
                                  double *out_ubs, const int *hints, const int n_hints, const int max_rounds,
                                  long long *ticket) {
    if (!s || !lbs || !ubs || !out_lbs || !out_ubs || !ticket) return GDP_ERROR;
    try {
        *ticket = gdp::server_submit_pinned(s->s, lbs, ubs, out_lbs, out_ubs, hints, n_hints, max_rounds);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerSubmitPinned", e);
    }
}

GDP_RETCODE gdpServerSetBase(gdp_server s, const double *lbs, const double *ubs) {
    if (!s || !lbs || !ubs) return GDP_ERROR;
    try {
        gdp::server_set_base(s->s, lbs, ubs);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerSetBase", e);
    }
}

GDP_RETCODE gdpServerSubmitDelta(gdp_server s, const int n_changes, const int *cols, const double *lbs,
                                 const double *ubs, const int max_rounds, long long *ticket) {
    if (!s || !ticket) return GDP_ERROR;
    try {
        *ticket = gdp::server_submit_delta(s->s, n_changes, cols, lbs, ubs, max_rounds);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerSubmitDelta", e);
    }
}

GDP_RETCODE gdpServerStep(gdp_server s, int *completed) {
    if (!s) return GDP_ERROR;
    try {
        const int c = gdp::server_step(s->s);
        if (completed) *completed = c;
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerStep", e);
    }
}

GDP_RETCODE gdpServerPending(gdp_server s, int *pending) {
    if (!s || !pending) return GDP_ERROR;
    *pending = gdp::server_pending(s->s);
    return GDP_OKAY;
}

GDP_RETCODE gdpServerCompleted(gdp_server s, long long *tickets, const int cap, int *count) {
    if (!s || !count || (cap > 0 && !tickets)) return GDP_ERROR;
    if (gdp::server_running(s->s)) {
        try {
            std::vector<int64_t> t(std::max(0, cap));
            const int n = gdp::server_take_completed(s->s, t.data(), cap);
            for (int k = 0; k < n; k++) tickets[k] = t[k];
            *count = n;
            return GDP_OKAY;
        } catch (const std::exception &e) {
            return fail("gdpServerCompleted", e);
        }
    }
    const std::vector<int64_t> &c = gdp::server_completed(s->s);
    const int n = static_cast<int>(c.size());
    for (int k = 0; k < n && k < cap; k++) tickets[k] = c[k];
    *count = n;
    return GDP_OKAY;
}

GDP_RETCODE gdpServerStart(gdp_server s) {
    if (!s) return GDP_ERROR;
    try {
        gdp::server_start(s->s);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerStart", e);
    }
}

GDP_RETCODE gdpServerStop(gdp_server s) {
    if (!s) return GDP_ERROR;
    try {
        gdp::server_stop(s->s);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerStop", e);
    }
}

GDP_RETCODE gdpServerResultChanges(gdp_server s, const long long ticket, int *cols, double *lbs, double *ubs,
                                   const int cap, int *count, gdp_run_info *info, int *ready) {
    if (!s || !count || !ready || cap < 0 || (cap > 0 && (!cols || !lbs || !ubs))) return GDP_ERROR;
    try {
        std::vector<int> c;
        std::vector<double> l, u;
        RunResult r;
        *ready = gdp::server_result_changes(s->s, ticket, c, l, u, &r);
        *count = 0;
        if (*ready) {
            *count = static_cast<int>(c.size());
            const size_t k = std::min<size_t>(c.size(), static_cast<size_t>(cap));
            std::copy(c.begin(), c.begin() + k, cols);
            std::copy(l.begin(), l.begin() + k, lbs);
            std::copy(u.begin(), u.begin() + k, ubs);
            fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        }
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpServerResultChanges", e);
    }
}

GDP_RETCODE gdpServerResult(gdp_server s, const long long ticket, double *lbs, double *ubs, gdp_run_info *info,
                            int *ready) {
    if (!s || !ready || (!lbs) != (!ubs)) return GDP_ERROR;
    RunResult r;
    *ready = gdp::server_result(s->s, ticket, lbs, ubs, &r) ? 1 : 0;
    if (*ready) fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
    return GDP_OKAY;
}

GDP_RETCODE gdpPropagate(gdp_problem p, double *lbs, double *ubs, const int sync_type, const int max_rounds,
                         gdp_run_info *info) {
    if (!p || !lbs || !ubs) return GDP_ERROR;
    try {
        const RunResult r = gdp::problem_propagate(p->p, lbs, ubs, sync_type, max_rounds);
        fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpPropagate", e);
    }
}

GDP_RETCODE gdpProblemWorkProfile(gdp_problem p, double *lbs, double *ubs, const int max_rounds, long long *rows,
                                  long long *nnz, long long *cols, long long *accepted, int *n_rounds,
                                  gdp_run_info *info) {
    if (!p || !lbs || !ubs || !rows || !nnz || !cols || !accepted || !n_rounds) return GDP_ERROR;
    try {
        std::vector<gdp::RoundWork> work;
        const RunResult r = gdp::problem_work_profile(p->p, lbs, ubs, max_rounds, work);
        for (size_t k = 0; k < work.size(); k++) {
            rows[k] = work[k].rows;
            nnz[k] = work[k].nnz;
            cols[k] = work[k].cols;
            accepted[k] = work[k].accepted;
        }
        *n_rounds = static_cast<int>(work.size());
        fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpProblemWorkProfile", e);
    }
}

GDP_RETCODE gdpPropagateDevice(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                               const int max_rounds, gdp_run_info *info) {
    if (!p || !d_lbs || !d_ubs) return GDP_ERROR;
    try {
        const RunResult r = gdp::problem_propagate_device(p->p, d_lbs, d_ubs, sync_type, max_rounds);
        fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpPropagateDevice", e);
    }
}

GDP_RETCODE gdpPropagateDeviceNoAlloc(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                                      const int max_rounds, gdp_run_info *info) {
    static_assert(GDP_NOALLOC_SCRATCH_BYTES == static_cast<long long>(gdp::kScratchBytes),
                  "the header states the size of the device array");
    if (!p || !d_lbs || !d_ubs) return GDP_ERROR;
    try {
        const RunResult r = gdp::problem_propagate_no_alloc(p->p, d_lbs, d_ubs, sync_type, max_rounds);
        fill_info(info, r.rounds, r.infeasible, r.capped, r.kernel_ms);
        return GDP_OKAY;
    } catch (const std::exception &e) {
        return fail("gdpPropagateDeviceNoAlloc", e);
    }
}

GDP_RETCODE gdpSequentialPropagate(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                                   const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                                   double *lbs, double *ubs, const int *vartypes, const int max_rounds,
                                   gdp_run_info *info) {
    return host_double(0, n_cons, n_vars, nnz, col_indices, row_ptrs, vals, lhss, rhss, lbs, ubs, vartypes,
                       max_rounds, 0, info);
}

GDP_RETCODE gdpOMPPropagate(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                            const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                            double *lbs, double *ubs, const int *vartypes, const int max_rounds, const int threads,
                            gdp_run_info *info) {
    return host_double(2, n_cons, n_vars, nnz, col_indices, row_ptrs, vals, lhss, rhss, lbs, ubs, vartypes,
                       max_rounds, threads, info);
}

GDP_RETCODE gdpDeviceCount(int *count) {
    if (!count) return GDP_ERROR;
    *count = 0;
    const hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) { *count = 0; return GDP_OKAY; }
    return hip_rc(e, "hipGetDeviceCount");
}

GDP_RETCODE gdpDeviceMalloc(void **ptr, const size_t bytes) {
    if (!ptr) return GDP_ERROR;
    return hip_rc(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}

GDP_RETCODE gdpDeviceFree(void *ptr) { return hip_rc(hipFree(ptr), "hipFree"); }

GDP_RETCODE gdpHostAlloc(void **ptr, const size_t bytes) {
    if (!ptr) return GDP_ERROR;
    return hip_rc(hipHostMalloc(ptr, bytes ? bytes : 1, 0), "hipHostMalloc");
}

GDP_RETCODE gdpHostFree(void *ptr) { return hip_rc(hipHostFree(ptr), "hipHostFree"); }

// hipMemcpy may return before a device-to-device copy has finished; these
// helpers promise completion, since the propagation runs on its own stream.
GDP_RETCODE gdpMemcpyHtoD(void *dst, const void *src, const size_t bytes) {
    const GDP_RETCODE rc = hip_rc(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return rc == GDP_OKAY ? hip_rc(hipDeviceSynchronize(), "hipDeviceSynchronize") : rc;
}

GDP_RETCODE gdpMemcpyDtoH(void *dst, const void *src, const size_t bytes) {
    const GDP_RETCODE rc = hip_rc(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return rc == GDP_OKAY ? hip_rc(hipDeviceSynchronize(), "hipDeviceSynchronize") : rc;
}

GDP_RETCODE gdpMemcpyDtoD(void *dst, const void *src, const size_t bytes) {
    const GDP_RETCODE rc = hip_rc(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    return rc == GDP_OKAY ? hip_rc(hipDeviceSynchronize(), "hipDeviceSynchronize") : rc;
}

GDP_RETCODE gdpDeviceName(char *buf, const int len) {
    if (!buf || len <= 0) return GDP_ERROR;
    int dev = 0;
    hipDeviceProp_t prop;
    GDP_RETCODE rc = hip_rc(hipGetDevice(&dev), "hipGetDevice");
    if (rc != GDP_OKAY) return rc;
    rc = hip_rc(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (rc != GDP_OKAY) return rc;
    std::snprintf(buf, static_cast<size_t>(len), "%s %s cu=%d", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
    return GDP_OKAY;
}

const char *gdpLastError(void) {
    thread_local std::string s;
    s = gdp::last_error();
    return s.c_str();
}

const char *gdpVersion(void) { return "gdp-amd 0.1 gfx950 fp64"; }

}  // extern "C"
