// gdp_host.h - host propagators of libGpuProp.so (cpu_seq, cpu_seq_dis,
// cpu_omp).  Plain C++; no HIP dependency.
#pragma once

#include <functional>

namespace gdp {

// Activities of one row in CSR order (computeActivities,
// src/propagation_methods.cuh:161-221).
struct RowActs {
    double minact, maxact, delta;
    int min_inf, max_inf;
};
RowActs row_activities(int i, const int *col, const int *row_ptr, const double *val, const double *lb,
                       const double *ub);

struct HostResult {
    int rounds = 0;
    int infeasible = 0;
    int capped = 0;
    double loop_ms = 0.0;  // wall time of the round loop alone (no CSC build or
                           // allocation: sequential_propagator.h:248-281)
};

// Called after every round with the round number and the bounds at its
// end; returning false stops the run (HostResult.rounds = that round).
using RoundObserver = std::function<bool(int round, const double *lb, const double *ub)>;

HostResult seq_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds, const RoundObserver *observer = nullptr);

HostResult seq_dis_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                             const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                             int max_rounds);

// threads <= 0: GDP_OMP_THREADS from the environment, else 8.
HostResult omp_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds, int threads);

int omp_default_threads();

// ---- progress measure (src/progressMeasureMethods/, gdp_measure.cpp) ----

// computeWeakestBounds (weakest_bounds.h:772-829): lb/ub in/out; bounds that
// start infinite receive the weakest finite value the search derives.
void weakest_bounds(int m, int n, const int *col, const int *row_ptr, const double *val, const double *lhs,
                    const double *rhs, const int *vartype, double *lb, double *ub);

struct Measure {
    double fin = 0.0;  // finite progress
    int inf = 0;       // infinite bounds made finite
};
// calcMaxMeasureValues (progress_function.h:118-150): the largest attainable
// measure of input bounds lb_in/ub_in with weakest wl/wu and limit ll/lu.
Measure measure_max(int n, const double *lb_in, const double *ub_in, const double *wl, const double *wu,
                    const double *ll, const double *lu);
// First variable with a finite limit upper bound but an infinite weakest
// upper bound (the reference stops the process on it), or -1.
int measure_defect(int n, const double *wu, const double *lu);
// One round (progress_function.h:45-116, :152-195), variables in index
// order; pl/pu are the previous round's bounds.
Measure measure_round(int n, const double *lb, const double *ub, const double *wl, const double *wu,
                      const double *ll, const double *lu, const double *pl, const double *pu);
// normalizeTo0_100 (progress_function.h:10-16)
double measure_normalize(double v, double max_v);

}  // namespace gdp
