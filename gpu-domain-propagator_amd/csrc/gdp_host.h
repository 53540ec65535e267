// gdp_host.h - host propagators of libGpuProp.so (cpu_seq, cpu_seq_dis,
// cpu_omp).  Plain C++; no HIP dependency.
#pragma once

namespace gdp {

struct HostResult {
    int rounds = 0;
    int infeasible = 0;
    int capped = 0;
};

HostResult seq_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds);

HostResult seq_dis_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                             const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                             int max_rounds);

// threads <= 0: GDP_OMP_THREADS from the environment, else 8.
HostResult omp_propagate(int m, int n, const int *col, const int *row_ptr, const double *val,
                         const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype,
                         int max_rounds, int threads);

int omp_default_threads();

}  // namespace gdp
