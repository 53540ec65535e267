// gdp_engine.h - host-side API of the MI355X propagation engine (C++).
//
// A Problem owns a device-resident copy of one constraint matrix laid out for
// the round kernels (rows grouped by length, columns relabelled by bound
// storage class, integrality folded into the column index) plus the per-run
// device state (bound snapshot and accumulator, change bitmap, per-round
// status words).  It is built once and can propagate any number of bound
// vectors that lie inside its bound envelope.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gdp_arith.h"

namespace gdp {

// sync types (GDP_SYNCTYPE values of the C interface)
constexpr int CPU_LOOP_ID = 0;
constexpr int GPU_LOOP_ID = 1;
constexpr int MEGAKERNEL_ID = 2;

struct Tuning {
    int item_nnz = 1024;      // target nonzeros per work item
    int block_threads = 256;  // threads per workgroup (multiple of 64)
    int mega_blocks_per_cu = 0;  // 0: occupancy limit
    int graph_rounds = 8;     // rounds per captured graph for GPU_LOOP
    int max_split_segs = -1;  // rows with more segments run on one wave (-1: split all, 0: none)
    int compact_bounds = 1;   // 0: keep every column in 16-byte FP64 records
    // per tree-width bin (width 2^b, b = 0..9): lanes per row as log2
    // default: K = min(width, 4) consecutive entries per lane
    int group_log[10] = {0, 0, 0, 1, 2, 3, 4, 5, 6, 6};
};

struct RunResult {
    int rounds = 0;
    int infeasible = 0;
    int capped = 0;
    double kernel_ms = 0.0;   // device time of the round loop (events)
};

struct ProblemStats {
    int m = 0, n = 0, nnz = 0;
    int m_nonempty = 0;
    int n_items = 0;
    int n_long_rows = 0;
    int64_t bin_rows[11] = {0};
    int64_t bin_nnz[11] = {0};
    int n16 = 0, n32 = 0, nf = 0;  // columns per bound storage class
};

struct Problem;

// Throws std::runtime_error on invalid input or device failure.
// env_lb / env_ub (may be null) bound every vector that will be propagated
// with this problem, e.g. the root bounds of a branch-and-bound tree: integer
// columns whose envelope is finite and integral get compact bound records.
// Without an envelope every column uses FP64 records.
Problem *problem_create(int m, int n, int nnz, const int *col, const int *row_ptr,
                        const double *val, const double *lhs, const double *rhs,
                        const int *vartype, const double *env_lb, const double *env_ub,
                        const Tuning *tuning);
void problem_destroy(Problem *p);
ProblemStats problem_stats(const Problem *p);

// Propagates lbs/ubs (host arrays of length n, in/out).
RunResult problem_propagate(Problem *p, double *lbs, double *ubs, int sync_type, int max_rounds);

// Same with device-resident bound arrays (length n each, device pointers);
// runs on the problem's stream and synchronizes before returning.
RunResult problem_propagate_device(Problem *p, double *d_lbs, double *d_ubs, int sync_type,
                                   int max_rounds);

std::string last_error();
void set_last_error(const std::string &msg);

}  // namespace gdp
