// gdp_engine.h - host-side API of the MI355X propagation engine (C++).
//
// A Problem owns a device-resident layout of a constraint matrix for the
// step kernels (rows grouped by length, columns relabelled by bound storage
// class, integrality folded into the column index, a column-to-row index for
// row marking) plus the device state of its propagation units:
//   * a single matrix with state for `capacity` bound vectors (replicas
//     mode): one propagation, many node vectors at once, or a server that
//     admits node requests into free slots between steps;
//   * a block-diagonal batch of independent instances (blocks mode): all
//     instances propagate in the same launches, each with its own rounds.
// Every unit is round-synchronous: bit-identical to the oracle's
// round-synchronous propagation of that unit alone.
#pragma once

#include <cstddef>
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
    int item_nnz = 0;         // target nonzeros per work item (0: 2048 for instances of
                              // at least 4M nonzeros, else 1024)
    int block_threads = 256;  // threads per workgroup (multiple of 64)
    int mega_blocks_per_cu = 0;  // 0: occupancy limit
    int mega_threads = 512;   // workgroup size of the persistent kernel: 256, 512 or 1024
    int graph_rounds = 8;     // rounds per captured graph for GPU_LOOP
    int max_split_segs = -1;  // rows with more segments run on one wave (-1: split all, 0: none)
    int compact_bounds = 1;   // 0: keep every column in 16-byte FP64 records
    int row_marking = 1;      // 0: every round processes every row of an active unit
    int mark_limit = 0;       // accepted candidates of a round from which the next round
                              // processes every row (0: max(64, m n / (4 nnz)))
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
    int n_units = 0;               // instances (batch) or unit capacity
    int64_t unit_state_bytes = 0;  // device bytes of one unit's state (replicas)
    int mega_threads = 0;          // persistent kernel: workgroup size
    int mega_grid = 0;             // and its workgroups
    int n_bin = 0;                 // binary columns (envelope in [0, 1]; single problems)
    int n_lds = 0;                 // of them, read from LDS by one-unit persistent launches
};

// One instance of a batch (caller's arrays, CSR, zero-based).
struct InstanceView {
    int m = 0, n = 0, nnz = 0;
    const int *col = nullptr, *row_ptr = nullptr;
    const double *val = nullptr, *lhs = nullptr, *rhs = nullptr;
    const int *vartype = nullptr;
    const double *env_lb = nullptr, *env_ub = nullptr;
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
// A block-diagonal batch of independent instances.
Problem *batch_create(const std::vector<InstanceView> &instances, const Tuning *tuning);
void problem_destroy(Problem *p);
ProblemStats problem_stats(const Problem *p);
// Replicas mode: device state for `capacity` bound vectors.
void problem_reserve(Problem *p, int capacity);

// Propagates lbs/ubs (host arrays of length n, in/out).
RunResult problem_propagate(Problem *p, double *lbs, double *ubs, int sync_type, int max_rounds);

// Same with device-resident bound arrays (length n each, device pointers);
// runs on the problem's stream and synchronizes before returning.
RunResult problem_propagate_device(Problem *p, double *d_lbs, double *d_ubs, int sync_type,
                                   int max_rounds);

// Same as problem_propagate_device with the run state (unit state, status
// words, launch tables) in a fixed device array of kScratchBytes that the
// library holds per device, instead of the problem's own: no device memory
// is allocated.  Calls on one device are serialised.  Throws when the
// problem's run state (problem_run_state_bytes) exceeds the array.
constexpr size_t kScratchBytes = 64u << 20;
RunResult problem_propagate_no_alloc(Problem *p, double *d_lbs, double *d_ubs, int sync_type, int max_rounds);
// Device bytes of the run state of one bound vector.
size_t problem_run_state_bytes(const Problem *p);

// Work of each round of one propagation of host vectors lbs/ubs (in/out,
// the CPU_LOOP launches, same result): the rows whose activities the round
// computes (rows marked by the previous round, or every row), their
// nonzeros, the distinct columns they hold, and the candidates it accepted
// (bound changes folded into the accumulator).  Diagnostic: reads the row
// marks back after every round.
struct RoundWork {
    int64_t rows = 0, nnz = 0, cols = 0, accepted = 0;
};
RunResult problem_work_profile(Problem *p, double *lbs, double *ubs, int max_rounds, std::vector<RoundWork> &work);

// Progress measure of every round of one propagation of host vectors
// lbs/ubs (in/out, the CPU_LOOP launches, same result): after each step a
// kernel compares the round's bounds with the weakest bounds wl/wu, the
// limit bounds ll/lu and the previous round's bounds (host arrays of
// length n).  fin[r], inf[r]: the unnormalised finite progress and the
// bounds that became finite in round r + 1.
RunResult problem_propagate_measured(Problem *p, double *lbs, double *ubs, int max_rounds, const double *wl,
                                     const double *wu, const double *ll, const double *lu, std::vector<double> &fin,
                                     std::vector<int> &inf);

// Replicas mode: `count` device-resident bound vectors at once (vector k's
// at d_lbs[k], d_ubs[k], host arrays of device pointers).  With hints (host
// CSR: vector k's columns hints[hint_ptr[k] .. hint_ptr[k+1])), the vectors
// differ from a propagated fixpoint only in those columns and round 1
// processes only their rows.  Results per vector in out[k]; the kernel time
// of the whole launch in each out[k].kernel_ms.
void nodes_propagate_device(Problem *p, int count, double *const *d_lbs, double *const *d_ubs,
                            const int *hint_ptr, const int *hints, int sync_type, int max_rounds,
                            RunResult *out);

// Blocks mode: every instance at once; d_lbs/d_ubs concatenate the
// instances' bound vectors (device).  Results per instance in out[i].
void batch_propagate_device(Problem *p, double *d_lbs, double *d_ubs, int sync_type, int max_rounds,
                            RunResult *out);

// Autotuning: builds the problem (one instance, or a batch of several) with
// candidate tunings and times the propagation of lbs/ubs (host, the
// instances' vectors concatenated) to the fixpoint; a coordinate search over
// the work-item size, the lanes per row of each tree-width bin, the
// persistent grid size (MEGAKERNEL) and row marking, starting from `start`
// (may be null).  Every trial is bit-identical; only the time differs.
struct AutotuneTrial {
    Tuning tuning;
    double kernel_ms = 0.0;
};
Tuning autotune(const std::vector<InstanceView> &instances, const double *lbs, const double *ubs, int sync_type,
                const Tuning *start, std::vector<AutotuneTrial> *trials);

// Serving: node requests against a device-resident matrix (replicas mode).
// Requests are queued by submit; each step admits queued requests into free
// slots (up to the slot count the HBM budget allows), advances every active
// slot by one round in one set of launches, and retires finished slots,
// whose bounds are copied back.  Requests admitted in different steps share
// the device: continuous batching.
struct Server;
Server *server_create(Problem *p, int64_t hbm_budget, int max_slots);
void server_destroy(Server *s);
int server_slots(const Server *s);
// Copies the request (full bound vectors; hints: columns that differ from a
// propagated fixpoint, may be empty); returns its ticket.
int64_t server_submit(Server *s, const double *lbs, const double *ubs, const int *hints, int n_hints,
                      int max_rounds);
// Zero-copy request: lbs/ubs and out_lbs/out_ubs are pinned host buffers
// owned by the caller and left untouched until the ticket completes; the
// vectors are transferred from lbs/ubs at admission and the result straight
// into out_lbs/out_ubs (server_result then copies nothing).  In place
// (out_lbs == lbs, out_ubs == ubs), only the columns that changed come back
// over PCIe and are written into the buffers.
int64_t server_submit_pinned(Server *s, const double *lbs, const double *ubs, double *out_lbs, double *out_ubs,
                             const int *hints, int n_hints, int max_rounds);
// Base vector of delta requests (kept in device memory), e.g. the
// propagated root of a branch-and-bound tree.
void server_set_base(Server *s, const double *lbs, const double *ubs);
// A request given as the base vector with `n_changes` columns set to new
// bounds; no bound vector crosses PCIe.  The base must be a propagated
// fixpoint: round 1 processes only the rows of the changed columns.
int64_t server_submit_delta(Server *s, int n_changes, const int *cols, const double *lbs, const double *ubs,
                            int max_rounds);
// One step; returns the number of requests completed by it.  Not while the
// stepping thread runs.
int server_step(Server *s);
// Requests queued or running.
int server_pending(Server *s);
// Tickets completed by the last step.
const std::vector<int64_t> &server_completed(const Server *s);
// The server's own stepping thread: while it runs, a step follows whenever
// requests are queued or running, and callers only submit and take results
// (from any thread).  Completed tickets are collected for
// server_take_completed.  server_stop ends the thread after its current
// step and reports an error the thread met.
void server_start(Server *s);
void server_stop(Server *s);
bool server_running(Server *s);
// Up to cap tickets completed by the stepping thread and not yet taken.
int server_take_completed(Server *s, int64_t *tickets, int cap);
// Completed request: copies its bounds and outcome and forgets it; false if
// the ticket is not (yet) complete.
bool server_result(Server *s, int64_t ticket, double *lbs, double *ubs, RunResult *r);
// Completed request as the columns whose bounds differ from the base vector
// (bit for bit) with their bounds; 1 if it was complete, 0 otherwise.
// Delta requests are returned this way without copying full vectors.
int server_result_changes(Server *s, int64_t ticket, std::vector<int> &cols, std::vector<double> &lbs,
                          std::vector<double> &ubs, RunResult *r);

std::string last_error();
void set_last_error(const std::string &msg);

}  // namespace gdp
