/*
 * gdp_ext.h - extension C API of libGpuProp.so (no counterpart in the
 * reference, whose calls are stateless: src/GPU_interface.cuh:11-75 uploads
 * and frees everything per call).
 *
 * A gdp_problem keeps one constraint matrix resident in device memory so that
 * any number of bound vectors can be propagated against it without
 * re-uploading, and reports the round count, infeasibility and device time
 * of every run.  All functions return GDP_OKAY or GDP_ERROR; gdpLastError()
 * gives the message of the last error on the calling thread.
 */
#ifndef GDP_AMD_EXT_H
#define GDP_AMD_EXT_H

#include <stddef.h>

#include "gdp_def.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdp_problem_s *gdp_problem;

typedef struct {
    int rounds;        /* executed rounds, counting the final no-change round */
    int infeasible;    /* 1 if some variable ended with lb > ub + GDP_EPS */
    int capped;        /* 1 if the round cap stopped a run that still changed */
    double kernel_ms;  /* device time of the round loop (host propagators: wall
                          time of the round loop alone, without the column index
                          build and the allocations, as the reference times it) */
} gdp_run_info;

typedef struct {
    int item_nnz;            /* target nonzeros per work item (0: default, 1024) */
    int mega_blocks_per_cu;  /* persistent grid: workgroups per CU (0: occupancy) */
    int mega_threads;        /* persistent kernel workgroup size: 256, 512 (default) or 1024 */
    int graph_rounds;        /* rounds per captured graph in GPU_LOOP (0: 8) */
    int max_split_segs;      /* rows longer than 256 entries with up to this many
                                256-entry segments are split into one wave per
                                segment (-1: default, all; 0: never split) */
    int compact_bounds;      /* 1 (default): integer columns with a finite integral
                                envelope use 4- or 8-byte bound records; 0: FP64 only */
    int row_marking;         /* 1 (default): a round processes only rows with a column
                                that moved in the previous round; 0: every row */
    int mark_limit;          /* accepted candidates of a round from which the next round
                                processes every row instead (0: default max(64, m n / (4 nnz))) */
    int group_log[10];       /* lanes per row (log2) for rows of width 2^b; -1: default */
} gdp_tuning;

typedef struct {
    int n_cons, n_vars, nnz;
    int n_cons_nonempty;
    int n_items;
    int n_long_rows;
    long long bin_rows[11];  /* rows per tree-width bin 2^0..2^9, then rows > 256 */
    long long bin_nnz[11];
    int n16, n32, nf;        /* columns with 2x int16, 2x int32, 2x double records */
    int n_units;             /* instances of a batch, or bound-vector capacity */
    long long unit_state_bytes; /* device bytes of one bound vector's state */
    int mega_threads;           /* persistent kernel: workgroup size */
    int mega_grid;              /* persistent kernel: workgroups */
    long long run_state_bytes;  /* device bytes of the run state of one propagation:
                                   allocated by the problem on its first propagation,
                                   taken from the fixed device array by
                                   gdpPropagateDeviceNoAlloc */
    int n_binary;               /* integer columns whose envelope lies in [0, 1]
                                   (single problems; 0 for batches) */
    int n_lds_binary;           /* of them, the columns whose bounds a persistent
                                   launch of one bound vector reads from LDS */
} gdp_problem_stats;

/* Fills *t with the defaults. */
void gdpTuningDefaults(gdp_tuning *t);

/* Uploads the matrix (same argument meaning as the functions of
 * interface.h).  env_lbs / env_ubs (may be NULL) are a bound envelope: every
 * bound vector later propagated with this problem must lie inside it (e.g.
 * the root bounds of a branch-and-bound tree); integer columns whose envelope
 * is finite and integral then get compact bound records, which makes the
 * rounds faster.  A propagation whose bounds leave the envelope returns
 * GDP_ERROR.  tuning may be NULL. */
GDP_RETCODE gdpProblemCreate(const int n_cons, const int n_vars, const int nnz,
                             const int *col_indices, const int *row_ptrs, const double *vals,
                             const double *lhss, const double *rhss, const int *vartypes,
                             const double *env_lbs, const double *env_ubs,
                             const gdp_tuning *tuning, gdp_problem *out);
GDP_RETCODE gdpProblemDestroy(gdp_problem p);
GDP_RETCODE gdpProblemGetStats(gdp_problem p, gdp_problem_stats *out);

/* Propagates host bound arrays lbs/ubs (length n_vars, in/out) to the fixpoint
 * or for at most max_rounds rounds (1..MAX_NUM_ROUNDS). info may be NULL. */
GDP_RETCODE gdpPropagate(gdp_problem p, double *lbs, double *ubs, const int sync_type,
                         const int max_rounds, gdp_run_info *info);

/* Work of each round of one propagation of host vectors lbs/ubs (in/out;
 * the CPU_LOOP launches, same result as gdpPropagate): rows[k], nnz[k] and
 * cols[k] (arrays of max_rounds entries) are the rows whose activities round
 * k+1 computes, their nonzeros and the distinct columns they hold; the
 * number of rounds run is stored in *n_rounds.  A diagnostic (it reads the
 * row marks back after every round), e.g. for the bytes a round moves.
 * accepted[k]: the bound candidates round k+1 accepted (each folded into the
 * bounds by an atomic).
 * No counterpart in the reference, whose rounds always process every row. */
GDP_RETCODE gdpProblemWorkProfile(gdp_problem p, double *lbs, double *ubs, const int max_rounds, long long *rows,
                                  long long *nnz, long long *cols, long long *accepted, int *n_rounds,
                                  gdp_run_info *info);

/* Same with device pointers (hipMalloc'd, length n_vars each); no host copy.
 * The library works on its own stream: work that produced d_lbs / d_ubs on
 * another stream must be complete before the call.  Returns after the device
 * has finished. */
GDP_RETCODE gdpPropagateDevice(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                               const int max_rounds, gdp_run_info *info);

/* Same as gdpPropagateDevice, allocating no device memory: no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The run
 * state (unit state, split-row partials, status words, counters, launch
 * tables; gdp_problem_stats.run_state_bytes) lives in one fixed array of
 * GDP_NOALLOC_SCRATCH_BYTES (64 MiB) of device memory that the library holds
 * on every device it runs on until the process ends, instead of the
 * problem's own state, which is never allocated for a problem run only this
 * way.  Problems whose run state exceeds the array return GDP_ERROR.  The
 * problem's device must be the current device.  The array is shared by all
 * problems on a device and keeps the last call's contents; every call resets
 * what it reads.  Calls on one device are serialised by the library (calls
 * that overlapped on one array would race).  Supports every sync type;
 * GPU_LOOP issues its self-stopping launches without capturing a graph. */
#define GDP_NOALLOC_SCRATCH_BYTES (64LL << 20)
GDP_RETCODE gdpPropagateDeviceNoAlloc(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                                      const int max_rounds, gdp_run_info *info);

/* ---- many bound vectors against one matrix ----
 * The problem keeps device state for `capacity` bound vectors (grown on
 * demand by gdpPropagateNodesDevice). */
GDP_RETCODE gdpProblemReserve(gdp_problem p, const int capacity);

/* Propagates `count` device-resident bound vectors at once (vector k at
 * d_lbs[k], d_ubs[k]: host arrays of device pointers) in shared launches;
 * each vector gets its own rounds and outcome in infos[k] and is
 * bit-identical to propagating it alone.  hint_ptr / hints (may be NULL):
 * vector k differs from a propagated fixpoint only in the columns
 * hints[hint_ptr[k] .. hint_ptr[k+1]), so its first round processes only their
 * rows (a branch-and-bound child of a propagated parent). */
GDP_RETCODE gdpPropagateNodesDevice(gdp_problem p, const int count, double *const *d_lbs, double *const *d_ubs,
                                    const int *hint_ptr, const int *hints, const int sync_type,
                                    const int max_rounds, gdp_run_info *infos);

/* ---- autotuning ----
 * Builds the problem with candidate tunings and times the propagation of
 * lbs/ubs to the fixpoint with sync_type: a coordinate search over the
 * work-item size, the lanes per row of each row-length bin, the persistent
 * kernel's workgroup size (MEGAKERNEL) and row marking, starting from *tuning (or the
 * defaults when it holds the defaults).  *tuning receives the fastest
 * candidate and *best_ms its device time.  The bounds double as the bound
 * envelope.  Every candidate gives the same bits. */
GDP_RETCODE gdpAutotune(const int n_cons, const int n_vars, const int nnz, const int *col_indices,
                        const int *row_ptrs, const double *vals, const double *lhss, const double *rhss,
                        const int *vartypes, const double *lbs, const double *ubs, const int sync_type,
                        gdp_tuning *tuning, double *best_ms);
/* The same for a batch (arrays as in gdpBatchCreate; lbs / ubs concatenate
 * the instances' bound vectors). */
GDP_RETCODE gdpBatchAutotune(const int n_inst, const int *n_cons, const int *n_vars, const int *nnz,
                             const int *const *col_indices, const int *const *row_ptrs,
                             const double *const *vals, const double *const *lhss, const double *const *rhss,
                             const int *const *vartypes, const double *lbs, const double *ubs,
                             const int sync_type, gdp_tuning *tuning, double *best_ms);

/* ---- a batch of independent instances in shared launches ----
 * Instance i is given by the i-th entries of the arrays (same meaning as in
 * gdpProblemCreate; env_lbs / env_ubs may be NULL, or hold NULL entries). */
GDP_RETCODE gdpBatchCreate(const int n_inst, const int *n_cons, const int *n_vars, const int *nnz,
                           const int *const *col_indices, const int *const *row_ptrs,
                           const double *const *vals, const double *const *lhss, const double *const *rhss,
                           const int *const *vartypes, const double *const *env_lbs,
                           const double *const *env_ubs, const gdp_tuning *tuning, gdp_problem *out);
/* Propagates every instance; lbs / ubs concatenate the instances' bound
 * vectors in instance order (host arrays, in/out; the Device variant takes
 * device pointers).  infos[i] (n_inst entries) gets instance i's outcome;
 * kernel_ms is the device time of the whole batch. */
GDP_RETCODE gdpBatchPropagate(gdp_problem p, double *lbs, double *ubs, const int sync_type,
                              const int max_rounds, gdp_run_info *infos);
GDP_RETCODE gdpBatchPropagateDevice(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                                    const int max_rounds, gdp_run_info *infos);

/* ---- serving node requests with continuous batching ----
 * A server admits submitted requests into free slots of the problem's
 * device state (as many slots as hbm_budget bytes allow, at most max_slots;
 * a slot takes the unit state of gdpProblemGetStats plus 64 bytes a column:
 * its bound vectors, the input of an in-place request, a staging entry and
 * a parking entry for a finished result its caller has not collected yet),
 * and each gdpServerStep advances every running request by one round in
 * shared launches: requests that arrive while others run join at the next
 * step.  A request's result is bit-identical to propagating it alone. */
typedef struct gdp_server_s *gdp_server;
GDP_RETCODE gdpServerCreate(gdp_problem p, const long long hbm_budget, const int max_slots, gdp_server *out);
GDP_RETCODE gdpServerDestroy(gdp_server s);
GDP_RETCODE gdpServerSlots(gdp_server s, int *slots);
/* Copies the request (host bound vectors of length n_vars; hints as in
 * gdpPropagateNodesDevice, may be NULL) and returns its ticket. */
GDP_RETCODE gdpServerSubmit(gdp_server s, const double *lbs, const double *ubs, const int *hints,
                            const int n_hints, const int max_rounds, long long *ticket);
/* Zero-copy request: lbs / ubs (input) and out_lbs / out_ubs (result) are
 * pinned host buffers of n_vars doubles (gdpHostAlloc) owned by the caller,
 * left untouched by it until the ticket completes.  The vectors are
 * transferred straight from lbs / ubs when the request is admitted and the
 * result straight into out_lbs / out_ubs when it completes; no host copy is
 * made.  In place (out_lbs == lbs and out_ubs == ubs) only the columns that
 * changed cross PCIe on the way back and are written into the buffers.
 * gdpServerResult with NULL buffers then returns the outcome. */
GDP_RETCODE gdpServerSubmitPinned(gdp_server s, const double *lbs, const double *ubs, double *out_lbs,
                                  double *out_ubs, const int *hints, const int n_hints, const int max_rounds,
                                  long long *ticket);
/* Base vector of delta requests, kept in device memory: a propagated
 * fixpoint, e.g. the root of a branch-and-bound tree. */
GDP_RETCODE gdpServerSetBase(gdp_server s, const double *lbs, const double *ubs);
/* A request given as the base vector with n_changes columns (cols) set to new
 * bounds (lbs, ubs); no bound vector crosses PCIe, and round 1 processes only
 * the rows of the changed columns. */
GDP_RETCODE gdpServerSubmitDelta(gdp_server s, const int n_changes, const int *cols, const double *lbs,
                                 const double *ubs, const int max_rounds, long long *ticket);
/* One step: admit, one round of every running request, retire.  An error
 * while the server's stepping thread runs. */
GDP_RETCODE gdpServerStep(gdp_server s, int *completed);
GDP_RETCODE gdpServerPending(gdp_server s, int *pending);
/* Tickets of the requests completed by the last gdpServerStep: up to cap of
 * them into tickets, their number into *count.  While the stepping thread
 * runs: up to cap of the tickets it completed since the last call, taken
 * (the rest stay for the next call), their number into *count. */
GDP_RETCODE gdpServerCompleted(gdp_server s, long long *tickets, const int cap, int *count);
/* The server's own stepping thread: while it runs, every queued or running
 * request is stepped without the caller, whose threads only submit and take
 * results (the submit, completed and result calls may come from any
 * thread).  Stop ends it after its current step and returns GDP_ERROR if the
 * thread met an error (message in gdpLastError).  Once the thread has failed,
 * the submit calls return GDP_ERROR, and so does gdpServerCompleted after
 * handing out the tickets the thread finished before.  No counterpart in the
 * reference, which has no server. */
GDP_RETCODE gdpServerStart(gdp_server s);
GDP_RETCODE gdpServerStop(gdp_server s);
/* *ready = 1 and the result copied out (and forgotten) if the request has
 * completed; *ready = 0 otherwise.  lbs and ubs may both be NULL: only the
 * outcome is returned (zero-copy requests, or results not wanted).
 * info->rounds is -1 for a request whose bounds left the problem's bound
 * envelope. */
GDP_RETCODE gdpServerResult(gdp_server s, const long long ticket, double *lbs, double *ubs,
                            gdp_run_info *info, int *ready);

/* Completed request as the columns whose bounds differ (bit for bit) from the
 * base vector of gdpServerSetBase, with their new bounds: *count of them, of
 * which min(*count, cap) are copied (cap = n_vars always suffices).  Delta
 * requests are returned this way without copying full vectors.  *ready as in
 * gdpServerResult. */
GDP_RETCODE gdpServerResultChanges(gdp_server s, const long long ticket, int *cols, double *lbs, double *ubs,
                                   const int cap, int *count, gdp_run_info *info, int *ready);

/* ---- MPS input ----
 * Parses an MPS model (fixed or free format text; decompress .gz first) into
 * the arrays of the propagation entry points: N rows dropped, L/G/E rows as
 * constraints with lhs/rhs (RANGES applied), variables in order of first
 * appearance, MARKER integer sections, bounds UP/LO/FX/FR/MI/PL/BV/LI/UI,
 * default bounds [0, +inf), infinite bounds as +-GDP_INF, types BINARY (an
 * integer in [0, 1]), INTEGER, CONTINUOUS.  Free the arrays with
 * gdpMpsFree. */
typedef struct {
    int n_cons, n_vars, nnz;
    int *row_ptrs, *col_indices;
    double *vals, *lhss, *rhss, *lbs, *ubs;
    int *vartypes;
    char name[128];
} gdp_mps;
GDP_RETCODE gdpParseMPS(const char *text, const size_t len, gdp_mps *out);
void gdpMpsFree(gdp_mps *p);

/* Host propagators with run information: cpu_seq (Gauss-Seidel), cpu_seq_dis
 * and cpu_omp (threads <= 0: GDP_OMP_THREADS or 8). */
GDP_RETCODE gdpSequentialPropagate(const int n_cons, const int n_vars, const int nnz,
                                   const int *col_indices, const int *row_ptrs, const double *vals,
                                   const double *lhss, const double *rhss, double *lbs, double *ubs,
                                   const int *vartypes, const int max_rounds, gdp_run_info *info);
GDP_RETCODE gdpOMPPropagate(const int n_cons, const int n_vars, const int nnz,
                            const int *col_indices, const int *row_ptrs, const double *vals,
                            const double *lhss, const double *rhss, double *lbs, double *ubs,
                            const int *vartypes, const int max_rounds, const int threads,
                            gdp_run_info *info);

/* ---- progress measure ----
 * The computation of sequentialPropagateWithMeasureDouble (method
 * GDP_MEASURE_CPU_SEQ) and atomicPropagateWithMeasureDouble (method
 * GDP_MEASURE_GPU) with the measures returned: p_fin[k] and p_inf[k] (arrays
 * of max_rounds entries, max_rounds in 1..100) receive round k+1's
 * normalised measures (0..100, as printed), *n_rounds the rounds run; lbs/ubs
 * in/out.  print != 0 also prints what the reference prints.
 * Reference: src/progressMeasureMethods/progressPropagators.cuh:11-272,
 * progress_function.h:10-195. */
#define GDP_MEASURE_CPU_SEQ 0
#define GDP_MEASURE_GPU 1
GDP_RETCODE gdpPropagateWithMeasure(const int n_cons, const int n_vars, const int nnz,
                                    const int *col_indices, const int *row_ptrs, const double *vals,
                                    const double *lhss, const double *rhss, double *lbs, double *ubs,
                                    const int *vartypes, const int method, const int max_rounds,
                                    const int print, double *p_fin, double *p_inf, int *n_rounds,
                                    gdp_run_info *info);
/* Weakest bounds of lbs/ubs (in/out): every bound that starts infinite gets
 * the weakest finite value the search of computeWeakestBounds
 * (src/progressMeasureMethods/weakest_bounds.h:772-829) derives for it. */
GDP_RETCODE gdpWeakestBounds(const int n_cons, const int n_vars, const int nnz,
                             const int *col_indices, const int *row_ptrs, const double *vals,
                             const double *lhss, const double *rhss, double *lbs, double *ubs,
                             const int *vartypes);

/* Device memory plumbing, so that callers need no second HIP runtime (a
 * framework bundling its own ROCm runtime cannot share a process with this
 * library's).  Synchronous; device 0 unless the caller selected another. */
GDP_RETCODE gdpDeviceCount(int *count);
GDP_RETCODE gdpDeviceMalloc(void **ptr, const size_t bytes);
GDP_RETCODE gdpDeviceFree(void *ptr);
/* Pinned (page-locked) host memory for zero-copy server requests. */
GDP_RETCODE gdpHostAlloc(void **ptr, const size_t bytes);
GDP_RETCODE gdpHostFree(void *ptr);
GDP_RETCODE gdpMemcpyHtoD(void *dst, const void *src, const size_t bytes);
GDP_RETCODE gdpMemcpyDtoH(void *dst, const void *src, const size_t bytes);
GDP_RETCODE gdpMemcpyDtoD(void *dst, const void *src, const size_t bytes);
GDP_RETCODE gdpDeviceName(char *buf, const int len);

const char *gdpLastError(void);

/* Library build identification ("gfx950 ..."). */
const char *gdpVersion(void);

#ifdef __cplusplus
}
#endif

#endif /* GDP_AMD_EXT_H */
