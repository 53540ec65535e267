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
    double kernel_ms;  /* device time of the round loop (host paths: wall time) */
} gdp_run_info;

typedef struct {
    int item_nnz;            /* target nonzeros per work item (0: default 1024) */
    int mega_blocks_per_cu;  /* persistent grid: workgroups per CU (0: occupancy) */
    int graph_rounds;        /* rounds per captured graph in GPU_LOOP (0: 8) */
    int max_split_segs;      /* rows longer than 256 entries with up to this many
                                256-entry segments are split into one wave per
                                segment (-1: default, all; 0: never split) */
    int compact_bounds;      /* 1 (default): integer columns with a finite integral
                                envelope use 4- or 8-byte bound records; 0: FP64 only */
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

/* Same with device pointers (hipMalloc'd, length n_vars each); no host copy.
 * The library works on its own stream: work that produced d_lbs / d_ubs on
 * another stream must be complete before the call.  Returns after the device
 * has finished. */
GDP_RETCODE gdpPropagateDevice(gdp_problem p, double *d_lbs, double *d_ubs, const int sync_type,
                               const int max_rounds, gdp_run_info *info);

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

/* Device memory plumbing, so that callers need no second HIP runtime (a
 * framework bundling its own ROCm runtime cannot share a process with this
 * library's).  Synchronous; device 0 unless the caller selected another. */
GDP_RETCODE gdpDeviceCount(int *count);
GDP_RETCODE gdpDeviceMalloc(void **ptr, const size_t bytes);
GDP_RETCODE gdpDeviceFree(void *ptr);
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
