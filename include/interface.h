/*
 * interface.h - drop-in C ABI of libGpuProp.so (MI355X domain propagator).
 *
 * Every function below replaces the reference function of the same name in
 * include/interface.h of Sofranac-Boro/gpu-domain-propagator; arguments have
 * the same order and meaning:
 *
 *   n_cons, n_vars, nnz       sizes (m rows, n columns, nonzeros)
 *   col_indices[nnz]          zero-based CSR column index of each nonzero
 *   row_ptrs[n_cons + 1]      zero-based CSR row pointers
 *   vals[nnz]                 coefficients (nonzero)
 *   lhss[n_cons], rhss[...]   row sides, lhs <= a.x <= rhs; +-1e20 or IEEE
 *                             +-inf mean "absent"
 *   lbs[n_vars], ubs[...]     variable bounds, read on entry and overwritten
 *                             with the propagated bounds on return
 *   vartypes[n_vars]          GDP_BINARY / GDP_INTEGER / GDP_CONTINUOUS
 *
 * The caller owns every array; nothing is retained between calls.  Calls are
 * synchronous.  Errors (bad sizes, device failures) return GDP_ERROR and leave
 * lbs/ubs untouched; the process is never terminated.  A problem with zero
 * rows, columns or nonzeros returns GDP_OKAY without touching anything.
 * Propagation stops early when a variable's bounds cross (lb > ub + GDP_EPS):
 * the returned bounds then show the crossing.  gdp_ext.h exposes the round
 * count and the infeasibility flag explicitly.
 */
#ifndef GDP_AMD_INTERFACE_H
#define GDP_AMD_INTERFACE_H

#include "gdp_def.h"

#ifdef __cplusplus
#define GDP_DEFAULT_ARG(v) = v
extern "C" {
#else
#define GDP_DEFAULT_ARG(v)
#endif

/* Replaces include/interface.h:10-22 (propagateConstraintsGPUReductionDouble,
 * defined at src/interface.cu:18-47).  Runs the round-synchronous GPU engine;
 * the reference's separate candidate-array algorithm is not reproduced. */
GDP_RETCODE propagateConstraintsGPUReductionDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *csr_col_indices, const int *csr_row_ptrs, const double *csr_vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

/* Replaces include/interface.h:33-45 (src/interface.cu:49-78).  Computes in
 * FP64 on the device and rounds tightened bounds outward to float. */
GDP_RETCODE propagateConstraintsGPUReductionFloat(
        const int n_cons, const int n_vars, const int nnz,
        const int *csr_col_indices, const int *csr_row_ptrs, const float *csr_vals,
        const float *lhss, const float *rhss, float *lbs, float *ubs,
        const int *vartypes);

/* Replaces include/interface.h:56-69 (src/interface.cu:80-109): the hot path.
 * sync_type is a GDP_SYNCTYPE; unknown values return GDP_ERROR (reference
 * src/propagators/GPU_propagator.cuh:256-259). */
GDP_RETCODE propagateConstraintsGPUAtomicDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *csr_col_indices, const int *csr_row_ptrs, const double *csr_vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes, const int sync_type GDP_DEFAULT_ARG(0));

/* Replaces include/interface.h:79-92 (src/interface.cu:111-140).  FP64 on the
 * device; tightened bounds are rounded outward to float. */
GDP_RETCODE propagateConstraintsGPUAtomicFloat(
        const int n_cons, const int n_vars, const int nnz,
        const int *csr_col_indices, const int *csr_row_ptrs, const float *csr_vals,
        const float *lhss, const float *rhss, float *lbs, float *ubs,
        const int *vartypes, const int sync_type GDP_DEFAULT_ARG(0));

/* Replaces include/interface.h:102-116 (src/interface.cu:142-170): the
 * sequential Gauss-Seidel propagator cpu_seq on the host. */
GDP_RETCODE propagateConstraintsSequentialDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const double *vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

/* Replaces include/interface.h:126-140 (src/interface.cu:172-200). */
GDP_RETCODE propagateConstraintsSequentialFloat(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const float *vals,
        const float *lhss, const float *rhss, float *lbs, float *ubs,
        const int *vartypes);

/* Replaces include/interface.h:150-164 (src/interface.cu:202-230): the
 * OpenMP host propagator cpu_omp (SHARED_MEM_THREADS threads unless the
 * environment variable GDP_OMP_THREADS says otherwise). */
GDP_RETCODE propagateConstraintsFullOMPDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const double *vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

/* Replaces include/interface.h:174-188 (src/interface.cu:232-260). */
GDP_RETCODE propagateConstraintsFullOMPFloat(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const float *vals,
        const float *lhss, const float *rhss, float *lbs, float *ubs,
        const int *vartypes);

/* Replaces include/interface.h:198-212 (src/interface.cu:262-291): cpu_seq
 * with activities computed once per round before the sweep. */
GDP_RETCODE propagateConstraintsSequentialDisjointDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const double *vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

/* Replace include/interface.h:222-236 and :246-260 (src/interface.cu:293-353,
 * src/progressMeasureMethods/progressPropagators.cuh:11-272): propagation to
 * the fixpoint that prints a progress measure after every round
 * ("round=%d, P_fin=%.10f, P_inf=%.2f"), measured against the weakest bounds
 * of the input and the cpu_seq fixpoint.  The sequential entry propagates
 * with cpu_seq on the host; the atomic entry with the GPU's rounds, the
 * measure of each round computed on the device.  lbs/ubs receive the
 * fixpoint.  Where the reference exit()s (a finite limit upper bound whose
 * weakest bound is infinite) these return GDP_ERROR. */
GDP_RETCODE sequentialPropagateWithMeasureDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const double *vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

GDP_RETCODE atomicPropagateWithMeasureDouble(
        const int n_cons, const int n_vars, const int nnz,
        const int *col_indices, const int *row_indices, const double *vals,
        const double *lhss, const double *rhss, double *lbs, double *ubs,
        const int *vartypes);

#ifdef __cplusplus
}
#endif

#endif /* GDP_AMD_INTERFACE_H */
