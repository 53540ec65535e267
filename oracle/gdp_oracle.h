/*
 * gdp_oracle.h - CPU oracle for the domain-propagation parity tests.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, and only as the checker.
 *
 * Two drivers restate the reference's propagation in plain C:
 *   oracle_seq_propagate     cpu_seq, the Gauss-Seidel fixpoint of
 *                            src/propagators/sequential_propagator.h:13-123,
 *                            :207-296 with the scalar arithmetic of
 *                            src/propagation_methods.cuh:40-276.
 *   oracle_jacobi_propagate  the same arithmetic in round-synchronous form:
 *                            every row of round r reads the bounds as they
 *                            were at the end of round r-1, and the new bound
 *                            of a variable is the best accepted candidate.
 *                            Row activities use the canonical summation order
 *                            (oracle_canonical_sum), which is the order the
 *                            GPU kernels implement, so GPU and this driver
 *                            agree bit for bit, round by round.
 *
 * Both drivers stop at the end of the first round after which some variable
 * has lb > ub + GDP_EPS (infeasible) when detect_infeasible is nonzero.
 */
#ifndef GDP_ORACLE_H
#define GDP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* results of a propagation run */
typedef struct {
    int rounds;      /* executed rounds, counting the final no-change round */
    int infeasible;  /* 1 if bounds crossed */
    int capped;      /* 1 if the round cap stopped a run that still changed */
    double loop_seconds; /* monotonic-clock time of the round loop alone (the
                            reference's timing scope: no CSC build, no
                            allocation; sequential_propagator.h:248-281) */
} oracle_result;

/* ---- scalar arithmetic (reference src/propagation_methods.cuh) ---- */
double oracle_adjust_upper(int is_cont, double ub);                 /* :40-47  */
double oracle_adjust_lower(int is_cont, double lb);                 /* :49-52  */
int oracle_is_lb_better(double lb, double ub, double newlb);        /* :54-58  */
int oracle_is_ub_better(double lb, double ub, double newub);        /* :60-64  */
int oracle_tighten_var_upper(double coeff, double slack, double surplus, int num_inf,
                             double lb, double ub, int is_cont, double *newb);  /* :78-106 */
int oracle_tighten_var_lower(double coeff, double slack, double surplus, int num_inf,
                             double lb, double ub, int is_cont, double *newb);  /* :108-134 */
int oracle_can_cons_be_tightened(double minact, double maxact, int nmininf, int nmaxinf,
                                 double lhs, double rhs, double maxactdelta);   /* :136-154 */
int oracle_is_cons_infeasible(double minact, double maxact, double rhs, double lhs); /* :156-159 */
/* tightenVariable (:223-276); outputs flags and values for both bounds */
void oracle_tighten_variable(double coeff, double lhs, double rhs, double minact, double maxact,
                             int nmininf, int nmaxinf, int is_cont, double lb, double ub,
                             int *lb_t, double *newlb, int *ub_t, double *newub);

/* Sequential activities of one row (computeActivities, :161-221).
 * out[0]=minact out[1]=maxact out[2]=maxactdelta; inf[0]=min inf count,
 * inf[1]=max inf count. */
void oracle_row_activities_seq(int row, const int *col, const int *rowptr, const double *val,
                               const double *lb, const double *ub, double *out, int *inf);

/* Canonical activity order used by the GPU kernels and the Jacobi driver:
 * the row is cut into segments of seg_len consecutive entries; each segment
 * is summed by the pairwise tree over its entries in natural order (leaf k
 * pairs with leaf k^1 first, then pairs of pairs, ...; missing leaves of the
 * power-of-two padding are +0.0); segment sums are added left to right
 * starting from +0.0. */
double oracle_canonical_sum(const double *x, int n, int seg_len);
void oracle_row_activities_canon(int row, const int *col, const int *rowptr, const double *val,
                                 const double *lb, const double *ub, int seg_len,
                                 double *out, int *inf);

/* ---- fixpoint drivers ---- */
int oracle_seq_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                         const double *val, const double *lhs, const double *rhs,
                         double *lb, double *ub, const int *vartype,
                         int max_rounds, int detect_infeasible, oracle_result *res);

int oracle_jacobi_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                            const double *val, const double *lhs, const double *rhs,
                            double *lb, double *ub, const int *vartype,
                            int max_rounds, int detect_infeasible, int seg_len,
                            oracle_result *res);

/* Disjoint variant (sequential_propagator.h:125-204): activities of all rows
 * computed at the start of each round, then the Gauss-Seidel sweep uses them. */
int oracle_seq_disjoint_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                                  const double *val, const double *lhs, const double *rhs,
                                  double *lb, double *ub, const int *vartype,
                                  int max_rounds, oracle_result *res);

/* ---- progress measure (src/progressMeasureMethods/) ---- */

/* computeWeakestBounds (weakest_bounds.h:772-829): lb/ub in/out, the
 * weakest finite bounds propagation can derive for bounds that start
 * infinite.  Returns 0, or -1 when out of memory. */
int oracle_weakest_bounds(int m, int n, int nnz, const int *col, const int *rowptr, const double *val,
                          const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype);
/* calcMaxMeasureValues (progress_function.h:118-150) */
void oracle_progress_max(int n, const double *lb_in, const double *ub_in, const double *wl, const double *wu,
                         const double *ll, const double *lu, double *p_fin, int *p_inf);
/* one round's unnormalised measure (progress_function.h:45-116, :152-195);
 * returns 0 or -1 - j for a variable j the reference exits on */
int oracle_progress_round(int n, const double *lb, const double *ub, const double *wl, const double *wu,
                          const double *ll, const double *lu, const double *pl, const double *pu,
                          double *p_fin, int *p_inf);
/* normalizeTo0_100 (progress_function.h:10-16) */
double oracle_progress_normalize(double val, double max_val);

/* CSR -> CSC by counting sort (structure only; rows ascending in a column). */
void oracle_csr_to_csc(int m, int n, int nnz, const int *col, const int *rowptr,
                       const double *val, int *colptr, int *rowidx, double *cval);

#ifdef __cplusplus
}
#endif
#endif
