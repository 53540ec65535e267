/*
 * gdp_oracle.c - CPU oracle for the domain-propagation parity tests.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the parity tests, never shipped
 * in or called by the product.  Plain C99, compiled with
 * -ffp-contract=off so every product and sum is rounded on its own, as in the
 * reference's host code.
 *
 * Each function cites the reference file:line it restates.
 */
#define _POSIX_C_SOURCE 199309L
#include "gdp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_seconds(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

#define OR_INF 1e20
#define OR_EPS 1e-9

/* src/def.h:69-79 */
static inline int epsle(double x, double y) { return (x - y) <= OR_EPS; }
static inline int epslt(double x, double y) { return (x - y) < -OR_EPS; }
static inline int epsge(double x, double y) { return (x - y) >= -OR_EPS; }
static inline int epsgt(double x, double y) { return (x - y) > OR_EPS; }
static inline int isposinf(double x) { return epsge(x, OR_INF); }
static inline int isneginf(double x) { return epsle(x, -OR_INF); }
/* std::min as used by canConsBeTightened: returns the first argument on ties */
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* scalar arithmetic: src/propagation_methods.cuh                           */
/* ------------------------------------------------------------------------ */

/* :40-47 */
double oracle_adjust_upper(int is_cont, double ub) { return is_cont ? ub : floor(ub + OR_EPS); }
/* :49-52 */
double oracle_adjust_lower(int is_cont, double lb) { return is_cont ? lb : ceil(lb - OR_EPS); }
/* :54-58 */
int oracle_is_lb_better(double lb, double ub, double newlb) { (void)ub; return epsgt(newlb, lb); }
/* :60-64 */
int oracle_is_ub_better(double lb, double ub, double newub) { (void)lb; return epslt(newub, ub); }

/* :78-106 */
int oracle_tighten_var_upper(double coeff, double slack, double surplus, int num_inf,
                             double lb, double ub, int is_cont, double *newb) {
    double nb;
    *newb = ub;
    if (num_inf == 0) {
        nb = coeff > 0 ? slack / coeff : surplus / coeff;
        nb += lb;
    } else if (num_inf == 1 && isneginf(lb)) {
        nb = coeff > 0 ? slack / coeff : surplus / coeff;
    } else {
        return 0;
    }
    nb = oracle_adjust_upper(is_cont, nb);
    if (oracle_is_ub_better(lb, ub, nb)) {
        *newb = nb;
        return 1;
    }
    return 0;
}

/* :108-134 */
int oracle_tighten_var_lower(double coeff, double slack, double surplus, int num_inf,
                             double lb, double ub, int is_cont, double *newb) {
    double nb;
    *newb = lb;
    if (num_inf == 0) {
        nb = coeff > 0 ? surplus / coeff : slack / coeff;
        nb += ub;
    } else if (num_inf == 1 && isposinf(ub)) {
        nb = coeff > 0 ? surplus / coeff : slack / coeff;
    } else {
        return 0;
    }
    nb = oracle_adjust_lower(is_cont, nb);
    if (oracle_is_lb_better(lb, ub, nb)) {
        *newb = nb;
        return 1;
    }
    return 0;
}

/* :136-154 */
int oracle_can_cons_be_tightened(double minact, double maxact, int nmininf, int nmaxinf,
                                 double lhs, double rhs, double maxactdelta) {
    if (nmininf > 1 && nmaxinf > 1) return 0;
    if (epslt(maxactdelta, OR_INF)) return !epsle(maxactdelta, dmin(rhs - minact, maxact - lhs));
    return 1;
}

/* :156-159 */
int oracle_is_cons_infeasible(double minact, double maxact, double rhs, double lhs) {
    return (minact > rhs || maxact < lhs);
}

/* :223-276 */
void oracle_tighten_variable(double coeff, double lhs, double rhs, double minact, double maxact,
                             int nmininf, int nmaxinf, int is_cont, double lb, double ub,
                             int *lb_t, double *newlb, int *ub_t, double *newub) {
    double slack = rhs - minact;
    double surplus = lhs - maxact;
    *lb_t = 0; *newlb = lb;
    *ub_t = 0; *newub = ub;
    if (coeff > 0.0) {
        if (epsgt(coeff * (ub - lb), rhs - minact) && !isposinf(rhs) && !isneginf(minact))
            *ub_t = oracle_tighten_var_upper(coeff, slack, surplus, nmininf, lb, ub, is_cont, newub);
        if (epsgt(coeff * (ub - lb), maxact - lhs) && !isneginf(lhs) && !isposinf(maxact))
            *lb_t = oracle_tighten_var_lower(coeff, slack, surplus, nmaxinf, lb, ub, is_cont, newlb);
    } else {
        if (epsgt(coeff * (lb - ub), rhs - minact) && !isposinf(rhs) && !isneginf(minact))
            *lb_t = oracle_tighten_var_lower(coeff, slack, surplus, nmininf, lb, ub, is_cont, newlb);
        if (epsgt(coeff * (lb - ub), maxact - lhs) && !isneginf(lhs) && !isposinf(maxact))
            *ub_t = oracle_tighten_var_upper(coeff, slack, surplus, nmaxinf, lb, ub, is_cont, newub);
    }
}

/* computeActivities, :161-221: sequential in CSR order. */
void oracle_row_activities_seq(int row, const int *col, const int *rowptr, const double *val,
                               const double *lb, const double *ub, double *out, int *inf) {
    double minact = 0.0, maxact = 0.0, delta_max = 0.0;
    int mininf = 0, maxinf = 0;
    for (int k = rowptr[row]; k < rowptr[row + 1]; k++) {
        double a = val[k], l = lb[col[k]], u = ub[col[k]];
        double d = fabs(a) * (u - l);
        if (epsgt(d, delta_max)) delta_max = d;
        int imin = epsgt(a, 0) ? isneginf(l) : isposinf(u);
        int imax = epsgt(a, 0) ? isposinf(u) : isneginf(l);
        mininf += imin;
        maxinf += imax;
        if (imin == 0) minact += a > 0 ? a * l : a * u;
        if (imax == 0) maxact += a > 0 ? a * u : a * l;
    }
    out[0] = minact; out[1] = maxact; out[2] = delta_max;
    inf[0] = mininf; inf[1] = maxinf;
}

/* ------------------------------------------------------------------------ */
/* canonical summation order (no reference counterpart: the order the GPU    */
/* kernels implement; see gdp_oracle.h)                                      */
/* ------------------------------------------------------------------------ */

/* pairwise tree over `width` (a power of two) leaves, the first n present */
static double pairwise(const double *x, int n, int width) {
    if (width == 1) return n >= 1 ? x[0] : 0.0;
    int h = width / 2;
    int nl = n < h ? n : h;
    int nr = n - h > 0 ? n - h : 0;
    double left = pairwise(x, nl, h);
    double right = pairwise(x + h, nr, h);
    return left + right;
}

static int next_pow2(int n) {
    int w = 1;
    while (w < n) w <<= 1;
    return w;
}

double oracle_canonical_sum(const double *x, int n, int seg_len) {
    double acc = 0.0;
    for (int s = 0; s < n; s += seg_len) {
        int len = n - s < seg_len ? n - s : seg_len;
        acc = acc + pairwise(x + s, len, next_pow2(len));
    }
    return acc;
}

/* see oracle_row_activities_canon; x buffers hold the row's contributions */
static void row_activities_canon_buf(int row, const int *col, const int *rowptr, const double *val,
                                     const double *lb, const double *ub, int seg_len,
                                     double *out, int *inf, double *xmin, double *xmax) {
    int b = rowptr[row], L = rowptr[row + 1] - b;
    double delta_max = 0.0;
    int mininf = 0, maxinf = 0;
    for (int k = 0; k < L; k++) {
        double a = val[b + k], l = lb[col[b + k]], u = ub[col[b + k]];
        double d = fabs(a) * (u - l);
        if (epsgt(d, delta_max)) delta_max = d;
        int imin = epsgt(a, 0) ? isneginf(l) : isposinf(u);
        int imax = epsgt(a, 0) ? isposinf(u) : isneginf(l);
        mininf += imin;
        maxinf += imax;
        xmin[k] = imin ? 0.0 : (a > 0 ? a * l : a * u);
        xmax[k] = imax ? 0.0 : (a > 0 ? a * u : a * l);
    }
    out[0] = oracle_canonical_sum(xmin, L, seg_len);
    out[1] = oracle_canonical_sum(xmax, L, seg_len);
    out[2] = delta_max;
    inf[0] = mininf; inf[1] = maxinf;
}

/* Activities of one row in the canonical order.  Contribution terms and
 * infinity tests are those of computeActivities (:198-213): an infinite
 * contribution counts in the infinity counter and enters the sum as +0.0.
 * The row's largest |a|(u-l) is cpu_seq's sequential epsilon-maximum. */
void oracle_row_activities_canon(int row, const int *col, const int *rowptr, const double *val,
                                 const double *lb, const double *ub, int seg_len,
                                 double *out, int *inf) {
    int L = rowptr[row + 1] - rowptr[row];
    double *buf = (double *)malloc((size_t)(L > 0 ? 2 * L : 2) * sizeof(double));
    row_activities_canon_buf(row, col, rowptr, val, lb, ub, seg_len, out, inf, buf, buf + (L > 0 ? L : 1));
    free(buf);
}

/* ------------------------------------------------------------------------ */
/* CSR -> CSC                                                               */
/* ------------------------------------------------------------------------ */

/* Plays the role of csr_to_csc (src/kernels/util_kernels.cuh:92-152), which
 * the reference runs through cuSPARSE; only the column structure is used. */
void oracle_csr_to_csc(int m, int n, int nnz, const int *col, const int *rowptr,
                       const double *val, int *colptr, int *rowidx, double *cval) {
    (void)nnz;
    for (int j = 0; j <= n; j++) colptr[j] = 0;
    for (int i = 0; i < m; i++)
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) colptr[col[k] + 1]++;
    for (int j = 0; j < n; j++) colptr[j + 1] += colptr[j];
    int *fill = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int j = 0; j < n; j++) fill[j] = colptr[j];
    for (int i = 0; i < m; i++)
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int p = fill[col[k]]++;
            rowidx[p] = i;
            if (cval) cval[p] = val[k];
        }
    free(fill);
}

/* markConstraints, src/propagation_methods.cu:3-13 */
static void mark_rows(int j, const int *colptr, const int *rowidx, int *marked) {
    for (int p = colptr[j]; p < colptr[j + 1]; p++) marked[rowidx[p]] = 1;
}

static int any_crossed(int n, const double *lb, const double *ub, const unsigned char *only) {
    for (int j = 0; j < n; j++)
        if ((!only || only[j]) && epsgt(lb[j], ub[j])) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* cpu_seq: src/propagators/sequential_propagator.h                         */
/* ------------------------------------------------------------------------ */

/* One Gauss-Seidel round, :13-123.  acts_given != NULL selects the disjoint
 * variant (activities precomputed for the round, :125-204). */
static int seq_round(int m, const int *col, const int *rowptr, const double *val,
                     const double *lhs, const double *rhs, double *lb, double *ub,
                     const int *vartype, const int *colptr, const int *rowidx,
                     int *marked, int *marked_next, unsigned char *chg,
                     const double *acts_given, const int *infs_given) {
    int change = 0;
    memset(marked_next, 0, (size_t)m * sizeof(int));
    for (int i = 0; i < m; i++) {
        if (!marked[i]) continue;
        double acts[3];
        int inf[2];
        if (acts_given) {
            acts[0] = acts_given[3 * i]; acts[1] = acts_given[3 * i + 1]; acts[2] = acts_given[3 * i + 2];
            inf[0] = infs_given[2 * i]; inf[1] = infs_given[2 * i + 1];
        } else {
            oracle_row_activities_seq(i, col, rowptr, val, lb, ub, acts, inf);
        }
        if (!oracle_can_cons_be_tightened(acts[0], acts[1], inf[0], inf[1], lhs[i], rhs[i], acts[2]))
            continue;
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int j = col[k];
            int lt, ut;
            double nl, nu;
            oracle_tighten_variable(val[k], lhs[i], rhs[i], acts[0], acts[1], inf[0], inf[1],
                                    vartype[j] == 3, lb[j], ub[j], &lt, &nl, &ut, &nu);
            if (lt) lb[j] = nl;
            if (ut) ub[j] = nu;
            if (lt || ut) {
                change = 1;
                if (chg) chg[j] = 1;
                mark_rows(j, colptr, rowidx, marked_next);
            }
        }
    }
    memcpy(marked, marked_next, (size_t)m * sizeof(int));
    return change;
}

typedef struct {
    int *colptr, *rowidx, *marked, *marked_next;
    unsigned char *chg;
} work_t;

static int work_alloc(work_t *w, int m, int n, int nnz) {
    w->colptr = (int *)malloc((size_t)(n + 1) * sizeof(int));
    w->rowidx = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    w->marked = (int *)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    w->marked_next = (int *)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    w->chg = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    return w->colptr && w->rowidx && w->marked && w->marked_next && w->chg;
}

static void work_free(work_t *w) {
    free(w->colptr); free(w->rowidx); free(w->marked); free(w->marked_next); free(w->chg);
}

/* :207-296 plus the optional infeasibility stop. */
int oracle_seq_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                         const double *val, const double *lhs, const double *rhs,
                         double *lb, double *ub, const int *vartype,
                         int max_rounds, int detect_infeasible, oracle_result *res) {
    res->rounds = 0; res->infeasible = 0; res->capped = 0; res->loop_seconds = 0.0;
    if (m == 0 || n == 0 || nnz == 0) return 1;
    if (detect_infeasible && any_crossed(n, lb, ub, NULL)) { res->infeasible = 1; return 1; }
    work_t w;
    if (!work_alloc(&w, m, n, nnz)) { work_free(&w); return 0; }
    oracle_csr_to_csc(m, n, nnz, col, rowptr, val, w.colptr, w.rowidx, NULL);
    for (int i = 0; i < m; i++) w.marked[i] = 1;
    int change = 1, round;
    const double t0 = now_seconds();
    for (round = 1; round <= max_rounds && change; round++) {
        memset(w.chg, 0, (size_t)n);
        change = seq_round(m, col, rowptr, val, lhs, rhs, lb, ub, vartype, w.colptr, w.rowidx,
                           w.marked, w.marked_next, w.chg, NULL, NULL);
        if (detect_infeasible && change && any_crossed(n, lb, ub, w.chg)) {
            res->infeasible = 1;
            round++;
            break;
        }
    }
    res->loop_seconds = now_seconds() - t0;
    res->rounds = round - 1;
    res->capped = change && !res->infeasible;
    work_free(&w);
    return 1;
}

/* :125-204 */
int oracle_seq_disjoint_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                                  const double *val, const double *lhs, const double *rhs,
                                  double *lb, double *ub, const int *vartype,
                                  int max_rounds, oracle_result *res) {
    res->rounds = 0; res->infeasible = 0; res->capped = 0; res->loop_seconds = 0.0;
    if (m == 0 || n == 0 || nnz == 0) return 1;
    work_t w;
    if (!work_alloc(&w, m, n, nnz)) { work_free(&w); return 0; }
    double *acts = (double *)malloc((size_t)m * 3 * sizeof(double));
    int *infs = (int *)malloc((size_t)m * 2 * sizeof(int));
    oracle_csr_to_csc(m, n, nnz, col, rowptr, val, w.colptr, w.rowidx, NULL);
    for (int i = 0; i < m; i++) w.marked[i] = 1;
    int change = 1, round;
    for (round = 1; round <= max_rounds && change; round++) {
        for (int i = 0; i < m; i++)
            oracle_row_activities_seq(i, col, rowptr, val, lb, ub, acts + 3 * i, infs + 2 * i);
        change = seq_round(m, col, rowptr, val, lhs, rhs, lb, ub, vartype, w.colptr, w.rowidx,
                           w.marked, w.marked_next, NULL, acts, infs);
    }
    res->rounds = round - 1;
    res->capped = change;
    free(acts);
    free(infs);
    work_free(&w);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* round-synchronous driver                                                 */
/* ------------------------------------------------------------------------ */

/* Every marked row reads the round-start bounds (slb, sub); accepted
 * candidates are folded into the next bounds (nlb, nub) by max / min.
 * Skipping rows none of whose variables changed in the previous round is
 * exact: their candidates and the bounds they are compared with are those of
 * the previous round, where none was accepted. */
int oracle_jacobi_propagate(int m, int n, int nnz, const int *col, const int *rowptr,
                            const double *val, const double *lhs, const double *rhs,
                            double *lb, double *ub, const int *vartype,
                            int max_rounds, int detect_infeasible, int seg_len,
                            oracle_result *res) {
    res->rounds = 0; res->infeasible = 0; res->capped = 0; res->loop_seconds = 0.0;
    if (m == 0 || n == 0 || nnz == 0) return 1;
    if (detect_infeasible && any_crossed(n, lb, ub, NULL)) { res->infeasible = 1; return 1; }
    work_t w;
    if (!work_alloc(&w, m, n, nnz)) { work_free(&w); return 0; }
    double *slb = (double *)malloc((size_t)n * sizeof(double));
    double *sub = (double *)malloc((size_t)n * sizeof(double));
    int maxlen = 1;
    for (int i = 0; i < m; i++)
        if (rowptr[i + 1] - rowptr[i] > maxlen) maxlen = rowptr[i + 1] - rowptr[i];
    double *xbuf = (double *)malloc((size_t)maxlen * 2 * sizeof(double));
    oracle_csr_to_csc(m, n, nnz, col, rowptr, val, w.colptr, w.rowidx, NULL);
    for (int i = 0; i < m; i++) w.marked[i] = 1;
    int change = 1, round;
    for (round = 1; round <= max_rounds && change; round++) {
        memcpy(slb, lb, (size_t)n * sizeof(double));
        memcpy(sub, ub, (size_t)n * sizeof(double));
        memset(w.chg, 0, (size_t)n);
        memset(w.marked_next, 0, (size_t)m * sizeof(int));
        change = 0;
        for (int i = 0; i < m; i++) {
            if (!w.marked[i]) continue;
            double acts[3];
            int inf[2];
            row_activities_canon_buf(i, col, rowptr, val, slb, sub, seg_len, acts, inf, xbuf, xbuf + maxlen);
            if (!oracle_can_cons_be_tightened(acts[0], acts[1], inf[0], inf[1], lhs[i], rhs[i], acts[2]))
                continue;
            for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
                int j = col[k];
                int lt, ut;
                double nl, nu;
                oracle_tighten_variable(val[k], lhs[i], rhs[i], acts[0], acts[1], inf[0], inf[1],
                                        vartype[j] == 3, slb[j], sub[j], &lt, &nl, &ut, &nu);
                if (lt) { if (nl > lb[j]) lb[j] = nl; w.chg[j] = 1; change = 1; }
                if (ut) { if (nu < ub[j]) ub[j] = nu; w.chg[j] = 1; change = 1; }
            }
        }
        for (int j = 0; j < n; j++)
            if (w.chg[j]) mark_rows(j, w.colptr, w.rowidx, w.marked_next);
        memcpy(w.marked, w.marked_next, (size_t)m * sizeof(int));
        if (detect_infeasible && change && any_crossed(n, lb, ub, w.chg)) {
            res->infeasible = 1;
            round++;
            break;
        }
    }
    res->rounds = round - 1;
    res->capped = change && !res->infeasible;
    free(slb);
    free(sub);
    free(xbuf);
    work_free(&w);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* progress measure: src/progressMeasureMethods/                             */
/* ------------------------------------------------------------------------ */

/* One candidate weakest bound (class NewWeakestBound, weakest_bounds.h:13-35):
 * variable, 'l' or 'u', value, CSR position of the entry that produced it. */
typedef struct {
    int var;
    char type;
    double newb;
    int pos;
} wbd_t;

typedef struct {
    int m, n;
    const int *col, *rowptr, *vt;
    const double *val, *lhs, *rhs;
    double *glb, *gub;        /* global weakest bounds (Bounds global_wbds) */
    unsigned char *expanded;  /* can_tighten_arr: one flag per nonzero */
} wb_ctx;

/* NEWcomputeInfiniteReduction (weakest_bounds.h:638-688): the candidates
 * of one row; the infinity count passed to both tightenings is the
 * minimum activity's, as in the reference. */
static int wb_row(const wb_ctx *c, int row, const double *lb, const double *ub, const double *act,
                  const int *inf, wbd_t *out) {
    int cnt = 0;
    const double minact = act[0], maxact = act[1];
    const int mininf = inf[0], maxinf = inf[1];
    const double lhs = c->lhs[row], rhs = c->rhs[row];
    for (int k = c->rowptr[row]; k < c->rowptr[row + 1]; k++) {
        const int j = c->col[k];
        const double a = c->val[k];
        const double slack = rhs - minact, surplus = lhs - maxact;
        const int cont = c->vt[j] == 3;
        if (isposinf(ub[j])) {
            const int c1 = a > 0 && !isposinf(rhs) && (mininf == 0 || (mininf == 1 && isneginf(lb[j])));
            const int c2 = a < 0 && !isneginf(lhs) && (maxinf == 0 || (maxinf == 1 && isneginf(lb[j])));
            if (c1 || c2) {
                double nb = ub[j];
                oracle_tighten_var_upper(a, slack, surplus, mininf, lb[j], ub[j], cont, &nb);
                out[cnt].var = j; out[cnt].type = 'u'; out[cnt].newb = nb; out[cnt].pos = k; cnt++;
            }
        }
        if (isneginf(lb[j])) {
            const int c1 = a > 0 && !isneginf(lhs) && (maxinf == 0 || (maxinf == 1 && isneginf(ub[j])));
            const int c2 = a < 0 && !isposinf(rhs) && (mininf == 0 || (mininf == 1 && isneginf(ub[j])));
            if (c1 || c2) {
                double nb = lb[j];
                oracle_tighten_var_lower(a, slack, surplus, mininf, lb[j], ub[j], cont, &nb);
                out[cnt].var = j; out[cnt].type = 'l'; out[cnt].newb = nb; out[cnt].pos = k; cnt++;
            }
        }
    }
    return cnt;
}

/* weakestBoundsAlgorithm (weakest_bounds.h:691-745): the local bounds are
 * a copy per call, the candidates of the call are collected over all rows
 * (each folded into the global weakest bounds, Bounds::updateWeakestBounds
 * :513-529) and expanded in order unless their entry was expanded before. */
static int wb_node(wb_ctx *c, const double *plb, const double *pub, const wbd_t *nw) {
    const int m = c->m, n = c->n;
    double *lb = malloc(sizeof(double) * (size_t)n);
    double *ub = malloc(sizeof(double) * (size_t)n);
    const int nnz = c->rowptr[m];
    wbd_t *all = malloc(sizeof(wbd_t) * (size_t)(2 * nnz + 1));
    if (!lb || !ub || !all) { free(lb); free(ub); free(all); return -1; }
    memcpy(lb, plb, sizeof(double) * (size_t)n);
    memcpy(ub, pub, sizeof(double) * (size_t)n);
    if (nw) {
        /* Bounds::updateLocalWeakestBounds (:496-510) */
        c->expanded[nw->pos] = 1;
        if (nw->type == 'l' && isneginf(lb[nw->var])) lb[nw->var] = nw->newb;
        if (nw->type == 'u' && isposinf(ub[nw->var])) ub[nw->var] = nw->newb;
    }
    int total = 0;
    for (int i = 0; i < m; i++) {
        double act[3];
        int inf[2];
        oracle_row_activities_seq(i, c->col, c->rowptr, c->val, lb, ub, act, inf);
        if ((inf[0] <= 1 && !isposinf(c->rhs[i])) || (inf[1] <= 1 && !isneginf(c->lhs[i]))) {
            const int k = wb_row(c, i, lb, ub, act, inf, all + total);
            for (int q = total; q < total + k; q++) {
                const int j = all[q].var;
                const double v = all[q].newb;
                if (all[q].type == 'l') {
                    if (isneginf(c->glb[j]) || (!isneginf(c->glb[j]) && epslt(v, c->glb[j]))) c->glb[j] = v;
                } else {
                    if (isposinf(c->gub[j]) || (!isposinf(c->gub[j]) && epsgt(v, c->gub[j]))) c->gub[j] = v;
                }
            }
            total += k;
        }
    }
    int rc = 0;
    for (int q = 0; q < total && rc == 0; q++)
        if (!c->expanded[all[q].pos]) rc = wb_node(c, lb, ub, &all[q]);
    free(lb);
    free(ub);
    free(all);
    return rc;
}

/* computeWeakestBounds (weakest_bounds.h:772-829). */
int oracle_weakest_bounds(int m, int n, int nnz, const int *col, const int *rowptr, const double *val,
                          const double *lhs, const double *rhs, double *lb, double *ub, const int *vartype) {
    wb_ctx c;
    c.m = m; c.n = n; c.col = col; c.rowptr = rowptr; c.vt = vartype;
    c.val = val; c.lhs = lhs; c.rhs = rhs;
    c.glb = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    c.gub = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    c.expanded = calloc((size_t)(nnz > 0 ? nnz : 1), 1);
    if (!c.glb || !c.gub || !c.expanded) { free(c.glb); free(c.gub); free(c.expanded); return -1; }
    memcpy(c.glb, lb, sizeof(double) * (size_t)n);
    memcpy(c.gub, ub, sizeof(double) * (size_t)n);
    const int rc = wb_node(&c, lb, ub, NULL);
    if (rc == 0) {
        memcpy(lb, c.glb, sizeof(double) * (size_t)n);
        memcpy(ub, c.gub, sizeof(double) * (size_t)n);
    }
    free(c.glb);
    free(c.gub);
    free(c.expanded);
    return rc;
}

/* calcMaxMeasureValues (progress_function.h:118-150). */
void oracle_progress_max(int n, const double *lb_in, const double *ub_in, const double *wl, const double *wu,
                         const double *ll, const double *lu, double *p_fin, int *p_inf) {
    double f = 0.0;
    int k = 0;
    for (int j = 0; j < n; j++) {
        if (epsgt(ll[j], -OR_INF) && !(fabs(ll[j] - wl[j]) <= OR_EPS)) f += 1.0;
        if (epslt(lu[j], OR_INF) && !(fabs(lu[j] - wu[j]) <= OR_EPS)) f += 1.0;
        if (epsgt(ll[j], -OR_INF) && epsle(lb_in[j], -OR_INF)) k += 1;
        if (epslt(lu[j], OR_INF) && epsge(ub_in[j], OR_INF)) k += 1;
    }
    *p_fin = f;
    *p_inf = k;
}

/* measureAndPrintProgress with calcProgressMeasureVar
 * (progress_function.h:152-195, :45-116): the unnormalised sums of one
 * round, variables in index order.  Returns 0, or -1 - j for the first
 * variable j whose limit upper bound is finite while its weakest upper bound
 * is infinite (the reference exits there). */
int oracle_progress_round(int n, const double *lb, const double *ub, const double *wl, const double *wu,
                          const double *ll, const double *lu, const double *pl, const double *pu,
                          double *p_fin, int *p_inf) {
    double f = 0.0;
    int k = 0;
    for (int j = 0; j < n; j++) {
        if (epslt(lu[j], OR_INF) && epsge(wu[j], OR_INF)) return -1 - j;
        double v = 0.0;
        if (epsgt(wl[j], -OR_INF) && epsgt(lb[j], -OR_INF) && epsge(lb[j], wl[j]) && !(fabs(ll[j] - wl[j]) <= OR_EPS))
            v += (lb[j] - wl[j]) / (ll[j] - wl[j]);
        if (epslt(wu[j], OR_INF) && epslt(ub[j], OR_INF) && epsle(ub[j], wu[j]) && !(fabs(wu[j] - lu[j]) <= OR_EPS))
            v += (wu[j] - ub[j]) / (wu[j] - lu[j]);
        f += v;
        if (epsle(pl[j], -OR_INF) && epsgt(lb[j], -OR_INF)) k += 1;
        if (epsge(pu[j], OR_INF) && epslt(ub[j], OR_INF)) k += 1;
    }
    *p_fin = f;
    *p_inf = k;
    return 0;
}

/* normalizeTo0_100 (progress_function.h:10-16). */
double oracle_progress_normalize(double val, double max_val) {
    return epsgt(max_val, 0.0) ? (100.0 * val) / max_val : 0.0;
}
