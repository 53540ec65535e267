// gdp_measure.cpp - the progress measure of a propagation run
// (reference src/progressMeasureMethods/): the weakest bounds of the input,
// and per round the normalised progress between the weakest and the limit
// (fixpoint) bounds.
//
// The weakest-bounds search is a depth-first search over bound changes that
// turn an infinite bound finite.  It runs on the host, as in the reference.
// At every node the reference copies the bounds and recomputes the
// activities of all rows; a node differs from its parent in one bound, so
// here only the rows of that column are recomputed (each from scratch, in
// CSR order, so every value equals the full recomputation's) and restored
// when the search returns, and the search keeps its own stack.  Candidates
// are taken and folded into the weakest bounds in the reference's order, and
// each matrix entry is expanded once, as there.
#include <set>
#include <vector>

#include "gdp_arith.h"
#include "gdp_host.h"

namespace gdp {

namespace {

// One candidate weakest bound (NewWeakestBound, weakest_bounds.h:13-35).
struct Candidate {
    int var;
    bool upper;
    double value;
    int pos;  // CSR position of the entry that produced it
};

// NEWcomputeInfiniteReduction (weakest_bounds.h:638-688): candidates of one
// row for the bounds of its variables that are infinite.  Both tightenings
// are given the minimum activity's infinity count, as in the reference; a
// candidate that does not tighten carries the current (infinite) bound.
void row_candidates(int i, const int *col, const int *row_ptr, const double *val, const double *lhs,
                    const double *rhs, const int *vartype, const double *lb, const double *ub, const RowActs &a,
                    std::vector<Candidate> &out) {
    const double slack = rhs[i] - a.minact, surplus = lhs[i] - a.maxact;
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
        const int j = col[k];
        const double c = val[k];
        const bool cont = vartype[j] == kVarContinuous;
        if (is_pos_inf(ub[j])) {
            const bool by_min = c > 0 && !is_pos_inf(rhs[i]) &&
                                (a.min_inf == 0 || (a.min_inf == 1 && is_neg_inf(lb[j])));
            const bool by_max = c < 0 && !is_neg_inf(lhs[i]) &&
                                (a.max_inf == 0 || (a.max_inf == 1 && is_neg_inf(lb[j])));
            if (by_min || by_max) {
                double v = ub[j];
                tighten_upper(c, slack, surplus, a.min_inf, lb[j], ub[j], cont, v);
                out.push_back({j, true, v, k});
            }
        }
        if (is_neg_inf(lb[j])) {
            const bool by_max = c > 0 && !is_neg_inf(lhs[i]) &&
                                (a.max_inf == 0 || (a.max_inf == 1 && is_neg_inf(ub[j])));
            const bool by_min = c < 0 && !is_pos_inf(rhs[i]) &&
                                (a.min_inf == 0 || (a.min_inf == 1 && is_neg_inf(ub[j])));
            if (by_max || by_min) {
                double v = lb[j];
                tighten_lower(c, slack, surplus, a.min_inf, lb[j], ub[j], cont, v);
                out.push_back({j, false, v, k});
            }
        }
    }
}

struct Node {
    std::vector<Candidate> cands;
    size_t next = 0;
    int var = -1;  // local bound set on entry, restored on exit
    bool upper = false;
    double old = 0.0;
    size_t undo = 0;  // row states saved on entry (size of the undo log before them)
};

struct RowState {
    int row;
    RowActs acts;
    std::vector<Candidate> cands;
};

}  // namespace

void weakest_bounds(int m, int n, const int *col, const int *row_ptr, const double *val, const double *lhs,
                    const double *rhs, const int *vartype, double *lb, double *ub) {
    const int nnz = row_ptr[m];
    std::vector<double> glb(lb, lb + n), gub(ub, ub + n);  // weakest bounds found so far
    std::vector<double> llb(lb, lb + n), lub(ub, ub + n);  // bounds of the current node
    std::vector<unsigned char> expanded(nnz > 0 ? nnz : 1, 0);
    // rows of every column, to find the rows whose activities a change touches
    std::vector<int> colptr(n + 1, 0), rows(nnz > 0 ? nnz : 1);
    for (int k = 0; k < nnz; k++) colptr[col[k] + 1]++;
    for (int j = 0; j < n; j++) colptr[j + 1] += colptr[j];
    {
        std::vector<int> fill(colptr.begin(), colptr.end() - 1);
        for (int i = 0; i < m; i++)
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) rows[fill[col[k]]++] = i;
    }
    // Activities and candidates of every row under the current node's
    // bounds.  A node differs from its parent in one bound, so only the rows
    // of that column are recomputed (from scratch, in CSR order: the values
    // are those a full recomputation gives) and restored on the way back.
    std::vector<RowActs> acts(m);
    std::vector<std::vector<Candidate>> cands(m);
    std::set<int> with_cands;  // rows holding candidates, in row order
    auto compute_row = [&](int i) {
        acts[i] = row_activities(i, col, row_ptr, val, llb.data(), lub.data());
        cands[i].clear();
        const RowActs &a = acts[i];
        if ((a.min_inf <= 1 && !is_pos_inf(rhs[i])) || (a.max_inf <= 1 && !is_neg_inf(lhs[i])))
            row_candidates(i, col, row_ptr, val, lhs, rhs, vartype, llb.data(), lub.data(), a, cands[i]);
        if (cands[i].empty())
            with_cands.erase(i);
        else
            with_cands.insert(i);
    };
    std::vector<RowState> undo;
    std::vector<Node> stack;
    // weakestBoundsAlgorithm (weakest_bounds.h:691-745): apply the node's
    // change (Bounds::updateLocalWeakestBounds, :496-510), take the
    // candidates of every row that has at most one infinite contribution on
    // a finite side, in row order, folding each into the weakest bounds
    // (Bounds::updateWeakestBounds, :513-529).
    auto open = [&](const Candidate *c) {
        Node node;
        node.undo = undo.size();
        if (c) {
            expanded[c->pos] = 1;
            double &b = c->upper ? lub[c->var] : llb[c->var];
            if (c->upper ? is_pos_inf(b) : is_neg_inf(b)) {
                node.var = c->var;
                node.upper = c->upper;
                node.old = b;
                for (int p = colptr[c->var]; p < colptr[c->var + 1]; p++) {
                    const int i = rows[p];
                    undo.push_back({i, acts[i], cands[i]});
                }
                b = c->value;
                for (int p = colptr[c->var]; p < colptr[c->var + 1]; p++) compute_row(rows[p]);
            }
        }
        for (const int i : with_cands)
            for (const Candidate &d : cands[i]) {
                if (d.upper) {
                    double &g = gub[d.var];
                    if (is_pos_inf(g) || eps_gt(d.value, g)) g = d.value;
                } else {
                    double &g = glb[d.var];
                    if (is_neg_inf(g) || eps_lt(d.value, g)) g = d.value;
                }
                node.cands.push_back(d);
            }
        stack.push_back(std::move(node));
    };
    for (int i = 0; i < m; i++) compute_row(i);
    open(nullptr);
    while (!stack.empty()) {
        Node &top = stack.back();
        if (top.next < top.cands.size()) {
            const Candidate c = top.cands[top.next++];
            if (!expanded[c.pos]) open(&c);
            continue;
        }
        if (top.var >= 0) {
            (top.upper ? lub : llb)[top.var] = top.old;
            while (undo.size() > top.undo) {
                RowState &r = undo.back();
                acts[r.row] = r.acts;
                cands[r.row] = std::move(r.cands);
                if (cands[r.row].empty())
                    with_cands.erase(r.row);
                else
                    with_cands.insert(r.row);
                undo.pop_back();
            }
        }
        stack.pop_back();
    }
    for (int j = 0; j < n; j++) {
        lb[j] = glb[j];
        ub[j] = gub[j];
    }
}

Measure measure_max(int n, const double *lb_in, const double *ub_in, const double *wl, const double *wu,
                    const double *ll, const double *lu) {
    Measure r;
    for (int j = 0; j < n; j++) {
        if (eps_gt(ll[j], -kInf) && !(fabs(ll[j] - wl[j]) <= kEps)) r.fin += 1.0;
        if (eps_lt(lu[j], kInf) && !(fabs(lu[j] - wu[j]) <= kEps)) r.fin += 1.0;
        if (eps_gt(ll[j], -kInf) && is_neg_inf(lb_in[j])) r.inf += 1;
        if (eps_lt(lu[j], kInf) && is_pos_inf(ub_in[j])) r.inf += 1;
    }
    return r;
}

int measure_defect(int n, const double *wu, const double *lu) {
    for (int j = 0; j < n; j++)
        if (eps_lt(lu[j], kInf) && is_pos_inf(wu[j])) return j;
    return -1;
}

Measure measure_round(int n, const double *lb, const double *ub, const double *wl, const double *wu,
                      const double *ll, const double *lu, const double *pl, const double *pu) {
    Measure r;
    for (int j = 0; j < n; j++) {
        double v = 0.0;
        if (eps_gt(wl[j], -kInf) && eps_gt(lb[j], -kInf) && eps_ge(lb[j], wl[j]) && !(fabs(ll[j] - wl[j]) <= kEps))
            v += (lb[j] - wl[j]) / (ll[j] - wl[j]);
        if (eps_lt(wu[j], kInf) && eps_lt(ub[j], kInf) && eps_le(ub[j], wu[j]) && !(fabs(wu[j] - lu[j]) <= kEps))
            v += (wu[j] - ub[j]) / (wu[j] - lu[j]);
        r.fin += v;
        if (is_neg_inf(pl[j]) && eps_gt(lb[j], -kInf)) r.inf += 1;
        if (is_pos_inf(pu[j]) && eps_lt(ub[j], kInf)) r.inf += 1;
    }
    return r;
}

double measure_normalize(double v, double max_v) { return eps_gt(max_v, 0.0) ? (100.0 * v) / max_v : 0.0; }

}  // namespace gdp
