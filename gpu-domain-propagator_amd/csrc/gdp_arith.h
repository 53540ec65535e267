// gdp_arith.h - scalar propagation arithmetic shared by the HIP kernels and
// the host propagators of libGpuProp.so.
//
// The formulas are those of the reference's sequential propagator
// (src/propagation_methods.cuh:40-276, src/def.h:68-79) because the bounds the
// GPU returns must agree bit for bit with that propagator's arithmetic.  All
// operations are plain IEEE double operations evaluated one at a time (the
// library is compiled with -ffp-contract=off), division is the correctly
// rounded IEEE divide, and floor/ceil are exact, so host and device produce
// identical results from identical inputs.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define GDP_HD __host__ __device__ __forceinline__
#else
#define GDP_HD static inline
#endif

namespace gdp {

constexpr double kInf = 1e20;
constexpr double kEps = 1e-9;
constexpr int kMaxRounds = 100;
constexpr int kVarContinuous = 3;

// Segment length of the canonical activity order: a row is summed segment by
// segment (pairwise tree over the entries of a segment in natural order,
// segment sums added left to right from +0.0).  A format constant shared by
// every kernel and the oracle.
constexpr int kSegLen = 256;

GDP_HD bool eps_le(double x, double y) { return (x - y) <= kEps; }
GDP_HD bool eps_lt(double x, double y) { return (x - y) < -kEps; }
GDP_HD bool eps_ge(double x, double y) { return (x - y) >= -kEps; }
GDP_HD bool eps_gt(double x, double y) { return (x - y) > kEps; }
GDP_HD bool is_pos_inf(double x) { return eps_ge(x, kInf); }
GDP_HD bool is_neg_inf(double x) { return eps_le(x, -kInf); }

// Integer rounding of a candidate bound (propagation_methods.cuh:40-52).
GDP_HD double round_upper(bool is_cont, double ub) { return is_cont ? ub : floor(ub + kEps); }
GDP_HD double round_lower(bool is_cont, double lb) { return is_cont ? lb : ceil(lb - kEps); }

// Contributions of a_k * x_k to the row's minimum and maximum activity and
// whether they are infinite (propagation_methods.cuh:198-213).  An infinite
// contribution is counted and contributes +0.0 to the finite part.
struct Contribution {
    double min_part, max_part, delta;
    int min_inf, max_inf;
};

GDP_HD Contribution contribution(double a, double l, double u) {
    Contribution c;
    const bool pos = eps_gt(a, 0.0);
    c.min_inf = pos ? is_neg_inf(l) : is_pos_inf(u);
    c.max_inf = pos ? is_pos_inf(u) : is_neg_inf(l);
    c.min_part = c.min_inf ? 0.0 : (a > 0 ? a * l : a * u);
    c.max_part = c.max_inf ? 0.0 : (a > 0 ? a * u : a * l);
    c.delta = fabs(a) * (u - l);
    return c;
}

// Row gate of the sequential propagator (propagation_methods.cuh:136-154):
// a row is skipped when no variable can move by more than the tolerance.
GDP_HD bool row_gate(double minact, double maxact, int min_inf, int max_inf,
                     double lhs, double rhs, double delta_max) {
    if (min_inf > 1 && max_inf > 1) return false;
    if (eps_lt(delta_max, kInf)) {
        const double a = rhs - minact, b = maxact - lhs;
        return !eps_le(delta_max, (b < a) ? b : a);
    }
    return true;
}

// Largest |a|(u-l) as the sequential propagator accumulates it: a later term
// replaces the running value only when larger by more than the tolerance
// (propagation_methods.cuh:198-201).
GDP_HD double eps_max_step(double running, double d) { return eps_gt(d, running) ? d : running; }

// New upper bound from one row (propagation_methods.cuh:78-106).  `num_inf`
// is the infinity count of the activity the bound is derived from.
GDP_HD bool tighten_upper(double a, double slack, double surplus, int num_inf,
                          double l, double u, bool is_cont, double &nb) {
    double v;
    if (num_inf == 0) {
        v = a > 0 ? slack / a : surplus / a;
        v += l;
    } else if (num_inf == 1 && is_neg_inf(l)) {
        v = a > 0 ? slack / a : surplus / a;
    } else {
        return false;
    }
    v = round_upper(is_cont, v);
    if (eps_lt(v, u)) { nb = v; return true; }
    return false;
}

// New lower bound from one row (propagation_methods.cuh:108-134).
GDP_HD bool tighten_lower(double a, double slack, double surplus, int num_inf,
                          double l, double u, bool is_cont, double &nb) {
    double v;
    if (num_inf == 0) {
        v = a > 0 ? surplus / a : slack / a;
        v += u;
    } else if (num_inf == 1 && is_pos_inf(u)) {
        v = a > 0 ? surplus / a : slack / a;
    } else {
        return false;
    }
    v = round_lower(is_cont, v);
    if (eps_gt(v, l)) { nb = v; return true; }
    return false;
}

// Both candidate bounds of one variable from one row, with the
// per-variable gates of propagation_methods.cuh:223-276.  Both candidates
// are computed from the same (l, u).  Returns bit 0 = lower tightened,
// bit 1 = upper tightened.
GDP_HD int tighten_variable(double a, double lhs, double rhs, double minact, double maxact,
                            int min_inf, int max_inf, bool is_cont, double l, double u,
                            double &nl, double &nu) {
    const double slack = rhs - minact;
    const double surplus = lhs - maxact;
    int r = 0;
    if (a > 0.0) {
        if (eps_gt(a * (u - l), rhs - minact) && !is_pos_inf(rhs) && !is_neg_inf(minact))
            r |= tighten_upper(a, slack, surplus, min_inf, l, u, is_cont, nu) ? 2 : 0;
        if (eps_gt(a * (u - l), maxact - lhs) && !is_neg_inf(lhs) && !is_pos_inf(maxact))
            r |= tighten_lower(a, slack, surplus, max_inf, l, u, is_cont, nl) ? 1 : 0;
    } else {
        if (eps_gt(a * (l - u), rhs - minact) && !is_pos_inf(rhs) && !is_neg_inf(minact))
            r |= tighten_lower(a, slack, surplus, min_inf, l, u, is_cont, nl) ? 1 : 0;
        if (eps_gt(a * (l - u), maxact - lhs) && !is_neg_inf(lhs) && !is_pos_inf(maxact))
            r |= tighten_upper(a, slack, surplus, max_inf, l, u, is_cont, nu) ? 2 : 0;
    }
    return r;
}

}  // namespace gdp
