// gdp_kernels.hip - CDNA4 (gfx950) kernels of the propagation rounds.
//
// A round has three phases.
//   sweep    every row reads the bounds of the round's snapshot, computes its
//            activities and the candidate bounds of its columns, and folds
//            accepted candidates into the accumulator with native atomics
//            (int32 max/min for integer columns, global_atomic_{max,min}_f64
//            for the others); a bitmap records the columns that moved.
//   apply    split rows (longer than kSegLen) whose gate passed apply the
//            candidates of their segments (long rows rarely pass the gate,
//            so this phase is mostly empty).
//   publish  the columns of the bitmap are copied from the accumulator into
//            the snapshot and checked for crossed bounds (lb > ub + eps).
// Rows therefore see exactly the bounds at the end of the previous round:
// the round-synchronous semantics of the oracle.
//
// Sweep work is split into 64-lane wave items:
//   group item  rows of tree width w = next_pow2(len) <= kSegLen, processed by
//               groups of g = 2^logg lanes with K = w / g consecutive entries
//               per lane; a row's activity is the pairwise tree over its
//               entries in natural order: K entries in registers, then an
//               xor butterfly over the g lanes.
//   seg item    one kSegLen-entry segment of a longer row; segment sums are
//               added left to right by the last segment wave to arrive.
//   long item   a row that is not split (a tuning option), reduced segment
//               by segment by one wave.
// The summation order is thus a function of the row length only, whatever
// (g, K) a bin uses, which makes results reproducible bit for bit.
//
// Reference counterpart: GPUAtomicDomainPropagation
// (src/kernels/atomic_kernel.cuh:9-450) and the megakernel
// (src/kernels/atomic_megakernel.cuh:14-490).
#include <hip/hip_runtime.h>

#include "gdp_arith.h"
#include "gdp_device.h"

namespace gdp {

__device__ __forceinline__ double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }

__device__ __forceinline__ void atomic_max_d(double *p, double v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_min_d(double *p, double v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_max_i(int *p, int v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_min_i(int *p, int v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_or_u(unsigned *p, unsigned v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_or_i(int *p, int v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent plain accesses.  Relaxed agent-scope atomic loads and
// stores are single memory operations marked device-coherent (no whole-cache
// write-back or invalidate, unlike acquire/release fences at agent scope,
// which on this target maintain the entire L2).  Data exchanged between
// waves of one launch uses these; ordering comes from waiting for the
// wave's outstanding memory operations (mem_wait) before the operation that
// publishes them.
template <typename T>
__device__ __forceinline__ T coherent_load(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void coherent_store(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double coherent_load_d(const double *p) {
    return __longlong_as_double(coherent_load(reinterpret_cast<const long long *>(p)));
}
__device__ __forceinline__ void coherent_store_d(double *p, double v) {
    coherent_store(reinterpret_cast<long long *>(p), __double_as_longlong(v));
}
// waits until every memory operation this wave issued has completed
__device__ __forceinline__ void mem_wait() { __builtin_amdgcn_s_waitcnt(0); }

// ---------------------------------------------------------------------------
// bounds snapshot and accumulator
// ---------------------------------------------------------------------------

// Bounds of internal column `id` at the start of the round.
__device__ __forceinline__ void load_bounds(const DevMatrix &M, const DevState &D, int id, double &l, double &u) {
    if (id < M.n16) {
        const unsigned w = D.S16[id];
        l = static_cast<double>(static_cast<short>(w & 0xffffu));
        u = static_cast<double>(static_cast<short>(w >> 16));
    } else if (id < M.n_int) {
        const int2 w = D.S32[id - M.n16];
        l = static_cast<double>(w.x);
        u = static_cast<double>(w.y);
    } else {
        const double2 w = D.SF[id - M.n_int];
        l = w.x;
        u = w.y;
    }
}

__device__ __forceinline__ bool fits_int32(double v) { return v >= -2147483647.0 && v <= 2147483647.0; }

// Folds an accepted lower / upper candidate into the accumulator.
__device__ __forceinline__ void accumulate(const DevMatrix &M, const DevState &D, int id, int t, double nl,
                                           double nu) {
    if (id < M.n_int) {
        if (t & 1) {
            if (fits_int32(nl)) atomic_max_i(&D.AI[id].x, static_cast<int>(nl));
            else atomic_max_d(&D.AX[id].x, nl);
        }
        if (t & 2) {
            if (fits_int32(nu)) atomic_min_i(&D.AI[id].y, static_cast<int>(nu));
            else atomic_min_d(&D.AX[id].y, nu);
        }
    } else {
        double2 *a = &D.AF[id - M.n_int];
        if (t & 1) atomic_max_d(&a->x, nl);
        if (t & 2) atomic_min_d(&a->y, nu);
    }
    atomic_or_u(&D.bm[id >> 5], 1u << (id & 31));
}

// Current accumulated bounds of internal column `id` (device-coherent).
__device__ __forceinline__ void accumulated(const DevMatrix &M, const DevState &D, int id, double &l, double &u) {
    if (id < M.n_int) {
        const int *ai = reinterpret_cast<const int *>(&D.AI[id]);
        l = static_cast<double>(coherent_load(ai));
        u = static_cast<double>(coherent_load(ai + 1));
        const double xl = coherent_load_d(&D.AX[id].x);
        const double xu = coherent_load_d(&D.AX[id].y);
        l = xl > l ? xl : l;
        u = xu < u ? xu : u;
    } else {
        const double2 *a = &D.AF[id - M.n_int];
        l = coherent_load_d(&a->x);
        u = coherent_load_d(&a->y);
    }
}

__device__ __forceinline__ short clamp16(double v) {
    return static_cast<short>(v < -32767.0 ? -32767.0 : (v > 32767.0 ? 32767.0 : v));
}

__device__ __forceinline__ unsigned pack16(double l, double u) {
    return static_cast<unsigned short>(clamp16(l)) | (static_cast<unsigned>(static_cast<unsigned short>(clamp16(u))) << 16);
}

// Writes bounds into the snapshot record of `id` (device-coherent).  Values
// out of a record's range only occur once bounds have crossed, after which
// the run stops and the snapshot is not read again.
__device__ __forceinline__ void store_snapshot(const DevMatrix &M, const DevState &D, int id, double l, double u) {
    if (id < M.n16) {
        coherent_store(&D.S16[id], pack16(l, u));
    } else if (id < M.n_int) {
        int *s = reinterpret_cast<int *>(&D.S32[id - M.n16]);
        coherent_store(s, fits_int32(l) ? static_cast<int>(l) : -2147483647);
        coherent_store(s + 1, fits_int32(u) ? static_cast<int>(u) : 2147483647);
    } else {
        double2 *s = &D.SF[id - M.n_int];
        coherent_store_d(&s->x, l);
        coherent_store_d(&s->y, u);
    }
}

// ---------------------------------------------------------------------------
// row gate
// ---------------------------------------------------------------------------

// Row gate with the sequential propagator's epsilon-maximum of |a|(u-l).
// That maximum depends on the order of the entries (a later term replaces
// the running value only when larger by more than the tolerance) but always
// lies in [lo, M] with M the exact maximum (computed in parallel, order-free)
// and lo = M minus a little more than the tolerance.  row_gate is monotone in
// its last argument, so when row_gate(lo) == row_gate(M) the answer is known;
// otherwise (entries within 2*eps of the decision threshold) the
// order-dependent maximum is computed exactly by a sequential scan.
struct GateBounds {
    bool hi, lo;
    __device__ __forceinline__ bool decided() const { return hi == lo; }
};

__device__ __forceinline__ GateBounds gate_bounds(double minact, double maxact, int mininf, int maxinf, double lhs,
                                                  double rhs, double dmax) {
    GateBounds g;
    g.hi = row_gate(minact, maxact, mininf, maxinf, lhs, rhs, dmax);
    double lo_d = dmax;
    if (isfinite(dmax)) lo_d = dmax - 4e-9 - fabs(dmax) * 0x1p-48;
    g.lo = row_gate(minact, maxact, mininf, maxinf, lhs, rhs, lo_d);
    return g;
}

// Sequential epsilon-maximum of a whole row read from memory, in CSR order
// (rows that are not split, whose entries are not held in registers).
__device__ __forceinline__ double seq_eps_max(const DevMatrix &M, const DevState &D, int start, int len) {
    double run = 0.0;
    for (int e = 0; e < len; e++) {
        double l, u;
        load_bounds(M, D, M.col[start + e] & kIdMask, l, u);
        run = eps_max_step(run, fabs(M.val[start + e]) * (u - l));
    }
    return run;
}

__device__ __forceinline__ bool decide_gate(const DevMatrix &M, const DevState &D, int start, int len,
                                            double minact, double maxact, int mininf, int maxinf, double lhs,
                                            double rhs, double dmax) {
    const GateBounds g = gate_bounds(minact, maxact, mininf, maxinf, lhs, rhs, dmax);
    if (g.decided()) return g.hi;
    const double m = seq_eps_max(M, D, start, len);
    return row_gate(minact, maxact, mininf, maxinf, lhs, rhs, m);
}

// Applies the candidates of one entry; returns true when a bound moved.
__device__ __forceinline__ bool apply_entry(const DevMatrix &M, const DevState &D, int jp, double a, double l,
                                            double u, double lhs, double rhs, double minact, double maxact,
                                            int mininf, int maxinf) {
    const bool is_cont = jp >= 0;  // bit 31 set marks an integer column
    double nl = l, nu = u;
    const int t = tighten_variable(a, lhs, rhs, minact, maxact, mininf, maxinf, is_cont, l, u, nl, nu);
    if (t == 0) return false;
    accumulate(M, D, jp & kIdMask, t, nl, nu);
    return true;
}

// ---------------------------------------------------------------------------
// sweep items
// ---------------------------------------------------------------------------

// First entry and length of row r (starts are multiples of kRowAlign; the
// low bits of row_ptr[r] count the padding entries after the row).
__device__ __forceinline__ void row_extent(const DevMatrix &M, int r, int &start, int &len) {
    const int a = M.row_ptr[r], b = M.row_ptr[r + 1];
    start = a & ~(kRowAlign - 1);
    len = (b & ~(kRowAlign - 1)) - start - (a & (kRowAlign - 1));
}

// K consecutive entries of one lane: loads, contributions, pairwise tree of
// the K contributions in natural order, and the cross-lane reduction.
template <int K>
struct LaneEntries {
    int jp[K];
    double a[K], l[K], u[K];
    double tmin, tmax, dmax;
    int mininf, maxinf;

    __device__ __forceinline__ void load(const DevMatrix &M, const DevState &D, int base, int first, int len) {
        double smin[K], smax[K];
        mininf = 0;
        maxinf = 0;
        dmax = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            jp[k] = 0;
            a[k] = 0.0;
        }
        // the lane's K entries in one go: base + first is a multiple of K and
        // the row is padded to a multiple of kRowAlign (a multiple of K)
        if (first < len) {
            const int *cp = M.col + base + first;
            const double *vp = M.val + base + first;
            if constexpr (K == 4) {
                const int4 c = *reinterpret_cast<const int4 *>(cp);
                const double2 v0 = *reinterpret_cast<const double2 *>(vp);
                const double2 v1 = *reinterpret_cast<const double2 *>(vp + 2);
                jp[0] = c.x; jp[1] = c.y; jp[2] = c.z; jp[3] = c.w;
                a[0] = v0.x; a[1] = v0.y; a[2] = v1.x; a[3] = v1.y;
            } else if constexpr (K == 2) {
                const int2 c = *reinterpret_cast<const int2 *>(cp);
                const double2 v0 = *reinterpret_cast<const double2 *>(vp);
                jp[0] = c.x; jp[1] = c.y;
                a[0] = v0.x; a[1] = v0.y;
            } else {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    jp[k] = cp[k];
                    a[k] = vp[k];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = first + k;
            smin[k] = 0.0;
            smax[k] = 0.0;
            l[k] = 0.0;
            u[k] = 0.0;
            if (e >= len) {
                jp[k] = 0;
                a[k] = 0.0;
            } else {
                load_bounds(M, D, jp[k] & kIdMask, l[k], u[k]);
                const Contribution q = contribution(a[k], l[k], u[k]);
                smin[k] = q.min_part;
                smax[k] = q.max_part;
                mininf += q.min_inf;
                maxinf += q.max_inf;
                dmax = fmax(dmax, q.delta);
            }
        }
#pragma unroll
        for (int w = 1; w < K; w <<= 1) {
#pragma unroll
            for (int k = 0; k < K; k += 2 * w) {
                smin[k] = smin[k] + smin[k + w];
                smax[k] = smax[k] + smax[k + w];
            }
        }
        tmin = smin[0];
        tmax = smax[0];
    }

    // xor butterfly over groups of g lanes (natural pairwise order)
    __device__ __forceinline__ void reduce(int g) {
        for (int h = 1; h < g; h <<= 1) {
            tmin = tmin + shfl_xor_d(tmin, h);
            tmax = tmax + shfl_xor_d(tmax, h);
            mininf += __shfl_xor(mininf, h, 64);
            maxinf += __shfl_xor(maxinf, h, 64);
            dmax = fmax(dmax, shfl_xor_d(dmax, h));
        }
    }

    // Sequential epsilon-maximum over the entries of each group of g lanes in
    // natural order (lane by lane, K entries each), starting from `run`;
    // all lanes of a group return the same value.  Entries past the end of
    // the row hold zeros, which never replace the running value.
    __device__ __forceinline__ double eps_scan(int g, double run) const {
        const int gbase = (threadIdx.x & 63) & ~(g - 1);
        double d[K];
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = fabs(a[k]) * (u[k] - l[k]);
        for (int j = 0; j < g; j++) {
#pragma unroll
            for (int k = 0; k < K; k++) run = eps_max_step(run, __shfl(d[k], gbase + j, 64));
        }
        return run;
    }

    __device__ __forceinline__ bool apply(const DevMatrix &M, const DevState &D, int first, int len, double lhs,
                                          double rhs, double minact, double maxact, int ni, int xi) {
        bool moved = false;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (first + k < len) moved |= apply_entry(M, D, jp[k], a[k], l[k], u[k], lhs, rhs, minact, maxact, ni, xi);
        return moved;
    }
};

// Rows [first, first + nrows) of one tree-width bin: groups of 2^logg lanes,
// K consecutive entries per lane.
template <int K>
__device__ bool group_item(const DevMatrix &M, const DevState &D, int first, int nrows, int logg) {
    const int lane = threadIdx.x & 63;
    const int g = 1 << logg;
    const int gl = lane & (g - 1);
    const int per_pass = 64 >> logg;
    const int end = first + nrows;
    bool moved = false;
    for (int base = first; base < end; base += per_pass) {
        const int r = base + (lane >> logg);
        const bool valid = r < end;
        int start = 0, len = 0;
        double lhs = 0.0, rhs = 0.0;
        if (valid) {
            row_extent(M, r, start, len);
            const double2 sd = M.sides[r];
            lhs = sd.x;
            rhs = sd.y;
        }
        LaneEntries<K> E;
        E.load(M, D, start, gl * K, len);
        E.reduce(g);
        const double minact = 0.0 + E.tmin;
        const double maxact = 0.0 + E.tmax;
        const bool active = valid && len > 0;
        const GateBounds gb = gate_bounds(minact, maxact, E.mininf, E.maxinf, lhs, rhs, E.dmax);
        bool gate = gb.hi;
        const bool undecided = active && !gb.decided();
        if (__any(undecided)) {
            const double dm = E.eps_scan(g, 0.0);
            if (undecided) gate = row_gate(minact, maxact, E.mininf, E.maxinf, lhs, rhs, dm);
        }
        if (active && gate) moved |= E.apply(M, D, gl * K, len, lhs, rhs, minact, maxact, E.mininf, E.maxinf);
    }
    return moved;
}

// One row with too many segments to split: one wave, segment by segment,
// then a second pass for the candidates.
__device__ bool long_item(const DevMatrix &M, const DevState &D, int r) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    int start, len;
    row_extent(M, r, start, len);
    const double2 sd = M.sides[r];
    double acc_min = 0.0, acc_max = 0.0, dmax = 0.0;
    int mininf = 0, maxinf = 0;
    for (int s0 = 0; s0 < len; s0 += kSegLen) {
        LaneEntries<K> E;
        E.load(M, D, start + s0, lane * K, len - s0);
        E.reduce(64);
        acc_min = acc_min + E.tmin;
        acc_max = acc_max + E.tmax;
        mininf += E.mininf;
        maxinf += E.maxinf;
        dmax = fmax(dmax, E.dmax);
    }
    bool moved = false;
    if (decide_gate(M, D, start, len, acc_min, acc_max, mininf, maxinf, sd.x, sd.y, dmax)) {
        for (int e = lane; e < len; e += 64) {
            const int jp = M.col[start + e];
            double l, u;
            load_bounds(M, D, jp & kIdMask, l, u);
            moved |= apply_entry(M, D, jp, M.val[start + e], l, u, sd.x, sd.y, acc_min, acc_max, mininf, maxinf);
        }
    }
    return moved;
}

// Sequential epsilon-maximum of a split row, computed by its last segment
// wave.  A segment whose exact maximum Ms satisfies !(Ms - run > eps) cannot
// replace the running value (floating-point subtraction is monotone), so it
// is skipped; the others are scanned entry by entry in order.
__device__ double split_eps_scan(const DevMatrix &M, const DevState &D, const int4 info, int start, int len) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    double run = 0.0;
    for (int b = 0; b < info.z; b += 64) {
        const int q = b + lane;
        const double ms = q < info.z ? coherent_load_d(&D.seg_parts[info.y + q].dmax) : 0.0;
        const int cnt = min(64, info.z - b);
        for (int t = 0; t < cnt; t++) {
            const double mt = __shfl(ms, t, 64);
            if (!eps_gt(mt, run)) continue;
            const int s0 = (b + t) * kSegLen;
            LaneEntries<K> E;
            E.load(M, D, start + s0, lane * K, len - s0);
            run = E.eps_scan(64, run);
        }
    }
    return run;
}

// Stores the activity partial of segment `part` (lane 0 of the wave).
__device__ __forceinline__ void store_part(const DevState &D, int part, const LaneEntries<kSegLen / 64> &E) {
    SegPart *p = &D.seg_parts[part];
    coherent_store_d(&p->smin, E.tmin);
    coherent_store_d(&p->smax, E.tmax);
    coherent_store_d(&p->dmax, E.dmax);
    coherent_store(&p->mininf, E.mininf);
    coherent_store(&p->maxinf, E.maxinf);
}

// Segments [s, s + cnt) of a split row (cnt <= kSegsPerItem).  The wave
// reduces each kSegLen-entry segment like a kSegLen-wide group row, stores
// the partials and counts them in; nobody waits.  The wave that completes the row's count adds the
// partials in segment order and decides the row gate; when it passes, the
// row's activity is stored and its segments are queued for the apply phase
// that follows the sweep.
__device__ void seg_item(const DevMatrix &M, const DevState &D, int lr, int s, int cnt, int round) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    const int4 info = M.seg_rows[lr];
    const int r = info.x;
    int start, len;
    row_extent(M, r, start, len);
    for (int q = 0; q < cnt; q++) {
        const int s0 = (s + q) * kSegLen;
        LaneEntries<K> E;
        E.load(M, D, start + s0, lane * K, len - s0);
        E.reduce(64);
        if (lane == 0) store_part(D, info.y + s + q, E);
    }
    unsigned old = 0;
    if (lane == 0) {
        mem_wait();
        old = __hip_atomic_fetch_add(&D.seg_count[lr], static_cast<unsigned>(cnt), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    }
    old = __shfl(old, 0, 64);
    if (old + cnt != static_cast<unsigned>(info.z)) return;

    // the row is complete: the lanes load 64 partials at a time
    // (independent loads), which are then added in segment order
    double am = 0.0, ax = 0.0, dm = 0.0;
    int ni = 0, xi = 0;
    for (int b = 0; b < info.z; b += 64) {
        const int q = b + lane;
        double pmin = 0.0, pmax = 0.0;
        if (q < info.z) {
            const SegPart *pq = &D.seg_parts[info.y + q];
            pmin = coherent_load_d(&pq->smin);
            pmax = coherent_load_d(&pq->smax);
            dm = fmax(dm, coherent_load_d(&pq->dmax));
            ni += coherent_load(&pq->mininf);
            xi += coherent_load(&pq->maxinf);
        }
        const int n_here = min(64, info.z - b);
        for (int t = 0; t < n_here; t++) {
            am = am + __shfl(pmin, t, 64);
            ax = ax + __shfl(pmax, t, 64);
        }
    }
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        ni += __shfl_xor(ni, h, 64);
        xi += __shfl_xor(xi, h, 64);
        dm = fmax(dm, shfl_xor_d(dm, h));
    }
    const double2 sd = M.sides[r];
    const GateBounds gb = gate_bounds(am, ax, ni, xi, sd.x, sd.y, dm);
    bool gate = gb.hi;
    if (!gb.decided()) gate = row_gate(am, ax, ni, xi, sd.x, sd.y, split_eps_scan(M, D, info, start, len));
    int pos = 0;
    if (lane == 0) {
        coherent_store(&D.seg_count[lr], 0u);
        if (gate) {
            SegState *stp = &D.seg_state[lr];
            coherent_store_d(&stp->minact, am);
            coherent_store_d(&stp->maxact, ax);
            coherent_store(&stp->mininf, ni);
            coherent_store(&stp->maxinf, xi);
            pos = __hip_atomic_fetch_add(&D.status[kStatusWords * round + kStatusApply], info.z, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!gate) return;
    pos = __shfl(pos, 0, 64);
    for (int q = lane; q < info.z; q += 64) coherent_store(&D.apply_tasks[pos + q], info.y + q);
}

// Applies the candidates of segment `part` of a split row whose gate passed
// (one wave; the row's activity is read back from its SegState).
__device__ bool apply_split_task(const DevMatrix &M, const DevState &D, int part) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    const int lr = M.part_row[part];
    const int4 info = M.seg_rows[lr];
    const int r = info.x;
    int start, len;
    row_extent(M, r, start, len);
    const int s0 = (part - info.y) * kSegLen;
    const SegState *stp = &D.seg_state[lr];
    const double minact = coherent_load_d(&stp->minact);
    const double maxact = coherent_load_d(&stp->maxact);
    const int ni = coherent_load(&stp->mininf);
    const int xi = coherent_load(&stp->maxinf);
    const double2 sd = M.sides[r];
    bool moved = false;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = s0 + lane * K + k;
        if (e < len) {
            const int jp = M.col[start + e];
            double l, u;
            load_bounds(M, D, jp & kIdMask, l, u);
            moved |= apply_entry(M, D, jp, M.val[start + e], l, u, sd.x, sd.y, minact, maxact, ni, xi);
        }
    }
    return moved;
}

// The apply phase: the waves of the grid take the queued segments in turn.
__device__ bool apply_phase(const DevMatrix &M, const DevState &D, int round, int wave, int nwaves) {
    const int count = coherent_load(&D.status[kStatusWords * round + kStatusApply]);
    bool moved = false;
    for (int t = wave; t < count; t += nwaves) moved |= apply_split_task(M, D, coherent_load(&D.apply_tasks[t]));
    return moved;
}

__device__ bool sweep_item(const DevMatrix &M, const DevState &D, int it, int round) {
    const int4 d = M.items[it];
    const int kind = d.x & 15;
    if (kind == kItemSeg) {
        seg_item(M, D, d.y, d.z, d.w, round);
        return false;
    }
    if (kind == kItemLong) return long_item(M, D, d.y);
    const int logg = (d.x >> 4) & 15;
    const int logk = (d.x >> 8) & 15;
    switch (logk) {
        case 0: return group_item<1>(M, D, d.y, d.z, logg);
        case 1: return group_item<2>(M, D, d.y, d.z, logg);
        default: return group_item<4>(M, D, d.y, d.z, logg);
    }
}

// ---------------------------------------------------------------------------
// publish items
// ---------------------------------------------------------------------------

// Copies the accumulated bounds of column `id` into the snapshot when it
// moved this round (bitmap), checks them for crossed bounds and clears the
// bitmap word (the lane of the word's first column, after the wave has read
// it; the ids of a wave are 64 consecutive columns starting at a multiple of
// 64).  Returns true on a crossing.
__device__ __forceinline__ bool publish_column(const DevMatrix &M, const DevState &D, int id) {
    if (id >= M.n) return false;
    const unsigned bits = coherent_load(&D.bm[id >> 5]);
    if (bits == 0u) return false;
    if ((id & 31) == 0) coherent_store(&D.bm[id >> 5], 0u);
    if (!((bits >> (id & 31)) & 1u)) return false;
    double l, u;
    accumulated(M, D, id, l, u);
    store_snapshot(M, D, id, l, u);
    return eps_gt(l, u);
}

__device__ __forceinline__ void flag_crossing(int *round_status, bool crossed) {
    if (__any(crossed) && (threadIdx.x & 63) == 0) atomic_or_i(&round_status[kStatusInfeasible], 1);
}

// Publish phase over the columns [first, n) with a stride that is a multiple
// of 64 (every wave sees whole bitmap words).
__device__ __forceinline__ bool publish_phase(const DevMatrix &M, const DevState &D, int first, int stride) {
    bool crossed = false;
    for (int id = first; id < M.n; id += stride) crossed |= publish_column(M, D, id);
    return crossed;
}

__device__ __forceinline__ void flag_change(int *round_status, bool moved) {
    if (__any(moved) && (threadIdx.x & 63) == 0) atomic_or_i(&round_status[kStatusChanged], 1);
}

// With `guard` (captured graphs), a launch first checks on the device whether
// the run already ended before this round and then does nothing.
__device__ __forceinline__ bool round_skipped(const DevState &D, int round) {
    if (D.status[kGlobalError] != 0 || D.status[kGlobalDone] != 0 || D.status[kGlobalEnvelope] != 0) return true;
    if (round <= 1) return false;
    const int *prev = D.status + kStatusWords * (round - 1);
    return prev[kStatusInfeasible] != 0 || prev[kStatusChanged] == 0;
}

__global__ __launch_bounds__(256) void sweep_kernel(DevMatrix M, DevState D, int round, int guard) {
    if (guard && round_skipped(D, round)) return;
    const int it = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    bool moved = false;
    if (it < M.n_items) moved = sweep_item(M, D, it, round);
    flag_change(D.status + kStatusWords * round, moved);
}

__global__ __launch_bounds__(256) void apply_split_kernel(DevMatrix M, DevState D, int round, int guard) {
    if (guard && round_skipped(D, round)) return;
    const int waves = blockDim.x >> 6;
    const bool moved = apply_phase(M, D, round, blockIdx.x * waves + (threadIdx.x >> 6), gridDim.x * waves);
    flag_change(D.status + kStatusWords * round, moved);
}

__global__ __launch_bounds__(256) void publish_kernel(DevMatrix M, DevState D, int round, int guard) {
    if (guard && round_skipped(D, round)) return;
    const bool crossed = publish_phase(M, D, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    flag_crossing(D.status + kStatusWords * round, crossed);
}

// ---------------------------------------------------------------------------
// persistent kernel
// ---------------------------------------------------------------------------

// Device-wide barrier.  Workgroups arrive at one of 8 group counters
// (blockIdx mod 8, each on its own cache line); the last arrival of a group
// arrives at the top counter, and the last of those advances the generation
// word that all others wait for.  Waiting is bounded by wall-clock time so
// that a fault elsewhere ends in an error flag instead of a hang.
// Everything the phases exchange is written with atomics or device-coherent
// stores, so no cache write-back is needed.  With `invalidate`, each
// workgroup invalidates its caches after the barrier so that its next
// gathers of the snapshot cannot hit lines cached before the publish.
__device__ bool grid_barrier(const DevState &D, unsigned nblocks, bool invalidate) {
    __shared__ int ok_sh;
    mem_wait();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned *top = D.barrier + kBarrierTop;
        unsigned *gen = D.barrier + kBarrierGen;
        const unsigned x = blockIdx.x & 7;
        unsigned *grp = D.barrier + kBarrierGroup + 32 * x;
        const unsigned per = nblocks / 8 + ((nblocks & 7) > x ? 1u : 0u);
        const unsigned groups = nblocks < 8 ? nblocks : 8;
        const unsigned g0 = coherent_load(gen);
        bool released = false;
        if (__hip_atomic_fetch_add(grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == per - 1) {
            coherent_store(grp, 0u);
            mem_wait();
            if (__hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == groups - 1) {
                coherent_store(top, 0u);
                mem_wait();
                __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                released = true;
            }
        }
        if (!released) {
            const long long t0 = wall_clock64();
            while (coherent_load(gen) == g0) {
                __builtin_amdgcn_s_sleep(1);
                if (coherent_load(&D.status[kGlobalError])) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > kWaitTimeoutTicks) {
                    atomic_or_i(&D.status[kGlobalError], 1);
                    ok = 0;
                    break;
                }
            }
        }
        if (invalidate) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_sh = ok;
    }
    __syncthreads();
    return ok_sh != 0;
}

// All rounds in one persistent launch (MEGAKERNEL).  Every workgroup walks
// the items of a phase with a grid stride (consecutive items on consecutive
// waves), all meet at a barrier, and all read the same status words, so all
// take the same decision to go on.  Phases of a round: sweep; apply of the
// split rows when the sweep queued any; publish.  Outcome words in status
// row 0.
__global__ __launch_bounds__(256) void mega_kernel(DevMatrix M, DevState D, int max_rounds) {
    const unsigned nblocks = gridDim.x;
    const int waves = blockDim.x >> 6;
    const int wave = blockIdx.x * waves + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * waves;
    int rounds = 0, infeasible = 0, capped = 0;
    for (int round = 1; round <= max_rounds; round++) {
        int *st = D.status + kStatusWords * round;
        bool moved = false;
        for (int it = wave; it < M.n_items; it += nwaves) moved |= sweep_item(M, D, it, round);
        if (!grid_barrier(D, nblocks, false)) { rounds = -1; break; }
        if (coherent_load(&st[kStatusApply]) > 0) {
            moved |= apply_phase(M, D, round, wave, nwaves);
            if (!grid_barrier(D, nblocks, false)) { rounds = -1; break; }
        }
        flag_change(st, moved);
        const bool crossed =
            publish_phase(M, D, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
        flag_crossing(st, crossed);
        if (!grid_barrier(D, nblocks, true)) { rounds = -1; break; }
        rounds = round;
        if (coherent_load(&st[kStatusInfeasible])) { infeasible = 1; break; }
        if (!coherent_load(&st[kStatusChanged])) break;
        if (round == max_rounds) capped = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.status[kOutRounds] = rounds;
        D.status[kOutInfeasible] = infeasible;
        D.status[kOutCapped] = capped;
    }
}

// ---------------------------------------------------------------------------
// start and end of a run
// ---------------------------------------------------------------------------

// Loads the input bounds (original column order) into accumulator and
// snapshot; clears bitmap, status words, barrier and split-row counters.
// Sets the envelope flag when input bounds do not fit a column's storage
// class (that word is cleared by the preceding clear_envelope_kernel).
__global__ __launch_bounds__(256) void init_kernel(DevMatrix M, DevState D, const double *lbs, const double *ubs) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int k = tid; k < kStatusRows * kStatusWords; k += stride)
        if (k != kGlobalEnvelope && k != kGlobalCrossed) D.status[k] = 0;
    for (int id = tid; id < M.n; id += stride) {
        const int j = M.perm[id];
        const double l = lbs[j], u = ubs[j];
        if (eps_gt(l, u)) atomic_or_i(&D.status[kGlobalCrossed], 1);
        if (id < M.n_int) {
            const double lim = id < M.n16 ? 32767.0 : 2147483647.0;
            const bool fits = l >= -lim && l <= lim && u >= -lim && u <= lim && floor(l) == l && floor(u) == u;
            if (!fits) atomic_or_i(&D.status[kGlobalEnvelope], 1);
            const int il = fits ? static_cast<int>(l) : 0, iu = fits ? static_cast<int>(u) : 0;
            D.AI[id] = make_int2(il, iu);
            D.AX[id] = make_double2(-INFINITY, INFINITY);
            if (id < M.n16) D.S16[id] = pack16(il, iu);
            else D.S32[id - M.n16] = make_int2(il, iu);
        } else {
            D.AF[id - M.n_int] = make_double2(l, u);
            D.SF[id - M.n_int] = make_double2(l, u);
        }
    }
    for (int w = tid; w < M.nwords; w += stride) D.bm[w] = 0u;
    for (int k = tid; k < M.n_seg_rows; k += stride) D.seg_count[k] = 0u;
    for (int k = tid; k < kBarrierWords; k += stride) D.barrier[k] = 0u;
}

// Clears the two words init_kernel sets from many threads (separate launch).
__global__ void clear_envelope_kernel(DevState D) {
    D.status[kGlobalEnvelope] = 0;
    D.status[kGlobalCrossed] = 0;
}

// End of a run: accumulated bounds back to original column order.
__global__ __launch_bounds__(256) void unpack_kernel(DevMatrix M, DevState D, double *lbs, double *ubs) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int id = tid; id < M.n; id += stride) {
        double l, u;
        if (id < M.n_int) {
            const int2 a = D.AI[id];
            const double2 x = D.AX[id];
            l = static_cast<double>(a.x);
            u = static_cast<double>(a.y);
            l = x.x > l ? x.x : l;
            u = x.y < u ? x.y : u;
        } else {
            const double2 a = D.AF[id - M.n_int];
            l = a.x;
            u = a.y;
        }
        const int j = M.perm[id];
        lbs[j] = l;
        ubs[j] = u;
    }
}

}  // namespace gdp
