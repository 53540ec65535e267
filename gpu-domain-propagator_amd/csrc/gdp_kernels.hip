// gdp_kernels.hip - CDNA4 (gfx950) kernels of one propagation round.
//
// One round reads the bounds at the end of the previous round (snapshot S)
// and folds every accepted candidate into the accumulator A with native FP64
// atomic max/min (global_atomic_{max,min}_f64).  S and A are two buffers that
// swap roles every round.  A starts each round equal to the result of two
// rounds ago; the variables changed in the previous round are marked in a
// bitmap, and this round also folds S into A for exactly those variables
// (max/min commute, so that can happen in any order with the candidates).
//
// Work is split into 64-lane wave items:
//   group item  rows of tree width w = next_pow2(len) <= 512, processed by
//               groups of g = 2^logg lanes with K = w/g consecutive entries
//               per lane; the activity of a row is the pairwise tree over its
//               entries in natural order: K entries in registers, then an
//               xor butterfly over the g lanes.
//   long item   one row longer than kSegLen: segments of kSegLen entries are
//               reduced like a 512-wide group row and summed left to right.
//   fixup item  64 bitmap words of the previous round's changed variables.
// The summation order is thus a function of the row length only, whatever
// (g, K) a bin uses, which makes every result reproducible bit for bit.
//
// Reference counterpart: GPUAtomicDomainPropagation
// (src/kernels/atomic_kernel.cuh:9-450) and the megakernel
// (src/kernels/atomic_megakernel.cuh:14-490).
#include <hip/hip_runtime.h>

#include "gdp_arith.h"
#include "gdp_device.h"

namespace gdp {

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
    return __shfl_xor(v, mask, 64);
}

__device__ __forceinline__ void atomic_max_d(double *p, double v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_min_d(double *p, double v) {
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

// Row gate with the sequential propagator's epsilon-maximum of |a|(u-l).
// That maximum m depends on the order of the entries but always lies in
// [lo, M] with M the exact maximum (computed in parallel, order-free) and
// lo = M minus a little more than the tolerance; row_gate is monotone in its
// last argument, so when row_gate(lo) == row_gate(M) the answer is known.
// Otherwise (entries within 2*eps of the decision threshold) the row is
// rescanned in CSR order exactly as the sequential propagator does.
__device__ __noinline__ double seq_eps_max(const DevMatrix &M, const double2 *S, int start, int len) {
    double run = 0.0;
    for (int e = 0; e < len; e++) {
        const int jp = M.col[start + e];
        const double2 b = S[jp & kColMask];
        run = eps_max_step(run, fabs(M.val[start + e]) * (b.y - b.x));
    }
    return run;
}

__device__ __forceinline__ bool decide_gate(const DevMatrix &M, const double2 *S, int start, int len,
                                            double minact, double maxact, int mininf, int maxinf,
                                            double lhs, double rhs, double dmax) {
    const bool hi = row_gate(minact, maxact, mininf, maxinf, lhs, rhs, dmax);
    double lo_d = dmax;
    if (isfinite(dmax)) lo_d = dmax - 4e-9 - fabs(dmax) * 0x1p-48;
    const bool lo = row_gate(minact, maxact, mininf, maxinf, lhs, rhs, lo_d);
    if (hi == lo) return hi;
    const double m = seq_eps_max(M, S, start, len);
    return row_gate(minact, maxact, mininf, maxinf, lhs, rhs, m);
}

struct RoundCtx {
    const double2 *S;
    double2 *A;
    unsigned *bm_cur;
    unsigned *bm_prev;
    int *status;  // status words of this round
};

// Applies the candidates of one entry; returns true when a bound moved.
__device__ __forceinline__ bool apply_entry(const RoundCtx &c, int jp, double a, double l, double u,
                                            double lhs, double rhs, double minact, double maxact,
                                            int mininf, int maxinf) {
    const int j = jp & kColMask;
    const bool is_cont = jp >= 0;  // bit 31 set marks an integer column
    double nl = l, nu = u;
    const int t = tighten_variable(a, lhs, rhs, minact, maxact, mininf, maxinf, is_cont, l, u, nl, nu);
    if (t == 0) return false;
    if (t & 1) atomic_max_d(&c.A[j].x, nl);
    if (t & 2) atomic_min_d(&c.A[j].y, nu);
    atomic_or_u(&c.bm_cur[j >> 5], 1u << (j & 31));
    return true;
}

// Rows [first, first + nrows) of one tree-width bin: groups of 2^logg lanes,
// K consecutive entries per lane.
template <int K>
__device__ bool group_item(const DevMatrix &M, const RoundCtx &c, int first, int nrows, int logg) {
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
            start = M.row_ptr[r];
            len = M.row_ptr[r + 1] - start;
            const double2 sd = M.sides[r];
            lhs = sd.x;
            rhs = sd.y;
        }
        int jp[K];
        double a[K], l[K], u[K], smin[K], smax[K];
        int mininf = 0, maxinf = 0;
        double dmax = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = gl * K + k;
            smin[k] = 0.0;
            smax[k] = 0.0;
            jp[k] = 0;
            a[k] = 0.0;
            l[k] = 0.0;
            u[k] = 0.0;
            if (e < len) {
                jp[k] = M.col[start + e];
                a[k] = M.val[start + e];
                const double2 b = c.S[jp[k] & kColMask];
                l[k] = b.x;
                u[k] = b.y;
                const Contribution q = contribution(a[k], l[k], u[k]);
                smin[k] = q.min_part;
                smax[k] = q.max_part;
                mininf += q.min_inf;
                maxinf += q.max_inf;
                dmax = fmax(dmax, q.delta);
            }
        }
        // pairwise tree over the K entries of this lane
#pragma unroll
        for (int w = 1; w < K; w <<= 1) {
#pragma unroll
            for (int k = 0; k < K; k += 2 * w) {
                smin[k] = smin[k] + smin[k + w];
                smax[k] = smax[k] + smax[k + w];
            }
        }
        double tmin = smin[0], tmax = smax[0];
        for (int h = 1; h < g; h <<= 1) {
            tmin = tmin + shfl_xor_d(tmin, h);
            tmax = tmax + shfl_xor_d(tmax, h);
            mininf += __shfl_xor(mininf, h, 64);
            maxinf += __shfl_xor(maxinf, h, 64);
            dmax = fmax(dmax, shfl_xor_d(dmax, h));
        }
        const double minact = 0.0 + tmin;
        const double maxact = 0.0 + tmax;
        if (valid && len > 0 &&
            decide_gate(M, c.S, start, len, minact, maxact, mininf, maxinf, lhs, rhs, dmax)) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (gl * K + k < len)
                    moved |= apply_entry(c, jp[k], a[k], l[k], u[k], lhs, rhs, minact, maxact, mininf,
                                         maxinf);
            }
        }
    }
    return moved;
}

// One row longer than kSegLen, processed by one wave: segment sums, then a
// second pass for the candidates.
__device__ bool long_item(const DevMatrix &M, const RoundCtx &c, int r) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    const int start = M.row_ptr[r];
    const int len = M.row_ptr[r + 1] - start;
    const double2 sd = M.sides[r];
    double acc_min = 0.0, acc_max = 0.0, dmax = 0.0;
    int mininf = 0, maxinf = 0;
    for (int s0 = 0; s0 < len; s0 += kSegLen) {
        double smin[K], smax[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = s0 + lane * K + k;
            smin[k] = 0.0;
            smax[k] = 0.0;
            if (e < len) {
                const int jp = M.col[start + e];
                const double a = M.val[start + e];
                const double2 b = c.S[jp & kColMask];
                const Contribution q = contribution(a, b.x, b.y);
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
        double tmin = smin[0], tmax = smax[0];
#pragma unroll
        for (int h = 1; h < 64; h <<= 1) {
            tmin = tmin + shfl_xor_d(tmin, h);
            tmax = tmax + shfl_xor_d(tmax, h);
        }
        acc_min = acc_min + tmin;
        acc_max = acc_max + tmax;
    }
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        mininf += __shfl_xor(mininf, h, 64);
        maxinf += __shfl_xor(maxinf, h, 64);
        dmax = fmax(dmax, shfl_xor_d(dmax, h));
    }
    bool moved = false;
    if (decide_gate(M, c.S, start, len, acc_min, acc_max, mininf, maxinf, sd.x, sd.y, dmax)) {
        for (int e = lane; e < len; e += 64) {
            const int jp = M.col[start + e];
            const double a = M.val[start + e];
            const double2 b = c.S[jp & kColMask];
            moved |= apply_entry(c, jp, a, b.x, b.y, sd.x, sd.y, acc_min, acc_max, mininf, maxinf);
        }
    }
    return moved;
}

// One segment of a split row.  The wave reduces its kSegLen entries exactly
// like a 512-wide group row, stores the partial, and counts itself in.  The
// last wave of the row to arrive adds the partials in segment order, decides
// the row gate and publishes the row's state; the others wait for it (their
// entries stay in registers) and then apply their candidates.
__device__ bool seg_item(const DevMatrix &M, const DevState &D, const RoundCtx &c, int lr, int s, int round) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    const int4 info = M.seg_rows[lr];
    const int r = info.x;
    const int start = M.row_ptr[r];
    const int len = M.row_ptr[r + 1] - start;
    const int s0 = s * kSegLen;
    int jp[K];
    double a[K], l[K], u[K], smin[K], smax[K];
    int mininf = 0, maxinf = 0;
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = s0 + lane * K + k;
        smin[k] = 0.0;
        smax[k] = 0.0;
        jp[k] = 0;
        a[k] = 0.0;
        l[k] = 0.0;
        u[k] = 0.0;
        if (e < len) {
            jp[k] = M.col[start + e];
            a[k] = M.val[start + e];
            const double2 b = c.S[jp[k] & kColMask];
            l[k] = b.x;
            u[k] = b.y;
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
    double tmin = smin[0], tmax = smax[0];
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        tmin = tmin + shfl_xor_d(tmin, h);
        tmax = tmax + shfl_xor_d(tmax, h);
        mininf += __shfl_xor(mininf, h, 64);
        maxinf += __shfl_xor(maxinf, h, 64);
        dmax = fmax(dmax, shfl_xor_d(dmax, h));
    }
    int ok = 1;
    unsigned old = 0;
    SegPart *part = &D.seg_parts[info.y + s];
    if (lane == 0) {
        coherent_store_d(&part->smin, tmin);
        coherent_store_d(&part->smax, tmax);
        coherent_store_d(&part->dmax, dmax);
        coherent_store(&part->mininf, mininf);
        coherent_store(&part->maxinf, maxinf);
        mem_wait();
        old = __hip_atomic_fetch_add(&D.seg_count[lr], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    old = __shfl(old, 0, 64);
    SegState *stp = &D.seg_state[lr];
    if (old == static_cast<unsigned>(info.z - 1)) {
        // last arrival: the lanes load 64 partials at a time (independent
        // loads), which are then added in segment order
        double am = 0.0, ax = 0.0, dm = 0.0;
        int ni = 0, xi = 0;
        for (int base = 0; base < info.z; base += 64) {
            const int q = base + lane;
            double pmin = 0.0, pmax = 0.0;
            if (q < info.z) {
                const SegPart *pq = &D.seg_parts[info.y + q];
                pmin = coherent_load_d(&pq->smin);
                pmax = coherent_load_d(&pq->smax);
                dm = fmax(dm, coherent_load_d(&pq->dmax));
                ni += coherent_load(&pq->mininf);
                xi += coherent_load(&pq->maxinf);
            }
            const int cnt = min(64, info.z - base);
            for (int t = 0; t < cnt; t++) {
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
        if (lane == 0) {
            const double2 sd = M.sides[r];
            const int gate = decide_gate(M, c.S, start, len, am, ax, ni, xi, sd.x, sd.y, dm) ? 1 : 0;
            coherent_store_d(&stp->minact, am);
            coherent_store_d(&stp->maxact, ax);
            coherent_store(&stp->mininf, ni);
            coherent_store(&stp->maxinf, xi);
            coherent_store(&stp->gate, gate);
            coherent_store(&D.seg_count[lr], 0u);
            mem_wait();
            coherent_store(&D.seg_flag[lr], round);
        }
    } else if (lane == 0) {
        const long long t0 = wall_clock64();
        while (coherent_load(&D.seg_flag[lr]) != round) {
            __builtin_amdgcn_s_sleep(2);
            if (coherent_load(&D.status[kGlobalError]) || wall_clock64() - t0 > kBarrierTimeoutTicks) {
                atomic_or_i(&D.status[kGlobalError], 1);
                ok = 0;
                break;
            }
        }
    }
    ok = __shfl(ok, 0, 64);
    if (!ok) return false;
    SegState st;
    st.minact = coherent_load_d(&stp->minact);
    st.maxact = coherent_load_d(&stp->maxact);
    st.mininf = coherent_load(&stp->mininf);
    st.maxinf = coherent_load(&stp->maxinf);
    st.gate = coherent_load(&stp->gate);
    bool moved = false;
    if (st.gate) {
        const double2 sd = M.sides[r];
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (s0 + lane * K + k < len)
                moved |= apply_entry(c, jp[k], a[k], l[k], u[k], sd.x, sd.y, st.minact, st.maxact, st.mininf,
                                     st.maxinf);
        }
    }
    return moved;
}

// Folds the previous round's result into the accumulator for the variables
// it changed, checks them for crossed bounds and clears the bitmap words.
__device__ void fixup_item(const DevMatrix &M, const RoundCtx &c, int f) {
    const int w = f * 64 + (threadIdx.x & 63);
    if (w >= M.nwords) return;
    unsigned bits = coherent_load(&c.bm_prev[w]);
    if (bits == 0) return;
    coherent_store(&c.bm_prev[w], 0u);
    bool crossed = false;
    while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1;
        const int j = w * 32 + b;
        double2 s;
        s.x = coherent_load_d(&c.S[j].x);
        s.y = coherent_load_d(&c.S[j].y);
        crossed |= eps_gt(s.x, s.y);
        atomic_max_d(&c.A[j].x, s.x);
        atomic_min_d(&c.A[j].y, s.y);
    }
    if (crossed) atomic_or_i(&c.status[kStatusInfeasPrev], 1);
}

__device__ __forceinline__ RoundCtx round_ctx(const DevState &D, int round) {
    RoundCtx c;
    const bool odd = round & 1;
    c.S = odd ? D.bnd0 : D.bnd1;
    c.A = odd ? D.bnd1 : D.bnd0;
    c.bm_cur = odd ? D.bm1 : D.bm0;
    c.bm_prev = odd ? D.bm0 : D.bm1;
    c.status = D.status + kStatusWords * round;
    return c;
}

// Processes work item `it` of a round (row items, then fixup items).
__device__ bool process_item(const DevMatrix &M, const DevState &D, const RoundCtx &c, int it, int round) {
    if (it < M.n_items) {
        const int4 d = M.items[it];
        const int kind = d.x & 15;
        if (kind == kItemSeg) return seg_item(M, D, c, d.y, d.z, round);
        if (kind == kItemLong) return long_item(M, c, d.y);
        const int logg = (d.x >> 4) & 15;
        const int logk = (d.x >> 8) & 15;
        switch (logk) {
            case 0: return group_item<1>(M, c, d.y, d.z, logg);
            case 1: return group_item<2>(M, c, d.y, d.z, logg);
            default: return group_item<4>(M, c, d.y, d.z, logg);
        }
    }
    const int f = it - M.n_items;
    if (f < M.n_fix_items) fixup_item(M, c, f);
    return false;
}

__device__ __forceinline__ void flag_change(const RoundCtx &c, bool moved) {
    if (__any(moved) && (threadIdx.x & 63) == 0) atomic_or_i(&c.status[kStatusChanged], 1);
}

// One round (CPU_LOOP and GPU_LOOP).  With `guard`, the kernel first checks
// on the device whether the previous round ended the run and then does
// nothing; a captured sequence of such launches stops by itself.
__global__ __launch_bounds__(256) void round_kernel(DevMatrix M, DevState D, int round, int guard) {
    if (guard && round > 1) {
        const int *prev = D.status + kStatusWords * (round - 1);
        if (D.status[kGlobalDone] != 0 || D.status[kGlobalError] != 0 ||
            prev[kStatusInfeasPrev] != 0 || prev[kStatusChanged] == 0)
            return;
    }
    const RoundCtx c = round_ctx(D, round);
    const int waves = blockDim.x >> 6;
    const int it = blockIdx.x * waves + (threadIdx.x >> 6);
    bool moved = false;
    if (it < M.n_items + M.n_fix_items) moved = process_item(M, D, c, it, round);
    flag_change(c, moved);
}

// Fixup pass only: checks the result of the last capped round.
__global__ __launch_bounds__(256) void fixup_kernel(DevMatrix M, DevState D, int round) {
    const RoundCtx c = round_ctx(D, round);
    const int waves = blockDim.x >> 6;
    const int it = blockIdx.x * waves + (threadIdx.x >> 6);
    if (it < M.n_fix_items) fixup_item(M, c, it);
}

// Device-wide barrier for the persistent kernel: arrival counter plus a
// generation word, waiting bounded by wall-clock time so that a fault
// elsewhere ends in an error flag instead of a hang.  Everything the rounds
// exchange is written with atomics or device-coherent stores, so no cache
// write-back is needed; each workgroup invalidates its caches once after the
// barrier so that the next round's gathers of the snapshot buffer cannot hit
// lines cached while that buffer was last read.
__device__ bool grid_barrier(const DevState &D, unsigned nblocks) {
    __shared__ int ok_sh;
    mem_wait();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned *count = D.barrier;
        unsigned *gen = D.barrier + 1;
        const unsigned g0 = coherent_load(gen);
        const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            coherent_store(count, 0u);
            mem_wait();
            __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const long long t0 = wall_clock64();
            while (coherent_load(gen) == g0) {
                __builtin_amdgcn_s_sleep(1);
                if (coherent_load(&D.status[kGlobalError])) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
                    atomic_or_i(&D.status[kGlobalError], 1);
                    ok = 0;
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_sh = ok;
    }
    __syncthreads();
    return ok_sh != 0;
}

__device__ __forceinline__ int load_status(const int *p) { return coherent_load(p); }

// All rounds in one persistent launch (MEGAKERNEL).  Every workgroup walks
// the items of a round with a grid stride, then all meet at a barrier and
// read the same status words, so all take the same decision to go on.
// Outcome words (status row 0): rounds, result buffer, infeasible, capped.
__global__ __launch_bounds__(256) void mega_kernel(DevMatrix M, DevState D, int max_rounds) {
    const unsigned nblocks = gridDim.x;
    const int waves = blockDim.x >> 6;
    const int wave = blockIdx.x * waves + (threadIdx.x >> 6);
    const int stride = gridDim.x * waves;
    const int total = M.n_items + M.n_fix_items;
    int rounds = 0, result_buf = 0, infeasible = 0, capped = 0;
    for (int round = 1; round <= max_rounds; round++) {
        const RoundCtx c = round_ctx(D, round);
        bool moved = false;
        for (int it = wave; it < total; it += stride) moved |= process_item(M, D, c, it, round);
        flag_change(c, moved);
        if (!grid_barrier(D, nblocks)) { rounds = -1; break; }
        const int *st = D.status + kStatusWords * round;
        if (load_status(st + kStatusInfeasPrev)) {
            rounds = round - 1; result_buf = (round - 1) & 1; infeasible = 1;
            break;
        }
        rounds = round;
        result_buf = round & 1;
        if (!load_status(st + kStatusChanged)) break;
        if (round == max_rounds) {
            // the cap stopped a run that still changed: check its result
            const RoundCtx cf = round_ctx(D, round + 1);
            for (int it = wave; it < M.n_fix_items; it += stride) fixup_item(M, cf, it);
            if (!grid_barrier(D, nblocks)) { rounds = -1; break; }
            infeasible = load_status(D.status + kStatusWords * (round + 1) + kStatusInfeasPrev) != 0;
            capped = 1;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.status[kOutRounds] = rounds;
        D.status[kOutResultBuf] = result_buf;
        D.status[kOutInfeasible] = infeasible;
        D.status[kOutCapped] = capped;
    }
}

}  // namespace gdp

namespace gdp {

// Start of a run: both bound buffers get the input (lb, ub) pairs; bitmaps,
// status words and the barrier are cleared.  One fill kernel instead of
// memsets keeps the sequence safe to capture.
__global__ __launch_bounds__(256) void init_kernel(DevMatrix M, DevState D, const double *lbs,
                                                   const double *ubs) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int j = tid; j < M.n; j += stride) {
        const double2 b = make_double2(lbs[j], ubs[j]);
        D.bnd0[j] = b;
        D.bnd1[j] = b;
    }
    for (int w = tid; w < M.nwords; w += stride) {
        D.bm0[w] = 0u;
        D.bm1[w] = 0u;
    }
    for (int k = tid; k < kStatusRows * kStatusWords; k += stride) D.status[k] = 0;
    for (int k = tid; k < M.n_seg_rows; k += stride) {
        D.seg_count[k] = 0u;
        D.seg_flag[k] = 0;
    }
    if (tid < 64) D.barrier[tid] = 0u;
}

// End of a run: copies the result buffer out to separate lb / ub arrays.
__global__ __launch_bounds__(256) void unpack_kernel(DevMatrix M, const double2 *bnd, double *lbs, double *ubs) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int j = tid; j < M.n; j += stride) {
        const double2 b = bnd[j];
        lbs[j] = b.x;
        ubs[j] = b.y;
    }
}

}  // namespace gdp
