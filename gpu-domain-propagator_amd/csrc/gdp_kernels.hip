// gdp_kernels.hip - CDNA4 (gfx950) kernels of the propagation steps.
//
// A step advances every active unit (bound vector) by one round, in three
// phases:
//   sweep    every marked row of every active unit reads the unit's bound
//            snapshot, computes its activities and the candidate bounds of
//            its columns, and folds accepted candidates into the unit's
//            accumulator with native atomics (int32 max/min for compact
//            integer columns, global_atomic_{max,min}_f64 for the others); the
//            first candidate of a 32-column word queues the word for publish.
//   apply    split rows (longer than kSegLen) whose gate passed apply the
//            candidates of their segments (long rows rarely pass the gate,
//            so this phase is mostly empty).
//   publish  the queued words' moved columns are copied from the accumulator
//            into the snapshot, checked for crossed bounds (lb > ub + eps),
//            and the rows of the moved columns are marked for the next step.
// Rows therefore see exactly the bounds at the end of the unit's previous
// round: the round-synchronous semantics of the oracle.  A row none of whose
// columns moved in the previous round is skipped; it would reproduce
// candidates that were already applied, so skipping it changes no result.
//
// Sweep work is split into 64-lane wave items:
//   group item  rows of tree width w = next_pow2(len) <= kSegLen, processed by
//               groups of g = 2^logg lanes with K = w / g consecutive entries
//               per lane; a row's activity is the pairwise tree over its
//               entries in natural order: K entries in registers, then an
//               xor butterfly over the g lanes.
//   seg item    up to kSegsPerItem kSegLen-entry segments of a longer row;
//               segment sums are added left to right by the wave that
//               completes the row.
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
__device__ __forceinline__ unsigned atomic_or_u(unsigned *p, unsigned v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int atomic_add_i(int *p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
// Two adjacent ints (8-byte aligned, first in the low word) in one
// device-coherent access.
__device__ __forceinline__ void coherent_store_2i(int *p, int a, int b) {
    coherent_store(reinterpret_cast<long long *>(p),
                   static_cast<long long>(static_cast<unsigned long long>(static_cast<unsigned>(a)) |
                                          (static_cast<unsigned long long>(static_cast<unsigned>(b)) << 32)));
}
__device__ __forceinline__ void coherent_load_2i(const int *p, int &a, int &b) {
    const unsigned long long w =
        static_cast<unsigned long long>(coherent_load(reinterpret_cast<const long long *>(p)));
    a = static_cast<int>(w & 0xffffffffull);
    b = static_cast<int>(w >> 32);
}
// waits until every memory operation this wave issued has completed
__device__ __forceinline__ void mem_wait() { __builtin_amdgcn_s_waitcnt(0); }

// ---------------------------------------------------------------------------
// bounds snapshot and accumulator
// ---------------------------------------------------------------------------

// Snapshot records are read with one aligned 16-byte load of the chunk that
// holds them, decoded without branches, so that the gathers of a lane issue
// back to back whatever the storage classes of their columns.
// Byte offset of the record of internal column `id` in the unit's block.
__device__ __forceinline__ unsigned record_offset(const DevMatrix &M, int id) {
    const unsigned uid = static_cast<unsigned>(id);
    const unsigned a16 = uid * 4u;
    const unsigned a32 = static_cast<unsigned>(M.snap_off32) + (uid - static_cast<unsigned>(M.n16)) * 8u;
    const unsigned af = static_cast<unsigned>(M.snap_offF) + (uid - static_cast<unsigned>(M.n_int)) * 16u;
    const unsigned a = id < M.n16 ? a16 : a32;
    return id < M.n_int ? a : af;
}

// Dynamic LDS of the persistent kernel: the unit filter's arrays in a launch
// of several units, else the 2-bit codes of the first D.n_lds binary columns
// (other kernels run with D.n_lds = 0 and never touch it).
extern __shared__ __align__(16) unsigned dyn_lds[];

// Bounds of binary column `id` < D.n_lds from its 2-bit code in LDS.
__device__ __forceinline__ void lds_binary_bounds(int id, double &l, double &u) {
    const unsigned code = dyn_lds[id >> 4] >> ((id & 15) * 2);
    l = (code & 1u) ? 1.0 : 0.0;
    u = (code & 2u) ? 0.0 : 1.0;
}

__device__ __forceinline__ uint4 load_chunk(const DevState &D, unsigned off) {
    return *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(D.S16) + (off & ~15u));
}

// Bounds of `id` from the chunk at its record offset `off`.
__device__ __forceinline__ void decode_bounds(const DevMatrix &M, int id, unsigned off, const uint4 c, double &l,
                                              double &u) {
    const unsigned odd = (off >> 2) & 1u, upper = (off >> 3) & 1u;
    // 4-byte record: dword (off >> 2) & 3; 8-byte record: that dword and the next
    const unsigned x01 = odd ? c.y : c.x;
    const unsigned x23 = odd ? c.w : c.z;
    const unsigned lo = upper ? x23 : x01;
    const unsigned hi = upper ? c.w : c.y;
    const double l16 = static_cast<double>(static_cast<short>(lo & 0xffffu));
    const double u16 = static_cast<double>(static_cast<short>(lo >> 16));
    const double l32 = static_cast<double>(static_cast<int>(lo));
    const double u32 = static_cast<double>(static_cast<int>(hi));
    const double lf = __hiloint2double(static_cast<int>(c.y), static_cast<int>(c.x));
    const double uf = __hiloint2double(static_cast<int>(c.w), static_cast<int>(c.z));
    const bool c16 = id < M.n16, c32 = id < M.n_int;
    const double li = c16 ? l16 : l32, ui = c16 ? u16 : u32;
    l = c32 ? li : lf;
    u = c32 ? ui : uf;
}

// Bounds of internal column `id` at the start of the round.
__device__ __forceinline__ void load_bounds(const DevMatrix &M, const DevState &D, int id, double &l, double &u) {
    if (id < D.n_lds) {
        lds_binary_bounds(id, l, u);
    } else {
        const unsigned off = record_offset(M, id);
        decode_bounds(M, id, off, load_chunk(D, off), l, u);
    }
}

__device__ __forceinline__ bool fits_int32(double v) { return v >= -2147483647.0 && v <= 2147483647.0; }

// Where accepted candidates go: the unit's accumulator and bitmap.
struct Sink {
    DevState V;
};

// Sentinels of the integer accumulator: a candidate beyond the int32 range
// goes to AX, and AI's side is pushed to its sentinel so that readers know
// to consult AX (AI.x == kAiHigh may also be a genuine bound; reading AX
// then is merely redundant).  Integer candidates are integral, so those
// beyond the range are >= 2^31 or <= -2^31 and the max / min with the
// sentinel is the candidate itself.
constexpr int kAiHigh = 2147483647;
constexpr int kAiLow = -2147483647 - 1;

// Folds an accepted lower / upper candidate into the accumulator and marks
// the column in the unit's bitmap of moved columns.
__device__ __forceinline__ void accumulate(const DevMatrix &M, const Sink &K, int id, int t, double nl, double nu) {
    const DevState &D = K.V;
    if (id < M.n_int) {
        if (t & 1) {
            if (fits_int32(nl)) {
                atomic_max_i(&D.AI[id].x, static_cast<int>(nl));
            } else {
                atomic_max_d(&D.AX[id].x, nl);
                atomic_max_i(&D.AI[id].x, kAiHigh);
            }
        }
        if (t & 2) {
            if (fits_int32(nu)) {
                atomic_min_i(&D.AI[id].y, static_cast<int>(nu));
            } else {
                atomic_min_d(&D.AX[id].y, nu);
                atomic_min_i(&D.AI[id].y, kAiLow);
            }
        }
    } else {
        double2 *a = &D.AF[id - M.n_int];
        if (t & 1) atomic_max_d(&a->x, nl);
        if (t & 2) atomic_min_d(&a->y, nu);
    }
    __hip_atomic_fetch_or(&D.bm[id >> 5], 1u << (id & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The accumulator words of internal column `id` as read (device-coherent):
// both int32 sides in one access, or the FP64 pair.  acc_resolve() turns
// them into bounds; issuing the reads of several columns first lets their
// latencies overlap.
struct AccRaw {
    long long w;
    double fl, fu;
};
__device__ __forceinline__ AccRaw acc_load(const DevMatrix &M, const DevState &D, int id) {
    AccRaw a{0, 0.0, 0.0};
    if (id < M.n_int) {
        a.w = coherent_load(reinterpret_cast<const long long *>(&D.AI[id]));
    } else {
        const double2 *p = &D.AF[id - M.n_int];
        a.fl = coherent_load_d(&p->x);
        a.fu = coherent_load_d(&p->y);
    }
    return a;
}
__device__ __forceinline__ void acc_resolve(const DevMatrix &M, const DevState &D, int id, const AccRaw &a,
                                            double &l, double &u) {
    if (id < M.n_int) {
        const int lo = static_cast<int>(static_cast<unsigned long long>(a.w) & 0xffffffffull);
        const int hi = static_cast<int>(static_cast<unsigned long long>(a.w) >> 32);
        l = static_cast<double>(lo);
        u = static_cast<double>(hi);
        if (lo == kAiHigh || hi == kAiLow) {  // AX only behind a sentinel
            const double xl = coherent_load_d(&D.AX[id].x);
            const double xu = coherent_load_d(&D.AX[id].y);
            l = xl > l ? xl : l;
            u = xu < u ? xu : u;
        }
    } else {
        l = a.fl;
        u = a.fu;
    }
}

// Current accumulated bounds of internal column `id` (device-coherent).
__device__ __forceinline__ void accumulated(const DevMatrix &M, const DevState &D, int id, double &l, double &u) {
    if (id < M.n_int) {
        // both int32 sides in one device-coherent access; AX only behind a sentinel
        const long long w = coherent_load(reinterpret_cast<const long long *>(&D.AI[id]));
        const int lo = static_cast<int>(static_cast<unsigned long long>(w) & 0xffffffffull);
        const int hi = static_cast<int>(static_cast<unsigned long long>(w) >> 32);
        l = static_cast<double>(lo);
        u = static_cast<double>(hi);
        if (lo == kAiHigh || hi == kAiLow) {
            const double xl = coherent_load_d(&D.AX[id].x);
            const double xu = coherent_load_d(&D.AX[id].y);
            l = xl > l ? xl : l;
            u = xu < u ? xu : u;
        }
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
        // both sides in one device-coherent 8-byte store (x in the low word)
        const int lo = fits_int32(l) ? static_cast<int>(l) : -2147483647;
        const int hi = fits_int32(u) ? static_cast<int>(u) : 2147483647;
        const unsigned long long w = static_cast<unsigned long long>(static_cast<unsigned>(lo)) |
                                     (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32);
        coherent_store(reinterpret_cast<long long *>(&D.S32[id - M.n16]), static_cast<long long>(w));
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
__device__ __forceinline__ bool apply_entry(const DevMatrix &M, const Sink &K, int jp, double a, double l, double u,
                                            double lhs, double rhs, double minact, double maxact, int mininf,
                                            int maxinf) {
    const bool is_cont = jp >= 0;  // bit 31 set marks an integer column
    double nl = l, nu = u;
    const int t = tighten_variable(a, lhs, rhs, minact, maxact, mininf, maxinf, is_cont, l, u, nl, nu);
    if (t == 0) return false;
    accumulate(M, K, jp & kIdMask, t, nl, nu);
    return true;
}

// ---------------------------------------------------------------------------
// lane entries
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
        // every gather first (entries past the end read column 0's bounds and
        // are discarded), then the decoding and the contributions; binary
        // columns held in LDS are read there instead
        unsigned off[K];
        uint4 chunk[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (first + k >= len) {
                jp[k] = 0;
                a[k] = 0.0;
            }
            off[k] = record_offset(M, jp[k] & kIdMask);
            chunk[k] = make_uint4(0u, 0u, 0u, 0u);
            if ((jp[k] & kIdMask) >= D.n_lds) chunk[k] = load_chunk(D, off[k]);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int id = jp[k] & kIdMask;
            if (id < D.n_lds) lds_binary_bounds(id, l[k], u[k]);
            else decode_bounds(M, id, off[k], chunk[k], l[k], u[k]);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            smin[k] = 0.0;
            smax[k] = 0.0;
            if (first + k >= len) {
                l[k] = 0.0;
                u[k] = 0.0;
            } else {
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

    // number of accepted candidates
    __device__ __forceinline__ int apply(const DevMatrix &M, const Sink &S, int first, int len, double lhs,
                                         double rhs, double minact, double maxact, int ni, int xi) {
        int moved = 0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (first + k < len) moved += apply_entry(M, S, jp[k], a[k], l[k], u[k], lhs, rhs, minact, maxact, ni, xi);
        return moved;
    }
};

// ---------------------------------------------------------------------------
// units and steps
// ---------------------------------------------------------------------------

// Value known to be the same on every lane of the wave (kept in a scalar
// register).
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int *status_row(const DevState &D, int step) {
    return D.status + kStatusWords * (step % kStatusRows);
}

// Whether unit u works in `step`, and whether it processes every row.
struct UnitStep {
    bool active;
    bool all_rows;
    int start;  // step of the unit's round 1
};

// The rule for a unit working in `step`, from its state words: loaded, not
// stopped by its load, under its round cap, and in its first round or moved a
// bound in the previous step without becoming infeasible.
__device__ __forceinline__ bool unit_works(int step, int start, int infeasible, int max_rounds, int crossed,
                                           int envelope, int prev_changed) {
    return start > 0 && step >= start && !crossed && !envelope && step - start < max_rounds &&
           (step == start || (prev_changed == step - 1 && infeasible == 0));
}

// unit_step for a different unit on every lane.
__device__ __forceinline__ UnitStep unit_step_lane(const DevMatrix &M, const DevState &D, int u, int step) {
    const long long *p = reinterpret_cast<const long long *>(&D.units[u]);
    const long long w0 = coherent_load(p), w1 = coherent_load(p + 1), w2 = coherent_load(p + 2),
                    w3 = coherent_load(p + 3);
    const int start = static_cast<int>(w0), infeasible = static_cast<int>(w0 >> 32);
    const int markall = static_cast<int>(w1), max_rounds = static_cast<int>(w1 >> 32);
    const int crossed = static_cast<int>(w2), envelope = static_cast<int>(w2 >> 32);
    const int prev_changed = static_cast<int>((step - 1) & 1 ? (w3 >> 32) : w3);
    UnitStep r;
    r.active = unit_works(step, start, infeasible, max_rounds, crossed, envelope, prev_changed);
    r.all_rows = !M.mark_rows || markall == step;
    r.start = start;
    return r;
}

// The same unit on every lane of the wave (result in scalar registers).
__device__ __forceinline__ UnitStep unit_step(const DevMatrix &M, const DevState &D, int u, int step) {
    const UnitStep l = unit_step_lane(M, D, uniform(u), step);
    UnitStep r;
    r.active = uniform(l.active ? 1 : 0) != 0;
    r.all_rows = uniform(l.all_rows ? 1 : 0) != 0;
    r.start = uniform(l.start);
    return r;
}

// Whether item b of a working unit (state V) has work in `step`: every item
// when the unit processes all rows; otherwise a group item whose item mark,
// or a seg / long item whose row mark, names the step.  Items still test
// their rows one by one, so an item visited needlessly costs time, never a
// different result.
__device__ __forceinline__ bool item_has_work(const DevMatrix &M, const DevState &V, int b, int step,
                                              bool all_rows) {
    if (all_rows) return true;
    const int row = M.item_row[b];
    const int mark = row >= 0 ? coherent_load(&V.marks[row]) : coherent_load(&V.item_mark[b]);
    return mark == step;
}

// Counts the wave's accepted candidates (`moved` per lane) of unit u in
// `step` in one of the unit's kMoveLanes counters of the step (one
// non-returning atomic per wave).  Whether the unit moved at all is recorded
// by the publish phase, from its bitmap, so that nothing writes the unit's
// state words while the items of a step read them.
__device__ __forceinline__ void note_change(const DevState &D, int u, int step, int moved) {
    if (!__any(moved != 0)) return;
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) moved += __shfl_xor(moved, h, 64);
    if ((threadIdx.x & 63) == 0) {
        const int lane_slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kMoveLanes;
        atomic_add_i(&D.moved[moved_index(u, step % kMoveSlots, lane_slot)], moved);
    }
}

__device__ __forceinline__ Sink make_sink(const DevState &V) {
    Sink k;
    k.V = V;
    return k;
}

// ---------------------------------------------------------------------------
// sweep items
// ---------------------------------------------------------------------------

// Rows [first, first + nrows) of one tree-width bin: groups of 2^logg lanes,
// K consecutive entries per lane.  Rows not marked for this step are
// skipped (a pass with no marked row loads nothing).
template <int K>
__device__ __forceinline__ int group_item(const DevMatrix &M, const Sink &S, int first, int nrows, int logg,
                                          int step, bool all_rows) {
    const DevState &V = S.V;
    const int lane = threadIdx.x & 63;
    const int g = 1 << logg;
    const int gl = lane & (g - 1);
    const int per_pass = 64 >> logg;
    const int end = first + nrows;
    int moved = 0;
    for (int base = first; base < end; base += per_pass) {
        const int r = base + (lane >> logg);
        const bool take = r < end && (all_rows || V.marks[r] == step);
        if (!__any(take)) continue;
        int start = 0, len = 0;
        double lhs = 0.0, rhs = 0.0;
        if (take) {
            row_extent(M, r, start, len);
            const double2 sd = M.sides[r];
            lhs = sd.x;
            rhs = sd.y;
        }
        LaneEntries<K> E;
        E.load(M, V, start, gl * K, len);
        E.reduce(g);
        const double minact = 0.0 + E.tmin;
        const double maxact = 0.0 + E.tmax;
        const bool active = take && len > 0;
        const GateBounds gb = gate_bounds(minact, maxact, E.mininf, E.maxinf, lhs, rhs, E.dmax);
        bool gate = gb.hi;
        const bool undecided = active && !gb.decided();
        if (__any(undecided)) {
            const double dm = E.eps_scan(g, 0.0);
            if (undecided) gate = row_gate(minact, maxact, E.mininf, E.maxinf, lhs, rhs, dm);
        }
        if (active && gate) moved += E.apply(M, S, gl * K, len, lhs, rhs, minact, maxact, E.mininf, E.maxinf);
    }
    return moved;
}

// One row that is not split: one wave, segment by segment, then a second
// pass for the candidates.
__device__ __forceinline__ int long_item(const DevMatrix &M, const Sink &S, int r) {
    constexpr int K = kSegLen / 64;
    const DevState &V = S.V;
    const int lane = threadIdx.x & 63;
    int start, len;
    row_extent(M, r, start, len);
    const double2 sd = M.sides[r];
    double acc_min = 0.0, acc_max = 0.0, dmax = 0.0;
    int mininf = 0, maxinf = 0;
    for (int s0 = 0; s0 < len; s0 += kSegLen) {
        LaneEntries<K> E;
        E.load(M, V, start + s0, lane * K, len - s0);
        E.reduce(64);
        acc_min = acc_min + E.tmin;
        acc_max = acc_max + E.tmax;
        mininf += E.mininf;
        maxinf += E.maxinf;
        dmax = fmax(dmax, E.dmax);
    }
    int moved = 0;
    if (decide_gate(M, V, start, len, acc_min, acc_max, mininf, maxinf, sd.x, sd.y, dmax)) {
        for (int e = lane; e < len; e += 64) {
            const int jp = M.col[start + e];
            double l, u;
            load_bounds(M, V, jp & kIdMask, l, u);
            moved += apply_entry(M, S, jp, M.val[start + e], l, u, sd.x, sd.y, acc_min, acc_max, mininf, maxinf);
        }
    }
    return moved;
}

// Sequential epsilon-maximum of a split row, computed by its last segment
// wave.  A segment whose exact maximum Ms satisfies !(Ms - run > eps) cannot
// replace the running value (floating-point subtraction is monotone), so it
// is skipped; the others are scanned entry by entry in order.
__device__ __forceinline__ double split_eps_scan(const DevMatrix &M, const DevState &V, const int4 info, int start, int len) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    double run = 0.0;
    for (int b = 0; b < info.z; b += 64) {
        const int q = b + lane;
        const double ms = q < info.z ? coherent_load_d(&V.seg_parts[info.y + q].dmax) : 0.0;
        const int cnt = min(64, info.z - b);
        for (int t = 0; t < cnt; t++) {
            const double mt = __shfl(ms, t, 64);
            if (!eps_gt(mt, run)) continue;
            const int s0 = (b + t) * kSegLen;
            LaneEntries<K> E;
            E.load(M, V, start + s0, lane * K, len - s0);
            run = E.eps_scan(64, run);
        }
    }
    return run;
}

// Stores the activity partial of segment `part` and the step that computed
// it (lane 0 of the wave).
__device__ __forceinline__ void store_part(const DevState &V, int part, const LaneEntries<kSegLen / 64> &E,
                                           int step) {
    SegPart *p = &V.seg_parts[part];
    coherent_store_d(&p->smin, E.tmin);
    coherent_store_d(&p->smax, E.tmax);
    coherent_store_d(&p->dmax, E.dmax);
    coherent_store_2i(&p->infs, E.mininf | (E.maxinf << 16), step);
}

// Segments [s, s + cnt) of split row lr (cnt <= kSegsPerItem) for unit u.
// The wave reduces each kSegLen-entry segment like a kSegLen-wide group row,
// stores the partials and counts them in; nobody waits.  The wave that
// completes the row's count adds the partials in segment order and decides
// the row gate; when it passes, the row's activity is stored and its
// segments are queued for the apply phase that follows the sweep.
__device__ __forceinline__ void seg_item(const DevMatrix &M, const DevState &D, const DevState &V, int u, int lr, int s, int cnt,
                         int *st, int step, bool all_rows, int unit_start) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    const int4 info = M.seg_rows[lr];
    const int r = info.x;
    int start, len;
    row_extent(M, r, start, len);
    // a segment computed earlier in this run none of whose columns moved in
    // the previous step keeps its partial (its columns' bounds have not
    // changed since it was computed: every change marks it); lane q reads
    // segment q's stamp and mark
    unsigned long long keep = 0ull;
    if (!all_rows) {
        bool kept = false;
        if (lane < cnt) {
            const int part = info.y + s + lane;
            const int stamp = coherent_load(&V.seg_parts[part].stamp);
            const int mark = coherent_load(&V.seg_marks[part]);
            kept = stamp >= unit_start && mark != step;
        }
        keep = __ballot(kept);
    }
    for (int q = 0; q < cnt; q++) {
        const int part = info.y + s + q;
        if (!all_rows && ((keep >> q) & 1ull)) continue;
        const int s0 = (s + q) * kSegLen;
        LaneEntries<K> E;
        E.load(M, V, start + s0, lane * K, len - s0);
        E.reduce(64);
        if (lane == 0) store_part(V, part, E, step);
    }
    unsigned old = 0;
    if (lane == 0) {
        mem_wait();
        old = __hip_atomic_fetch_add(&V.seg_count[lr], static_cast<unsigned>(cnt), __ATOMIC_RELAXED,
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
            const SegPart *pq = &V.seg_parts[info.y + q];
            pmin = coherent_load_d(&pq->smin);
            pmax = coherent_load_d(&pq->smax);
            dm = fmax(dm, coherent_load_d(&pq->dmax));
            int infs, stamp;
            coherent_load_2i(&pq->infs, infs, stamp);
            (void)stamp;
            ni += infs & 0xffff;
            xi += infs >> 16;
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
    if (!gb.decided()) gate = row_gate(am, ax, ni, xi, sd.x, sd.y, split_eps_scan(M, V, info, start, len));
    int pos = 0;
    if (lane == 0) {
        coherent_store(&V.seg_count[lr], 0u);
        if (gate) {
            SegState *stp = &V.seg_state[lr];
            coherent_store_d(&stp->minact, am);
            coherent_store_d(&stp->maxact, ax);
            coherent_store_2i(&stp->mininf, ni, xi);
            pos = atomic_add_i(&st[kStatusApply], info.z);
        }
    }
    if (!gate) return;
    pos = __shfl(pos, 0, 64);
    const int base = u * M.n_seg_parts + info.y;
    for (int q = lane; q < info.z; q += 64) coherent_store(&D.apply_tasks[pos + q], base + q);
}

// Unit and item of work index `it` (replicas: the launch's listed units,
// n_items each; blocks: the item, whose unit it names).
__device__ __forceinline__ void decode_work(const DevMatrix &M, const DevState &D, int it, int &u, int &b) {
    if (M.replicas) {
        const int k = it / M.n_items;
        u = D.active[k];
        b = it - k * M.n_items;
    } else {
        b = it;
        u = M.items[b].x >> kItemUnitShift;
    }
}

// Sweep item `it` of `step` for a unit working in the step (us: its state):
// skips unmarked rows, runs the item and records the unit's change.
__device__ __forceinline__ void sweep_item(const DevMatrix &M, const DevState &D, int it, int step, int *st,
                                           const UnitStep &us) {
    int u, b;
    decode_work(M, D, uniform(it), u, b);
    u = uniform(u);
    b = uniform(b);
    const DevState V = unit_view(M, D, u);
    const Sink S = make_sink(V);
    int4 d = M.items[b];
    d.x = uniform(d.x);
    d.y = uniform(d.y);
    d.z = uniform(d.z);
    d.w = uniform(d.w);
    const int kind = d.x & 15;
    int moved = 0;
    if (kind == kItemSeg) {
        const int r = M.seg_rows[d.y].x;
        if (us.all_rows || V.marks[r] == step) seg_item(M, D, V, u, d.y, d.z, d.w, st, step, us.all_rows, us.start);
        return;
    }
    if (kind == kItemLong) {
        if (us.all_rows || V.marks[d.y] == step) moved = long_item(M, S, d.y);
    } else {
        const int logg = (d.x >> 4) & 15;
        const int logk = (d.x >> 8) & 15;
        switch (logk) {
            case 0: moved = group_item<1>(M, S, d.y, d.z, logg, step, us.all_rows); break;
            case 1: moved = group_item<2>(M, S, d.y, d.z, logg, step, us.all_rows); break;
            default: moved = group_item<4>(M, S, d.y, d.z, logg, step, us.all_rows); break;
        }
    }
    note_change(D, u, step, moved);
}

// Applies the candidates of queued task `task` = unit * n_seg_parts + segment
// (one wave; the row's activity is read back from its SegState).
__device__ __forceinline__ void apply_split_task(const DevMatrix &M, const DevState &D, int task, int step, int *st) {
    constexpr int K = kSegLen / 64;
    const int lane = threadIdx.x & 63;
    task = uniform(task);
    const int u = task / M.n_seg_parts;
    const int part = task - u * M.n_seg_parts;
    const DevState V = unit_view(M, D, u);
    const Sink S = make_sink(V);
    const int lr = M.part_row[part];
    const int4 info = M.seg_rows[lr];
    const int r = info.x;
    int start, len;
    row_extent(M, r, start, len);
    const int s0 = (part - info.y) * kSegLen;
    const SegState *stp = &V.seg_state[lr];
    const double minact = coherent_load_d(&stp->minact);
    const double maxact = coherent_load_d(&stp->maxact);
    int ni, xi;
    coherent_load_2i(&stp->mininf, ni, xi);
    const double2 sd = M.sides[r];
    int moved = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = s0 + lane * K + k;
        if (e < len) {
            const int jp = M.col[start + e];
            double l, uu;
            load_bounds(M, V, jp & kIdMask, l, uu);
            moved += apply_entry(M, S, jp, M.val[start + e], l, uu, sd.x, sd.y, minact, maxact, ni, xi);
        }
    }
    note_change(D, u, step, moved);
}

// The apply phase: the waves of the grid take the queued segments in turn.
__device__ __forceinline__ void apply_phase(const DevMatrix &M, const DevState &D, int step, int wave, int nwaves) {
    int *st = status_row(D, step);
    const int count = coherent_load(&st[kStatusApply]);
    for (int t = wave; t < count; t += nwaves) apply_split_task(M, D, coherent_load(&D.apply_tasks[t]), step, st);
}

// ---------------------------------------------------------------------------
// publish phase
// ---------------------------------------------------------------------------

// Publishes bitmap word w of unit u (one half-wave of 32 lanes, lane b owns
// column 32 w + b; `bits` is the word as read by the scan): copies the
// accumulated bounds of the moved columns into the unit's snapshot, records
// a crossing as the unit's infeasible step and marks the rows of the moved
// columns for the next step.  A unit whose accepted candidates in this step
// reach its threshold processes all rows next step instead (no marks are
// pushed then).  Clears the word.
// Whether unit u processes all rows in the step after `step` (its accepted
// candidates of `step` reached its limit; the unit's all-rows step is set
// from the same counters by the publish phase's per-unit pass).  Lanes 0..15
// of each half-wave read the unit's counters: call with every lane of the
// wave, `valid` per half.
__device__ __forceinline__ bool decide_all_rows(const DevMatrix &M, const DevState &D, int u, bool valid, int step) {
    const int b = threadIdx.x & 31;
    int cnt = 0;
    if (valid && M.mark_rows && b < kMoveLanes)
        cnt = coherent_load(&D.moved[moved_index(u, step % kMoveSlots, b)]);
#pragma unroll
    for (int h = 1; h < kMoveLanes; h <<= 1) cnt += __shfl_xor(cnt, h, 64);
    cnt = __shfl(cnt, threadIdx.x & 32, 64);
    const int limit = valid ? coherent_load(&D.units[u].mark_limit) : 0;
    return valid && M.mark_rows && cnt >= limit;
}

__device__ __forceinline__ void publish_word(const DevMatrix &M, const DevState &D, int u, int w, unsigned bits,
                                             bool valid, int step, int all_rows_known) {
    const int lane = threadIdx.x & 63;
    const int b = lane & 31;
    if (valid && !M.replicas) u = M.word_unit[w];
    const DevState V = unit_view(M, D, u);
    if (valid && b == 0) coherent_store(&V.bm[w], 0u);
    const bool all_rows = all_rows_known >= 0 ? all_rows_known != 0 : decide_all_rows(M, D, u, valid, step);
    const bool moved = valid && ((bits >> b) & 1u);
    const int id = w * 32 + b;
    unsigned code = 0u;
    if (moved) {
        double l, uu;
        accumulated(M, V, id, l, uu);
        store_snapshot(M, V, id, l, uu);
        if (eps_gt(l, uu)) atomic_max_i(&D.units[u].infeasible, step);
        if (id < D.n_lds) code = (l >= 1.0 ? 1u : 0u) | (uu < 1.0 ? 2u : 0u);
    }
    if (D.n_lds > 0) {
        // the launch reads these columns from LDS (one unit, reloaded at the
        // start of every propagation): the new code bits of each 16
        // columns, ORed into their table word
        unsigned cw = code << ((id & 15) * 2);
#pragma unroll
        for (int h = 1; h < 16; h <<= 1) cw |= static_cast<unsigned>(__shfl_xor(static_cast<int>(cw), h, 64));
        if ((b & 15) == 0 && cw != 0u) atomic_or_u(&V.SB[id >> 4], cw);
    }
    if (!M.mark_rows || !valid || all_rows) return;
    // rows of the moved columns, flattened over the 32 lanes of the half:
    // exclusive prefix of the columns' lengths, then lane i marks entries
    // i, i + 32, ... found by a binary search over the prefix
    const int c0 = moved ? M.col_ptr[id] : 0;
    const int cn = moved ? M.col_ptr[id + 1] - c0 : 0;
    int pre = cn;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const int v = __shfl_up(pre, d, 32);
        if (b >= d) pre += v;
    }
    const int total = __shfl(pre, 31, 32);
    pre -= cn;  // exclusive
    // kMarkBatch entries per lane at a time: their index loads are all issued
    // before the marks they name are stored
    constexpr int kMarkBatch = 4;
    for (int base = 0; base < total; base += 32 * kMarkBatch) {  // same trip count on every lane of the half
        int row[kMarkBatch], sp[kMarkBatch], ci[kMarkBatch];
#pragma unroll
        for (int q = 0; q < kMarkBatch; q++) {
            const int i = base + q * 32 + b;
            int lo = 0;
#pragma unroll
            for (int st = 16; st > 0; st >>= 1) {
                const int cand = lo + st;
                if (cand < 32 && __shfl(pre, cand, 32) <= i) lo = cand;
            }
            const int k = __shfl(c0, lo, 32) + i - __shfl(pre, lo, 32);
            row[q] = -1;
            sp[q] = -1;
            ci[q] = -1;
            if (i < total) {
                row[q] = M.col_rows[k];
                sp[q] = M.col_seg[k];
                ci[q] = M.col_item[k];
            }
        }
#pragma unroll
        for (int q = 0; q < kMarkBatch; q++) {
            if (row[q] >= 0) coherent_store(&V.marks[row[q]], step + 1);
            if (sp[q] >= 0) coherent_store(&V.seg_marks[sp[q]], step + 1);
            if (ci[q] >= 0) coherent_store(&V.item_mark[ci[q]], step + 1);
        }
    }
}

// Publish phase of `step` over the waves of the grid: every wave scans 8
// bitmap words of one unit (replicas: the launch's units; blocks: the one
// shared bitmap) and publishes the nonzero ones two at a time; a unit with
// accepted candidates records that it, and the step, changed.  Also clears
// the next step's status row and the candidate counters two steps ahead.
__device__ __forceinline__ void publish_phase(const DevMatrix &M, const DevState &D, int step, int wave,
                                              int nwaves, int n_units) {
    constexpr int kChunk = 8;  // bitmap words per wave
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int chunks = (M.nwords + kChunk - 1) / kChunk;
    const int n_chunks = (M.replicas ? D.n_active : 1) * chunks;
    // One unit in the launch (a single problem): the all-rows decision is
    // taken once per workgroup (its first wave reads the unit's counters,
    // final after the sweep's barrier) and shared through LDS instead of
    // once per chunk after the chunk's words arrive.
    __shared__ int single_all_rows;
    const bool single = M.replicas && D.n_active == 1;
    if (single) {
        if (threadIdx.x < 64) {
            const bool all_rows = decide_all_rows(M, D, uniform(D.active[0]), true, step);
            if (threadIdx.x == 0) single_all_rows = all_rows ? 1 : 0;
        }
        __syncthreads();
    }
    for (int c = wave; c < n_chunks; c += nwaves) {
        const int k = c / chunks;
        const int q = c - k * chunks;
        const int u = M.replicas ? uniform(D.active[k]) : 0;
        const DevState V = unit_view(M, D, u);
        const int w = q * kChunk + lane;
        const unsigned bits = lane < kChunk && w < M.nwords ? coherent_load(&V.bm[w]) : 0u;
        unsigned long long pending = __ballot(bits != 0u);
        if (!pending) continue;
        // one decision for the chunk's unit (replicas; blocks: when the
        // chunk's words all belong to one unit, as in a single problem)
        int known = -1;
        if (single) {
            known = uniform(single_all_rows);
        } else if (M.replicas) {
            known = decide_all_rows(M, D, u, true, step) ? 1 : 0;
        } else {
            const int u0 = uniform(M.word_unit[q * kChunk]);
            const bool other = lane < kChunk && w < M.nwords && M.word_unit[w] != u0;
            if (!__ballot(other)) known = decide_all_rows(M, D, u0, true, step) ? 1 : 0;
        }
        if (M.replicas && (known == 1 || !M.mark_rows)) {
            // no row marks to push: the chunk's nonzero words all at once,
            // their accumulator reads issued before any is used (lane b of
            // each half takes bit b of the half's word of every pair)
            constexpr int kPairs = kChunk / 2;
            int ids[kPairs];
            bool mv[kPairs];
            AccRaw raw[kPairs];
#pragma unroll
            for (int i = 0; i < kPairs; i++) {
                int t0 = -1, t1 = -1;
                if (pending) {
                    t0 = __ffsll(static_cast<long long>(pending)) - 1;
                    pending &= pending - 1;
                }
                if (pending) {
                    t1 = __ffsll(static_cast<long long>(pending)) - 1;
                    pending &= pending - 1;
                }
                const int t = half ? t1 : t0;
                const int wt = q * kChunk + (t < 0 ? 0 : t);
                const unsigned wb = __shfl(bits, t < 0 ? 0 : t, 64);
                if (t >= 0 && (lane & 31) == 0) coherent_store(&V.bm[wt], 0u);
                ids[i] = wt * 32 + (lane & 31);
                mv[i] = t >= 0 && ((wb >> (lane & 31)) & 1u);
                raw[i] = mv[i] ? acc_load(M, V, ids[i]) : AccRaw{0, 0.0, 0.0};
            }
#pragma unroll
            for (int i = 0; i < kPairs; i++) {
                unsigned code = 0u;
                if (mv[i]) {
                    double l, uu;
                    acc_resolve(M, V, ids[i], raw[i], l, uu);
                    store_snapshot(M, V, ids[i], l, uu);
                    if (eps_gt(l, uu)) atomic_max_i(&D.units[u].infeasible, step);
                    if (ids[i] < D.n_lds) code = (l >= 1.0 ? 1u : 0u) | (uu < 1.0 ? 2u : 0u);
                }
                if (D.n_lds > 0) {
                    unsigned cw = code << ((ids[i] & 15) * 2);
#pragma unroll
                    for (int h = 1; h < 16; h <<= 1)
                        cw |= static_cast<unsigned>(__shfl_xor(static_cast<int>(cw), h, 64));
                    if ((lane & 15) == 0 && cw != 0u) atomic_or_u(&V.SB[ids[i] >> 4], cw);
                }
            }
            continue;
        }
        while (pending) {
            const int t0 = __ffsll(static_cast<long long>(pending)) - 1;
            pending &= pending - 1;
            int t1 = -1;
            if (pending) {
                t1 = __ffsll(static_cast<long long>(pending)) - 1;
                pending &= pending - 1;
            }
            const int t = half ? t1 : t0;
            const unsigned wb = __shfl(bits, t < 0 ? 0 : t, 64);
            publish_word(M, D, u, q * kChunk + (t < 0 ? 0 : t), wb, t >= 0, step, known);
        }
    }
    // The units that accepted candidates in this step moved a bound: their
    // candidate counters (final since the sweep's barrier) set the unit's
    // and the step's changed flags, and the unit's all-rows step when they
    // reach its limit, one wave per unit, so that the words above need no
    // flag reads.
    const int n_flag_units = M.replicas ? D.n_active : n_units;
    for (int k = wave; k < n_flag_units; k += nwaves) {
        const int uu = M.replicas ? uniform(D.active[k]) : k;
        int cnt = lane < kMoveLanes
                      ? coherent_load(&D.moved[moved_index(uu, step % kMoveSlots, lane)])
                      : 0;
        const int limit = lane == 0 ? coherent_load(&D.units[uu].mark_limit) : 0;
#pragma unroll
        for (int h = 1; h < kMoveLanes; h <<= 1) cnt += __shfl_xor(cnt, h, 64);
        if (lane == 0 && cnt > 0) {
            coherent_store(&D.units[uu].changed[step & 1], step);
            if (M.mark_rows && cnt >= limit) coherent_store(&D.units[uu].markall, step + 1);
            // read first: many units set it
            if (coherent_load(&status_row(D, step)[kStatusChanged]) == 0)
                atomic_or_i(&status_row(D, step)[kStatusChanged], 1);
        }
    }
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    if (tid < kStatusWords) coherent_store(&status_row(D, step + 1)[tid], 0);
    for (int k = tid; k < n_units * kMoveLanes; k += stride) {
        const int uu = k / kMoveLanes, j = k - uu * kMoveLanes;
        coherent_store(&D.moved[moved_index(uu, (step + 2) % kMoveSlots, j)], 0);
    }
}

// ---------------------------------------------------------------------------
// step kernels (CPU_LOOP, GPU_LOOP)
// ---------------------------------------------------------------------------

// With `guard` (captured graphs), a launch first checks on the device whether
// the previous step changed anything; if not, the run is over and the launch
// does nothing.
__device__ __forceinline__ bool step_skipped(const DevState &D, int step, int first_step) {
    if (coherent_load(&D.global[kGlobalError]) != 0) return true;
    if (step <= first_step) return false;
    return coherent_load(&status_row(D, step - 1)[kStatusChanged]) == 0;
}

// The sweep of one step over a grid sized to the device (grid stride): each
// wave tests 64 of its work items at once (one lane each: the unit's state
// and the item's mark) and visits those with work, as the persistent kernel
// does.  Many units' items then cost a few loads each instead of a wave.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void sweep_kernel(
    DevMatrix M, DevState D_arg, int step, int guard, int first_step) {
    // no LDS in the step kernels (a constant, so that the LDS reads compile out)
    DevState D = D_arg;
    D.n_lds = 0;
    if (guard && step_skipped(D, step, first_step)) return;
    const int waves = blockDim.x >> 6;
    const int wave = uniform(blockIdx.x * waves + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * waves;
    const int n_work = M.replicas ? M.n_items * D.n_active : M.n_items;
    const int lane = threadIdx.x & 63;
    int *st = status_row(D, step);
    for (int base = wave; base < n_work; base += 64 * nwaves) {
        const int it = base + lane * nwaves;
        bool on = false;
        if (it < n_work) {
            int u, b;
            decode_work(M, D, it, u, b);
            const UnitStep us = unit_step_lane(M, D, u, step);
            on = us.active && item_has_work(M, unit_view(M, D, u), b, step, us.all_rows);
        }
        unsigned long long pending = __ballot(on);
        while (pending) {
            const int t = __ffsll(static_cast<long long>(pending)) - 1;
            pending &= pending - 1;
            const int itt = uniform(base + t * nwaves);
            int uu, bb;
            decode_work(M, D, itt, uu, bb);
            sweep_item(M, D, itt, step, st, unit_step(M, D, uu, step));
        }
    }
}

__global__ __launch_bounds__(256) void apply_split_kernel(DevMatrix M, DevState D_arg, int step, int guard,
                                                          int first_step) {
    DevState D = D_arg;
    D.n_lds = 0;
    if (guard && step_skipped(D, step, first_step)) return;
    const int waves = blockDim.x >> 6;
    apply_phase(M, D, step, uniform(blockIdx.x * waves + (threadIdx.x >> 6)), gridDim.x * waves);
}

__global__ __launch_bounds__(256) void publish_kernel(DevMatrix M, DevState D_arg, int step, int guard,
                                                      int first_step, int n_units) {
    DevState D = D_arg;
    D.n_lds = 0;
    if (guard && step_skipped(D, step, first_step)) return;
    const int waves = blockDim.x >> 6;
    publish_phase(M, D, step, uniform(blockIdx.x * waves + (threadIdx.x >> 6)), gridDim.x * waves, n_units);
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
// Always inlined: as a call it makes the caller spill its DevState argument
// to scratch and lose a wave of occupancy in the persistent kernel.
__device__ __forceinline__ bool grid_barrier(const DevState &D, unsigned nblocks, bool invalidate) {
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
                if (coherent_load(&D.global[kGlobalError])) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > kWaitTimeoutTicks) {
                    atomic_or_i(&D.global[kGlobalError], 1);
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

// Steps [first_step, first_step + max_steps) in one persistent launch
// (MEGAKERNEL).  Every workgroup walks the work of a phase with a grid
// stride, all meet at a barrier, and all read the same status words, so all
// take the same decision to go on: the run ends after a step in which no
// unit moved a bound.  Phases of a step: sweep; apply of the split rows when
// the sweep queued any; publish.  The number of steps run goes to the global
// status block (-1 after a failed wait).
template <int kThreads, bool kLds>
__global__ __launch_bounds__(kThreads) void mega_kernel(DevMatrix M, DevState D_arg, int first_step, int max_steps,
                                                        int n_units) {
    // kLds: a launch of one unit that reads binary columns from LDS; without
    // it n_lds is a constant 0, and the LDS reads and code updates compile out
    DevState D = D_arg;
    if constexpr (!kLds) D.n_lds = 0;
    const unsigned nblocks = gridDim.x;
    const int waves = blockDim.x >> 6;
    const int wave = uniform(blockIdx.x * waves + (threadIdx.x >> 6));
    const int nwaves = gridDim.x * waves;
    const int n_work = M.replicas ? M.n_items * D.n_active : M.n_items;
    // Several units: at the start of a step every workgroup notes in LDS
    // which units work in it, which of them process all rows, and their
    // first steps (the state words do not change during the sweep), so that
    // an item's unit costs no device-coherent loads.  One unit: every wave
    // reads its state once per step.  Then each wave tests 64 of its items at
    // once (lanes): items of idle units, and items none of whose rows are
    // marked when their unit processes marked rows only, are not visited.
    // (in the dynamic LDS, kFilterLdsBytes of it in launches of several units)
    unsigned *works = dyn_lds;
    unsigned *all_rows_of = dyn_lds + kMaxFilterUnits / 32;
    int *start_of = reinterpret_cast<int *>(dyn_lds + 2 * (kMaxFilterUnits / 32));
    const int n_slots = M.replicas ? D.n_active : n_units;  // replicas: positions in the launch's list
    const bool filter = n_slots > 1 && n_slots <= kMaxFilterUnits;
    const int lane = threadIdx.x & 63;
    int steps = 0;
    for (int step = first_step; step < first_step + max_steps; step++) {
        int *st = status_row(D, step);
        UnitStep one{};
        if (n_slots == 1) {
            const int u1 = M.replicas ? uniform(D.active[0]) : 0;
            one = unit_step(M, D, u1, step);
            if (D.n_lds > 0) {
                // the unit's binary codes as published by the previous step
                // (the waves of this workgroup that read the LDS in the
                // previous step are past its barriers)
                const uint4 *src = reinterpret_cast<const uint4 *>(unit_view(M, D, u1).SB);
                uint4 *dst = reinterpret_cast<uint4 *>(dyn_lds);
                for (int i = threadIdx.x; i < D.n_lds / 64; i += blockDim.x) dst[i] = src[i];
                __syncthreads();
            }
        } else if (filter) {
            // the previous step's waves of this workgroup are past its
            // barriers, so no wave still reads the entries overwritten here
            for (int g = (threadIdx.x >> 6) * 64; g < n_slots; g += blockDim.x) {
                const int k = g + lane;
                UnitStep us{};
                if (k < n_slots) us = unit_step_lane(M, D, M.replicas ? D.active[k] : k, step);
                const unsigned long long bw = __ballot(k < n_slots && us.active);
                const unsigned long long ba = __ballot(k < n_slots && us.all_rows);
                if (lane == 0) {
                    works[g >> 5] = static_cast<unsigned>(bw);
                    works[(g >> 5) + 1] = static_cast<unsigned>(bw >> 32);
                    all_rows_of[g >> 5] = static_cast<unsigned>(ba);
                    all_rows_of[(g >> 5) + 1] = static_cast<unsigned>(ba >> 32);
                }
                if (k < n_slots) start_of[k] = us.start;
            }
            __syncthreads();
        }
        // the wave's items in the order wave, wave + nwaves, ...: 64 at a time
        if (n_slots != 1 || one.active) {
            for (int base = wave; base < n_work; base += 64 * nwaves) {
                const int it = base + lane * nwaves;
                bool on = it < n_work;
                int u = 0, b = 0;
                UnitStep us = one;
                if (on) {
                    decode_work(M, D, it, u, b);
                    if (filter) {
                        const int k = M.replicas ? it / M.n_items : u;
                        us.active = (works[k >> 5] >> (k & 31)) & 1u;
                        us.all_rows = (all_rows_of[k >> 5] >> (k & 31)) & 1u;
                        us.start = start_of[k];
                    } else if (n_slots != 1) {
                        us = unit_step_lane(M, D, u, step);
                    }
                    on = us.active && item_has_work(M, unit_view(M, D, u), b, step, us.all_rows);
                }
                unsigned long long pending = __ballot(on);
                while (pending) {
                    const int t = __ffsll(static_cast<long long>(pending)) - 1;
                    pending &= pending - 1;
                    // the picked item's unit state again, wave-uniform (no
                    // per-lane state stays live across the item)
                    const int itt = uniform(base + t * nwaves);
                    UnitStep ut = one;
                    if (n_slots != 1) {
                        int uu, bb;
                        decode_work(M, D, itt, uu, bb);
                        if (filter) {
                            const int k = uniform(M.replicas ? itt / M.n_items : uu);
                            ut.active = true;
                            ut.all_rows = (all_rows_of[k >> 5] >> (k & 31)) & 1u;
                            ut.start = start_of[k];
                        } else {
                            ut = unit_step(M, D, uu, step);
                        }
                    }
                    sweep_item(M, D, itt, step, st, ut);
                }
            }
        }
        if (!grid_barrier(D, nblocks, false)) { steps = -1; break; }
        if (coherent_load(&st[kStatusApply]) > 0) {
            apply_phase(M, D, step, wave, nwaves);
            if (!grid_barrier(D, nblocks, false)) { steps = -1; break; }
        }
        publish_phase(M, D, step, wave, nwaves, n_units);
        if (!grid_barrier(D, nblocks, true)) { steps = -1; break; }
        steps = step - first_step + 1;
        if (!coherent_load(&st[kStatusChanged])) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) D.global[kGlobalSteps] = steps;
}

template __global__ void mega_kernel<256, false>(DevMatrix, DevState, int, int, int);
template __global__ void mega_kernel<512, false>(DevMatrix, DevState, int, int, int);
template __global__ void mega_kernel<1024, false>(DevMatrix, DevState, int, int, int);
template __global__ void mega_kernel<256, true>(DevMatrix, DevState, int, int, int);
template __global__ void mega_kernel<512, true>(DevMatrix, DevState, int, int, int);
template __global__ void mega_kernel<1024, true>(DevMatrix, DevState, int, int, int);

// ---------------------------------------------------------------------------
// loading and reading bound vectors
// ---------------------------------------------------------------------------

// Resets the state words of the listed units, params[k] = {unit, mark
// limit, all rows in round 1, round cap}: round 1 in `start_step`, every row
// processed in that step unless marks come from changed-column hints.  With
// `clear_status`, also clears the step-status ring, the global words and the
// barrier.
__global__ void unit_reset_kernel(DevState D, int n_list, const int4 *params, int start_step, int clear_status) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int k = tid; k < n_list; k += stride) {
        const int4 q = params[k];
        UnitState s = {};
        s.start = start_step;
        s.markall = q.z ? start_step : 0;
        s.max_rounds = q.w;
        s.mark_limit = q.y;
        D.units[q.x] = s;
        for (int sl = 0; sl < kMoveSlots; sl++)
            for (int j = 0; j < kMoveLanes; j++) D.moved[moved_index(q.x, sl, j)] = 0;
    }
    if (clear_status) {
        for (int k = tid; k < kStatusRows * kStatusWords; k += stride) D.status[k] = 0;
        for (int k = tid; k < kGlobalWords; k += stride) D.global[k] = 0;
        for (int k = tid; k < kBarrierWords; k += stride) D.barrier[k] = 0u;
    }
}

// Loads the bound vectors of the listed units (caller's column order: unit
// k's bounds at lbp[k], ubp[k]; blocks mode: one concatenated vector) into
// snapshot and accumulator, clears their bitmaps, marks and split-row
// counters, and flags crossed input bounds and bounds outside a column's
// storage class.  Grid: x over internal ids, y over listed units.
__global__ __launch_bounds__(256) void load_kernel(DevMatrix M, DevState D, int n_list, const int *list,
                                                   const double *const *lbp, const double *const *ubp) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    const int n_pass = M.replicas ? n_list : 1;  // blocks mode: one pass over all columns
    for (int k = blockIdx.y; k < n_pass; k += gridDim.y) {
        const int u = list[k];
        const DevState V = unit_view(M, D, u);
        const double *lbs = lbp[M.replicas ? k : 0];
        const double *ubs = ubp[M.replicas ? k : 0];
        for (int id = tid; id < M.n; id += stride) {
            const int j = M.perm[id];
            const int cu = M.replicas ? u : M.word_unit[id >> 5];
            double l = 0.0, uu = 0.0;
            if (j >= 0) {
                l = lbs[j];
                uu = ubs[j];
                if (eps_gt(l, uu)) atomic_or_i(&D.units[cu].crossed, 1);
            }
            if (id < M.n_int) {
                const double lim = id < M.n16 ? 32767.0 : 2147483647.0;
                const double lo_lim = id < M.n_bin ? 0.0 : -lim, hi_lim = id < M.n_bin ? 1.0 : lim;
                const bool fits = l >= lo_lim && l <= hi_lim && uu >= lo_lim && uu <= hi_lim && floor(l) == l &&
                                  floor(uu) == uu;
                // crossed input bounds make the vector infeasible, whatever
                // the column's storage class
                if (!fits && !eps_gt(l, uu)) atomic_or_i(&D.units[cu].envelope, 1);
                const int il = fits ? static_cast<int>(l) : 0, iu = fits ? static_cast<int>(uu) : 0;
                V.AI[id] = make_int2(il, iu);
                V.AX[id] = make_double2(-INFINITY, INFINITY);
                if (id < M.n16) V.S16[id] = pack16(il, iu);
                else V.S32[id - M.n16] = make_int2(il, iu);
            } else {
                V.AF[id - M.n_int] = make_double2(l, uu);
                V.SF[id - M.n_int] = make_double2(l, uu);
            }
        }
        // 2-bit codes of the binary columns, padding words zero
        for (int w = tid; w < M.sb_words; w += stride) {
            unsigned cw = 0u;
            for (int t = 0; t < 16; t++) {
                const int id = w * 16 + t;
                if (id < M.n_bin) {
                    const int j = M.perm[id];
                    const double l = lbs[j], uu = ubs[j];
                    cw |= ((l >= 1.0 ? 1u : 0u) | (uu < 1.0 ? 2u : 0u)) << (2 * t);
                }
            }
            V.SB[w] = cw;
        }
        {
            for (int w = tid; w < M.nwords; w += stride) V.bm[w] = 0u;
            for (int r = tid; r < M.m; r += stride) V.marks[r] = 0;
            for (int q = tid; q < M.n_items; q += stride) V.item_mark[q] = 0;
            for (int q = tid; q < M.n_seg_rows; q += stride) V.seg_count[q] = 0u;
            // segment marks too: their step stamps would otherwise match a
            // step of this run (steps restart with every run) and make it
            // recompute segments whose columns did not move
            for (int q = tid; q < M.n_seg_parts; q += stride) {
                V.seg_parts[q].stamp = 0;
                V.seg_marks[q] = 0;
            }
        }
    }
}

// Marks the rows of hinted columns (caller's column indices; unit k's hints
// at hints[hint_ptr[k] .. hint_ptr[k + 1])) for the listed units' first step:
// the bound vector differs from a propagated fixpoint only in these columns.
__global__ __launch_bounds__(256) void hint_kernel(DevMatrix M, DevState D, int n_list, const int *list,
                                                   const int *hint_ptr, const int *hints, const int *inv_perm,
                                                   int start_step) {
    for (int k = blockIdx.y; k < n_list; k += gridDim.y) {
        const DevState V = unit_view(M, D, list[k]);
        for (int h = hint_ptr[k] + blockIdx.x * blockDim.x + threadIdx.x; h < hint_ptr[k + 1];
             h += gridDim.x * blockDim.x) {
            const int id = inv_perm[hints[h]];
            for (int q = M.col_ptr[id]; q < M.col_ptr[id + 1]; q++) {
                V.marks[M.col_rows[q]] = start_step;
                if (M.col_item[q] >= 0) V.item_mark[M.col_item[q]] = start_step;
            }
        }
    }
}

// Writes the accumulated bounds of the listed units back to the caller's
// column order.
__global__ __launch_bounds__(256) void unpack_kernel(DevMatrix M, DevState D, int n_list, const int *list,
                                                     double *const *lbp, double *const *ubp) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    const int n_pass = M.replicas ? n_list : 1;  // blocks mode: one pass over all columns
    for (int k = blockIdx.y; k < n_pass; k += gridDim.y) {
        const DevState V = unit_view(M, D, list[k]);
        double *lbs = lbp[M.replicas ? k : 0];
        double *ubs = ubp[M.replicas ? k : 0];
        for (int id = tid; id < M.n; id += stride) {
            const int j = M.perm[id];
            if (j < 0) continue;
            double l, u;
            if (id < M.n_int) {
                const int2 a = V.AI[id];
                const double2 x = V.AX[id];
                l = static_cast<double>(a.x);
                u = static_cast<double>(a.y);
                l = x.x > l ? x.x : l;
                u = x.y < u ? x.y : u;
            } else {
                const double2 a = V.AF[id - M.n_int];
                l = a.x;
                u = a.y;
            }
            lbs[j] = l;
            ubs[j] = u;
        }
    }
}

// Writes the changed columns of delta requests into their bound vectors:
// vector k gets lbs[cols[h]] = lv[h], ubs[cols[h]] = uv[h] for h in
// [ptr[k], ptr[k + 1]) (caller's column order).
__global__ __launch_bounds__(256) void delta_kernel(int n_list, double *const *lbp, double *const *ubp,
                                                    const int *ptr, const int *cols, const double *lv,
                                                    const double *uv) {
    for (int k = blockIdx.y; k < n_list; k += gridDim.y)
        for (int h = ptr[k] + blockIdx.x * blockDim.x + threadIdx.x; h < ptr[k + 1]; h += gridDim.x * blockDim.x) {
            lbp[k][cols[h]] = lv[h];
            ubp[k][cols[h]] = uv[h];
        }
}

// Copies `count` bound-vector pairs of n2 doubles: ptrs[k] to ptrs[count + k]
// (16-byte accesses; n2 is even).  Grid: x over elements, y over pairs.
__global__ __launch_bounds__(256) void copy_pairs_kernel(long long n2, int count, double *const *ptrs) {
    const long long h = n2 / 2;
    for (int k = blockIdx.y; k < count; k += gridDim.y) {
        const double2 *src = reinterpret_cast<const double2 *>(ptrs[k]);
        double2 *dst = reinterpret_cast<double2 *>(ptrs[count + k]);
        for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < h;
             i += static_cast<long long>(gridDim.x) * blockDim.x)
            dst[i] = src[i];
    }
}

// Columns in which vector k (lbp[k], ubp[k]) differs from the base vector,
// bit for bit: their indices and bounds go to region k (capacity cap) of
// cols / vl / vu, their number to cnt[k] (which may exceed cap).  cnt must be
// zero on entry.
__global__ __launch_bounds__(256) void diff_kernel(int n, int n_list, double *const *lbp, double *const *ubp,
                                                   const double *const *blp, const double *const *bup, int cap,
                                                   int *cnt, int *cols, double *vl, double *vu) {
    for (int k = blockIdx.y; k < n_list; k += gridDim.y) {
        const double *l = lbp[k];
        const double *u = ubp[k];
        const double *base_l = blp[k];
        const double *base_u = bup[k];
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
            const double a = l[j], b = u[j];
            if (__double_as_longlong(a) == __double_as_longlong(base_l[j]) &&
                __double_as_longlong(b) == __double_as_longlong(base_u[j]))
                continue;
            const int pos = atomic_add_i(&cnt[k], 1);
            if (pos < cap) {
                const size_t q = static_cast<size_t>(k) * cap + pos;
                cols[q] = j;
                vl[q] = a;
                vu[q] = b;
            }
        }
    }
}

// Progress measure of one round (src/progressMeasureMethods/progress_function.h:45-116,
// :152-195): for every column the finite progress of its bounds between the
// weakest bounds (wl, wu) and the limit bounds (ll, lu), and the bounds that
// became finite since the previous round (pl, pu, then overwritten with this
// round's).  Bounds in the caller's column order.  Each workgroup writes the
// sums of its columns to fin_part / inf_part[blockIdx.x] in a fixed order.
__global__ __launch_bounds__(256) void measure_kernel(int n, const double *lb, const double *ub, const double *wl,
                                                      const double *wu, const double *ll, const double *lu,
                                                      double *pl, double *pu, double *fin_part, int *inf_part) {
    __shared__ double sf[4];
    __shared__ int si[4];
    double f = 0.0;
    int k = 0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double l = lb[j], u = ub[j], a = wl[j], b = wu[j], c = ll[j], d = lu[j];
        double v = 0.0;
        if (eps_gt(a, -kInf) && eps_gt(l, -kInf) && eps_ge(l, a) && !(fabs(c - a) <= kEps)) v += (l - a) / (c - a);
        if (eps_lt(b, kInf) && eps_lt(u, kInf) && eps_le(u, b) && !(fabs(b - d) <= kEps)) v += (b - u) / (b - d);
        f += v;
        k += (is_neg_inf(pl[j]) && eps_gt(l, -kInf)) ? 1 : 0;
        k += (is_pos_inf(pu[j]) && eps_lt(u, kInf)) ? 1 : 0;
        pl[j] = l;
        pu[j] = u;
    }
    for (int h = 32; h >= 1; h >>= 1) {
        f += shfl_xor_d(f, h);
        k += __shfl_xor(k, h, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sf[wave] = f;
        si[wave] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        fin_part[blockIdx.x] = (sf[0] + sf[1]) + (sf[2] + sf[3]);
        inf_part[blockIdx.x] = si[0] + si[1] + si[2] + si[3];
    }
}

// Sums the workgroup partials of measure_kernel in index order into
// fin[0], inf[0] (one wave).
__global__ __launch_bounds__(64) void measure_total_kernel(int parts, const double *fin_part, const int *inf_part,
                                                           double *fin, int *inf) {
    double f = 0.0;
    int k = 0;
    for (int q = threadIdx.x; q < parts; q += 64) {
        f += fin_part[q];
        k += inf_part[q];
    }
    for (int h = 32; h >= 1; h >>= 1) {
        f += shfl_xor_d(f, h);
        k += __shfl_xor(k, h, 64);
    }
    if (threadIdx.x == 0) {
        fin[0] = f;
        inf[0] = k;
    }
}


// Packs the changed columns that diff_kernel found for nd vectors (vector
// k's cnt[k] entries at [k * cap, ...), none when cnt[k] > cap) into
// consecutive positions: vector k's start into off[k], the total into
// off[nd].  One workgroup per vector; each sums the counts before its own.
__global__ __launch_bounds__(256) void compact_changes_kernel(int nd, int cap, const int *cnt, const int *cols,
                                                              const double *vl, const double *vu, int *off,
                                                              int *kcols, double *kl, double *ku) {
    __shared__ int part[256];
    for (int k = blockIdx.x; k < nd; k += gridDim.x) {
        int sum = 0;
        for (int j = threadIdx.x; j < k; j += blockDim.x) {
            const int c = cnt[j];
            sum += c <= cap ? c : 0;
        }
        part[threadIdx.x] = sum;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (static_cast<int>(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
            __syncthreads();
        }
        const int o = part[0];
        const int c = cnt[k] <= cap ? cnt[k] : 0;
        if (threadIdx.x == 0) {
            off[k] = o;
            if (k == nd - 1) off[nd] = o + c;
        }
        const size_t src = static_cast<size_t>(k) * cap;
        for (int q = threadIdx.x; q < c; q += blockDim.x) {
            kcols[o + q] = cols[src + q];
            kl[o + q] = vl[src + q];
            ku[o + q] = vu[src + q];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// problem layout on the device (single instance): the passes over the
// nonzeros.  The host orders the rows (tree-width bins), pads them and
// relabels the columns; these kernels move the entries.
// ---------------------------------------------------------------------------

// Row r of the layout takes the entries col_in / val_in [src_start[r],
// src_start[r] + len), len from the padded row starts row_ptr_out, of
// instance row_inst[r] (instance 0 when row_inst is null), whose columns are
// global columns inst_col_off[i] .. inst_col_off[i + 1] - 1.  Each entry goes
// to its padded place as (internal id | integer flag, coefficient) and counts
// one for its column in col_cnt (by internal id).  Padding entries are left
// as written before (zeros).  A column index outside the instance's range
// sets *bad and the entry is skipped.  One wave per row, grid-stride.
__global__ __launch_bounds__(256) void layout_rows_kernel(int mp, const long long *src_start, const int *row_inst,
                                                          const int *inst_col_off, const int *col_in,
                                                          const double *val_in, const int *row_ptr_out,
                                                          const int *internal, const int *vartype, int *col_out,
                                                          double *val_out, int *col_cnt, int *bad) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < mp; r += waves) {
        const long long k0 = src_start[r];
        const int a = row_ptr_out[r], b = row_ptr_out[r + 1];
        const int dst = a & ~(kRowAlign - 1);
        const int len = (b & ~(kRowAlign - 1)) - dst - (a & (kRowAlign - 1));
        const int inst = row_inst ? row_inst[r] : 0;
        const int base = inst_col_off[inst], lim = inst_col_off[inst + 1] - base;
        for (int e = lane; e < len; e += 64) {
            const int j = col_in[k0 + e];
            if (j < 0 || j >= lim) {
                atomic_or_i(bad, 1);
                continue;
            }
            const int g = base + j;
            const int id = internal[g];
            col_out[dst + e] = id | (vartype[g] == kVarContinuous ? 0 : kIntFlag);
            val_out[dst + e] = val_in[k0 + e];
            atomic_add_i(&col_cnt[id], 1);
        }
    }
}

// Column-to-row index: every entry of layout row r appends r to its column's
// list (fill[] starts at col_ptr[]), for split rows (row_part[r] >= 0) the
// entry's segment, and the group item holding the row (row_item[r], -1 for
// split and long rows).  The order inside a column's list is not fixed; the
// lists are used as sets (row, segment and item marking).  One wave per row.
__global__ __launch_bounds__(256) void layout_csc_kernel(int mp, const int *row_ptr_out, const int *col_out,
                                                         const int *row_part, const int *row_item, int *fill,
                                                         int *col_rows, int *col_seg, int *col_item) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < mp; r += waves) {
        const int a = row_ptr_out[r], b = row_ptr_out[r + 1];
        const int st = a & ~(kRowAlign - 1);
        const int len = (b & ~(kRowAlign - 1)) - st - (a & (kRowAlign - 1));
        const int part = row_part[r];
        const int item = row_item[r];
        for (int e = lane; e < len; e += 64) {
            const int id = col_out[st + e] & kIdMask;
            const int slot = atomic_add_i(&fill[id], 1);
            col_rows[slot] = r;
            col_seg[slot] = part >= 0 ? part + e / kSegLen : -1;
            col_item[slot] = item;
        }
    }
}

// Rows of at most 64 entries (the tail of the layout: rows are ordered by
// decreasing tree width), 64 rows per wave.  Lane t holds the start, length,
// source position and instance of row r0 + t; the wave walks the padded
// entries of its rows 64 at a time, and each lane finds the row of its entry
// by a binary search over the starts the lanes hold.  The padding entries
// (zeroed beforehand) are skipped.  Same results as the kernels above.
struct ShortRows {
    int nrows, st, len, inst, S0, S1;
    long long k0;
};

__device__ __forceinline__ ShortRows short_rows(int r0, int mp, const int *row_ptr_out, const long long *src_start,
                                                const int *row_inst) {
    const int lane = threadIdx.x & 63;
    ShortRows R;
    R.nrows = min(64, mp - r0);
    const int r = r0 + min(lane, R.nrows - 1);
    const int a = row_ptr_out[r], b = row_ptr_out[r + 1];
    R.st = a & ~(kRowAlign - 1);
    R.len = (b & ~(kRowAlign - 1)) - R.st - (a & (kRowAlign - 1));
    R.k0 = src_start ? src_start[r] : 0;
    R.inst = row_inst ? row_inst[r] : 0;
    R.S0 = __shfl(R.st, 0, 64);
    R.S1 = __shfl(b & ~(kRowAlign - 1), R.nrows - 1, 64);
    return R;
}

// Lane of the row holding padded entry e (S0 <= e): the largest j < nrows
// with start_j <= e.  Every lane of the wave takes part.
__device__ __forceinline__ int short_row_of(const ShortRows &R, int e) {
    int lo = 0;
    for (int step = 32; step > 0; step >>= 1) {
        const int j = lo + step;
        const int sj = __shfl(R.st, j < R.nrows ? j : R.nrows - 1, 64);
        if (j < R.nrows && sj <= e) lo = j;
    }
    return lo;
}

__global__ __launch_bounds__(256) void layout_rows_short_kernel(int r_begin, int mp, const long long *src_start,
                                                                const int *row_inst, const int *inst_col_off,
                                                                const int *col_in, const double *val_in,
                                                                const int *row_ptr_out, const int *internal,
                                                                const int *vartype, int *col_out, double *val_out,
                                                                int *col_cnt, int *bad) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int r0 = r_begin + 64 * (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); r0 < mp; r0 += 64 * waves) {
        const ShortRows R = short_rows(r0, mp, row_ptr_out, src_start, row_inst);
        for (int base = R.S0; base < R.S1; base += 64) {
            const int e = base + lane;
            const int j = short_row_of(R, e);
            const int rs = __shfl(R.st, j, 64), rl = __shfl(R.len, j, 64), inst = __shfl(R.inst, j, 64);
            const long long k0 = __shfl(R.k0, j, 64);
            const int k = e - rs;
            if (e >= R.S1 || k >= rl) continue;
            const int cbase = inst_col_off[inst], lim = inst_col_off[inst + 1] - cbase;
            const int c = col_in[k0 + k];
            if (c < 0 || c >= lim) {
                atomic_or_i(bad, 1);
                continue;
            }
            const int g = cbase + c;
            const int id = internal[g];
            col_out[e] = id | (vartype[g] == kVarContinuous ? 0 : kIntFlag);
            val_out[e] = val_in[k0 + k];
            atomic_add_i(&col_cnt[id], 1);
        }
    }
}

// Column-to-row index entries of the short rows (never split: segment -1).
__global__ __launch_bounds__(256) void layout_csc_short_kernel(int r_begin, int mp, const int *row_ptr_out,
                                                               const int *col_out, const int *row_item, int *fill,
                                                               int *col_rows, int *col_seg, int *col_item) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int r0 = r_begin + 64 * (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); r0 < mp; r0 += 64 * waves) {
        const ShortRows R = short_rows(r0, mp, row_ptr_out, nullptr, nullptr);
        const int item = row_item[r0 + min(lane, R.nrows - 1)];
        for (int base = R.S0; base < R.S1; base += 64) {
            const int e = base + lane;
            const int j = short_row_of(R, e);
            const int rs = __shfl(R.st, j, 64), rl = __shfl(R.len, j, 64), it = __shfl(item, j, 64);
            if (e >= R.S1 || e - rs >= rl) continue;
            const int id = col_out[e] & kIdMask;
            const int slot = atomic_add_i(&fill[id], 1);
            col_rows[slot] = r0 + j;
            col_seg[slot] = -1;
            col_item[slot] = it;
        }
    }
}

}  // namespace gdp
