// gdp_engine.hip - host driver of the MI355X propagation engine.
//
// Reference counterpart: propagateConstraintsGPUAtomic<T>
// (src/propagators/GPU_propagator.cuh:98-268) and the row blocking of
// fill_row_blocks (src/kernels/util_kernels.cu:12-62).  Differences by
// design: the matrix is laid out once per Problem (rows grouped by tree
// width, columns relabelled by bound storage class, integrality folded into
// the column index, a column-to-row index for row marking); rounds are
// round-synchronous steps (sweep, apply of split rows, publish); GPU_LOOP is
// a captured HIP graph of step launches that stop themselves on the device;
// several bound vectors or instances can share the launches (units).
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "gdp_arith.h"
#include "gdp_copy_pool.h"
#include "gdp_device.h"
#include "gdp_devmem.h"
#include "gdp_engine.h"
#include "gdp_host.h"
#include "gdp_order.h"

namespace gdp {

__global__ void sweep_kernel(DevMatrix M, DevState D, int step, int guard, int first_step);
__global__ void apply_split_kernel(DevMatrix M, DevState D, int step, int guard, int first_step);
__global__ void publish_kernel(DevMatrix M, DevState D, int step, int guard, int first_step, int n_units);
template <int kThreads, bool kLds>
__global__ void mega_kernel(DevMatrix M, DevState D, int first_step, int max_steps, int n_units);
__global__ void unit_reset_kernel(DevState D, int n_list, const int4 *params, int start_step, int clear_status);
__global__ void load_kernel(DevMatrix M, DevState D, int n_list, const int *list, const double *const *lbp,
                            const double *const *ubp);
__global__ void hint_kernel(DevMatrix M, DevState D, int n_list, const int *list, const int *hint_ptr,
                            const int *hints, const int *inv_perm, int start_step);
__global__ void unpack_kernel(DevMatrix M, DevState D, int n_list, const int *list, double *const *lbp,
                              double *const *ubp);
__global__ void delta_kernel(int n_list, double *const *lbp, double *const *ubp, const int *ptr, const int *cols,
                             const double *lv, const double *uv);
__global__ void copy_pairs_kernel(long long n2, int count, double *const *ptrs);
__global__ void diff_kernel(int n, int n_list, double *const *lbp, double *const *ubp, const double *const *blp,
                            const double *const *bup, int cap, int *cnt, int *cols, double *vl, double *vu);
__global__ void compact_changes_kernel(int nd, int cap, const int *cnt, const int *cols, const double *vl,
                                       const double *vu, int *off, int *kcols, double *kl, double *ku);
__global__ void layout_rows_kernel(int mp, const long long *src_start, const int *row_inst, const int *inst_col_off,
                                   const int *col_in, const double *val_in, const int *row_ptr_out,
                                   const int *internal, const int *vartype, int *col_out, double *val_out,
                                   int *col_cnt, int *bad);
__global__ void layout_csc_kernel(int mp, const int *row_ptr_out, const int *col_out, const int *row_part,
                                  const int *row_item, int *fill, int *col_rows, int *col_seg, int *col_item);
__global__ void layout_rows_short_kernel(int r_begin, int mp, const long long *src_start, const int *row_inst,
                                         const int *inst_col_off, const int *col_in, const double *val_in,
                                         const int *row_ptr_out, const int *internal, const int *vartype,
                                         int *col_out, double *val_out, int *col_cnt, int *bad);
__global__ void layout_csc_short_kernel(int r_begin, int mp, const int *row_ptr_out, const int *col_out,
                                        const int *row_item, int *fill, int *col_rows, int *col_seg, int *col_item);
__global__ void measure_kernel(int n, const double *lb, const double *ub, const double *wl, const double *wu,
                               const double *ll, const double *lu, double *pl, double *pu, double *fin_part,
                               int *inf_part);
__global__ void measure_total_kernel(int parts, const double *fin_part, const int *inf_part, double *fin, int *inf);

namespace {

thread_local std::string g_last_error;

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T *dev_alloc(size_t count, std::vector<void *> &owned) {
    void *p = nullptr;
    if (count == 0) count = 1;
    check(cached_malloc(&p, count * sizeof(T)), "hipMalloc");
    owned.push_back(p);
    return static_cast<T *>(p);
}

// Copies of 1 MiB and more from pageable host memory (the caller's arrays): the runtime
// would stage them through its own pinned buffers, copying on one thread.
// Here several threads copy each 32 MiB chunk into one of two pinned staging
// buffers while the DMA engine drains the other.  One copy at a time uses
// them, and without them the copy is left to the runtime.  Returns when
// every copy out of them has completed.
//
// Lifetime: the buffers (64 MiB of pinned memory) and the copy threads exist
// while at least one problem does.  The first bulk copy creates them, and
// the destruction of the last problem joins the threads and frees the
// buffers, so nothing of them is left for the runtime's (or a profiler's)
// teardown at process exit.
constexpr size_t kStageChunk = 32u << 20;
constexpr size_t kBulkMin = 1u << 20;

class Staging {
  public:
    std::mutex use;  // held for a whole bulk copy
    char *buf[2] = {nullptr, nullptr};

    // A problem starts / ends using the staging.
    void acquire() {
        std::lock_guard<std::mutex> lk(life_);
        users_++;
    }
    void release() {
        std::lock_guard<std::mutex> use_lk(use);
        std::lock_guard<std::mutex> lk(life_);
        if (--users_ > 0) return;
        pool_.stop();
        for (char *&b : buf) {
            if (b) (void)hipHostFree(b);
            b = nullptr;
        }
    }
    // With `use` held: whether both buffers exist (allocated here).
    bool ready() {
        for (char *&b : buf)
            if (!b && hipHostMalloc(reinterpret_cast<void **>(&b), kStageChunk, hipHostMallocDefault) != hipSuccess) {
                b = nullptr;
                (void)hipGetLastError();
                return false;
            }
        return true;
    }
    // With `use` held: copies `bytes` from src to dst with the copy threads.
    void copy(void *dst, const void *src, size_t bytes) { pool_.copy(dst, src, bytes, copy_threads()); }

    // Threads of a bulk copy: GDP_COPY_THREADS if set (1..64), else 8, at
    // most the CPUs this process may run on.
    static int copy_threads() {
        static const int k = [] {
            int usable = static_cast<int>(std::thread::hardware_concurrency());
            cpu_set_t set;
            if (sched_getaffinity(0, sizeof(set), &set) == 0) usable = CPU_COUNT(&set);
            int want = 8;
            if (const char *e = std::getenv("GDP_COPY_THREADS")) want = std::atoi(e);
            want = std::max(1, std::min(64, want));
            return std::max(1, std::min(want, std::max(1, usable)));
        }();
        return k;
    }

  private:
    std::mutex life_;
    int users_ = 0;
    CopyPool pool_;
};

Staging &staging() {
    static Staging st;
    return st;
}

void upload_bulk(void *dst, const void *src, size_t bytes, hipStream_t s) {
    Staging &st = staging();
    std::lock_guard<std::mutex> lock(st.use);
    if (!st.ready()) {  // no pinned memory to spare: the runtime's own staged copy
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "H2D");
        return;
    }
    char *const *buf = st.buf;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    auto release = [&]() {
        for (int k = 0; k < 2; k++) {
            if (used[k]) (void)hipEventSynchronize(ev[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        }
    };
    try {
        for (hipEvent_t &e : ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        int k = 0;
        for (size_t off = 0; off < bytes; off += kStageChunk, k ^= 1) {
            const size_t len = std::min(kStageChunk, bytes - off);
            if (used[k]) check(hipEventSynchronize(ev[k]), "staging wait");  // its previous chunk has left
            st.copy(buf[k], static_cast<const char *>(src) + off, len);
            check(hipMemcpyAsync(static_cast<char *>(dst) + off, buf[k], len, hipMemcpyHostToDevice, s), "H2D");
            check(hipEventRecord(ev[k], s), "event");
            used[k] = true;
        }
    } catch (...) {
        release();
        throw;
    }
    release();
}

// The reverse of upload_bulk, into pageable host memory: chunk i's DMA into
// one staging buffer runs while chunk i - 1 is copied out of the other.
// Synchronous.
void download_bulk(void *dst, const void *src, size_t bytes, hipStream_t s) {
    Staging &st = staging();
    std::lock_guard<std::mutex> lock(st.use);
    if (!st.ready()) {
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "D2H");
        return;
    }
    char *const *buf = st.buf;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    auto release = [&]() {
        for (int k = 0; k < 2; k++) {
            if (used[k]) (void)hipEventSynchronize(ev[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        }
    };
    try {
        for (hipEvent_t &e : ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        int prev = -1;
        size_t prev_off = 0, prev_len = 0;
        int k = 0;
        // a buffer is filled again only after its previous chunk was copied out
        for (size_t off = 0; off < bytes; off += kStageChunk, k ^= 1) {
            const size_t len = std::min(kStageChunk, bytes - off);
            check(hipMemcpyAsync(buf[k], static_cast<const char *>(src) + off, len, hipMemcpyDeviceToHost, s), "D2H");
            check(hipEventRecord(ev[k], s), "event");
            used[k] = true;
            if (prev >= 0) {
                check(hipEventSynchronize(ev[prev]), "staging wait");
                st.copy(static_cast<char *>(dst) + prev_off, buf[prev], prev_len);
            }
            prev = k;
            prev_off = off;
            prev_len = len;
        }
        if (prev >= 0) {
            check(hipEventSynchronize(ev[prev]), "staging wait");
            st.copy(static_cast<char *>(dst) + prev_off, buf[prev], prev_len);
        }
    } catch (...) {
        release();
        throw;
    }
    release();
}

// Device to pageable host memory; synchronous.
template <typename T>
void download(T *dst, const T *src, size_t count, hipStream_t s) {
    if (!count) return;
    if (count * sizeof(T) >= kBulkMin) {
        download_bulk(dst, src, count * sizeof(T), s);
    } else {
        check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "D2H");
    }
}

template <typename T>
void upload(T *dst, const T *src, size_t count, hipStream_t s) {
    if (!count) return;
    if (count * sizeof(T) >= kBulkMin)
        upload_bulk(dst, src, count * sizeof(T), s);
    else
        check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
}

// Phase times of a problem's creation, printed to stderr when the
// environment variable GDP_TRACE_CREATE is set (a diagnostic of where the
// creation time goes: host ordering, allocations, copies, device layout).
class CreateTrace {
  public:
    CreateTrace() : on_(std::getenv("GDP_TRACE_CREATE") != nullptr), t0_(Clock::now()), last_(t0_) {}
    ~CreateTrace() {
        if (on_) std::fprintf(stderr, "create:%s total=%.2f ms\n", line_.c_str(), ms(t0_, Clock::now()));
    }
    void mark(const char *phase) {
        if (!on_) return;
        const Clock::time_point t = Clock::now();
        char buf[96];
        std::snprintf(buf, sizeof(buf), " %s=%.2f", phase, ms(last_, t));
        line_ += buf;
        last_ = t;
    }

  private:
    using Clock = std::chrono::steady_clock;
    static double ms(Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    }
    bool on_;
    Clock::time_point t0_, last_;
    std::string line_;
};

// Storage class of a column from its bound envelope: 0 = 2 x int16,
// 1 = 2 x int32, 2 = 2 x double.
int column_class(int vartype, const double *env_lb, const double *env_ub, int j) {
    if (vartype == kVarContinuous || !env_lb || !env_ub) return 2;
    const double l = env_lb[j], u = env_ub[j];
    // both within the int32 range (false for NaN and infinities), then
    // integral: exact through a 64-bit integer round trip in that range,
    // without the library calls of floor()
    if (!(l >= -2147483647.0 && l <= 2147483647.0 && u >= -2147483647.0 && u <= 2147483647.0)) return 2;
    if (static_cast<double>(static_cast<long long>(l)) != l || static_cast<double>(static_cast<long long>(u)) != u)
        return 2;
    return l >= -32767.0 && l <= 32767.0 && u >= -32767.0 && u <= 32767.0 ? 0 : 1;
}

// Input contract of the row structure (the column indices are checked on
// the device while the layout is built).
void validate(const InstanceView &v) {
    if (v.m <= 0 || v.n <= 0 || v.nnz <= 0) throw std::runtime_error("empty problem");
    if (!v.col || !v.row_ptr || !v.val || !v.lhs || !v.rhs || !v.vartype) throw std::runtime_error("null input array");
    if (v.row_ptr[0] != 0 || v.row_ptr[v.m] != v.nnz) throw std::runtime_error("row_ptrs do not span nnz");
    for (int i = 0; i < v.m; i++)
        if (v.row_ptr[i + 1] < v.row_ptr[i]) throw std::runtime_error("row_ptrs not monotone");
}

// Outcome of one unit from its state words (the reference's round-count
// convention: the final no-change round counts).
RunResult outcome(const UnitState &s) {
    RunResult r;
    if (s.envelope) throw std::runtime_error("bounds outside the problem's bound envelope");
    if (s.crossed) {
        r.infeasible = 1;
        return r;
    }
    if (s.infeasible) {
        r.rounds = s.infeasible - s.start + 1;
        r.infeasible = 1;
        return r;
    }
    const int last_change = std::max(s.changed[0], s.changed[1]);
    const int changed_rounds = last_change ? last_change - s.start + 1 : 0;
    if (changed_rounds >= s.max_rounds) {
        r.rounds = s.max_rounds;
        r.capped = 1;
    } else {
        r.rounds = changed_rounds + 1;
    }
    return r;
}

// Whether a unit works in the step after `step` (host mirror of unit_step).
bool active_after(const UnitState &s, int step) {
    if (s.start <= 0 || s.crossed || s.envelope || s.infeasible) return false;
    if (step + 1 - s.start >= s.max_rounds) return false;
    return s.changed[step & 1] == step;
}

int grid_for(long long waves) { return static_cast<int>(std::max(1LL, (waves + 3) / 4)); }

}  // namespace

std::string last_error() { return g_last_error; }
void set_last_error(const std::string &msg) { g_last_error = msg; }

// Launch tables of a run state (filled by the host, read by the kernels).
struct UnitTables {
    int *list = nullptr;      // listed units of a launch
    int4 *params = nullptr;   // unit_reset parameters
    double **lbp = nullptr;   // bound-vector pointer tables
    double **ubp = nullptr;
    int *hint_ptr = nullptr;  // hint starts
};

struct Problem {
    bool replicas = true;
    int m = 0, n = 0, nnz = 0;       // caller's totals (all instances)
    int n_inst = 1;
    Tuning tuning;
    ProblemStats stats;
    DevMatrix M{};
    DevState D{};
    std::vector<void *> owned;       // matrix layout
    void *state_buf = nullptr;       // run state owned by the problem (carve_state), or null
    int capacity = 0;                // units with state arrays (replicas)
    int n_units = 0;                 // entries of D.units
    std::vector<int> inst_ncols;     // blocks: columns per instance
    std::vector<int> mark_limit;     // per instance (blocks) or for every unit (replicas)
    int *d_list = nullptr;           // listed units of a launch
    int4 *d_params = nullptr;        // unit_reset parameters
    double **d_lbp = nullptr;        // bound-vector pointer tables
    double **d_ubp = nullptr;
    int *d_inv_perm = nullptr;
    int *d_hint_ptr = nullptr, *d_hints = nullptr;
    size_t hints_cap = 0;
    std::vector<void *> hint_owned;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int device = 0;
    int cu_count = 0;
    int mega_grid = 0;
    int n_lds = 0;                    // binary columns a one-unit persistent launch reads from LDS
    int mega_threads = 0;             // workgroup size of the persistent kernel
    int sweep_grid = 0;               // workgroups of the step sweep kernel (all resident)
    const void *mega_fn = nullptr;
    const void *mega_fn_lds = nullptr;  // the same reading binary columns from LDS (one unit)
    // captured GPU_LOOP graphs, one per chunk of steps, for one launch shape
    std::vector<hipGraphExec_t> graphs;
    std::vector<int> graph_last;
    long long graph_key = -1;
    std::mutex run_mutex;
    int servers = 0;                  // servers attached (they own the unit slots)
    bool graphs_enabled = true;       // GPU_LOOP may capture graphs of the current run state
    bool staging_held = false;

    void drop_graphs() {
        for (auto g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        graphs.clear();
        graph_last.clear();
        graph_key = -1;
    }

    ~Problem() {
        drop_graphs();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        if (state_buf) (void)cached_free(state_buf);
        for (void *p : hint_owned) (void)cached_free(p);
        for (void *p : owned) (void)cached_free(p);
        if (staging_held) staging().release();
    }
};

namespace {

// The run state of `cap` bound vectors (replicas) or of the batch (blocks),
// in one buffer at `base` with every array on a 256-byte boundary: the unit
// state the kernels read and write (snapshot, accumulators, bitmap, row and
// segment marks, split-row partials, apply queue, per-unit words and
// counters), the problem-wide status ring, global words and barrier, and
// the launch tables the host fills (listed units, reset parameters,
// bound-vector pointers, hint starts).  With base null only the size is
// computed.  Returns the bytes needed.
size_t carve_state(const DevMatrix &M, bool replicas, int n_inst, int cap, char *base, DevState &D,
                   UnitTables &T) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return p;
    };
    const size_t k = replicas ? static_cast<size_t>(cap) : 1;
    const size_t units = replicas ? static_cast<size_t>(cap) : static_cast<size_t>(n_inst);
    const size_t nf = static_cast<size_t>(M.n - M.n_int);
    char *snap = take(k * static_cast<size_t>(M.snap_stride));
    D.S16 = reinterpret_cast<unsigned *>(snap);
    D.S32 = reinterpret_cast<int2 *>(snap ? snap + M.snap_off32 : nullptr);
    D.SF = reinterpret_cast<double2 *>(snap ? snap + M.snap_offF : nullptr);
    D.AI = reinterpret_cast<int2 *>(take(k * M.n_int * sizeof(int2)));
    D.AX = reinterpret_cast<double2 *>(take(k * M.n_int * sizeof(double2)));
    D.AF = reinterpret_cast<double2 *>(take(k * nf * sizeof(double2)));
    D.bm = reinterpret_cast<unsigned *>(take(k * M.nwords * sizeof(unsigned)));
    D.SB = reinterpret_cast<unsigned *>(take(k * M.sb_words * sizeof(unsigned)));
    D.marks = reinterpret_cast<int *>(take(k * M.m * sizeof(int)));
    D.item_mark = reinterpret_cast<int *>(take(k * M.n_items * sizeof(int)));
    D.seg_parts = reinterpret_cast<SegPart *>(take(k * M.n_seg_parts * sizeof(SegPart)));
    D.seg_marks = reinterpret_cast<int *>(take(k * M.n_seg_parts * sizeof(int)));
    D.seg_state = reinterpret_cast<SegState *>(take(k * M.n_seg_rows * sizeof(SegState)));
    D.seg_count = reinterpret_cast<unsigned *>(take(k * M.n_seg_rows * sizeof(unsigned)));
    D.apply_tasks = reinterpret_cast<int *>(take(k * M.n_seg_parts * sizeof(int)));
    D.units = reinterpret_cast<UnitState *>(take(units * sizeof(UnitState)));
    D.moved = reinterpret_cast<int *>(take(units * kMoveSlots * kMoveLanes * kMoveStride * sizeof(int)));
    D.status = reinterpret_cast<int *>(take(kStatusRows * kStatusWords * sizeof(int)));
    D.global = reinterpret_cast<int *>(take(kGlobalWords * sizeof(int)));
    D.barrier = reinterpret_cast<unsigned *>(take(kBarrierWords * sizeof(unsigned)));
    T.list = reinterpret_cast<int *>(take(units * sizeof(int)));
    T.params = reinterpret_cast<int4 *>(take(units * sizeof(int4)));
    T.lbp = reinterpret_cast<double **>(take(units * sizeof(double *)));
    T.ubp = reinterpret_cast<double **>(take(units * sizeof(double *)));
    T.hint_ptr = reinterpret_cast<int *>(take((units + 1) * sizeof(int)));
    D.active = T.list;
    return off;
}

// Points the problem's run state at `base` (cap units, see carve_state).
void bind_state(Problem *P, char *base, int cap) {
    UnitTables T;
    carve_state(P->M, P->replicas, P->n_inst, cap, base, P->D, T);
    P->d_list = T.list;
    P->d_params = T.params;
    P->d_lbp = T.lbp;
    P->d_ubp = T.ubp;
    P->d_hint_ptr = T.hint_ptr;
    P->capacity = cap;
    P->n_units = P->replicas ? cap : P->n_inst;
}

// Device state of `cap` units (replicas) or of the batch (blocks), owned by
// the problem.
void alloc_units(Problem *P, int cap) {
    check(hipStreamSynchronize(P->stream), "sync");
    P->drop_graphs();
    if (P->state_buf) (void)cached_free(P->state_buf);
    P->state_buf = nullptr;
    P->capacity = 0;
    P->n_units = 0;
    DevState D{};
    UnitTables T;
    const size_t bytes = carve_state(P->M, P->replicas, P->n_inst, cap, nullptr, D, T);
    check(cached_malloc(&P->state_buf, bytes), "hipMalloc run state");
    bind_state(P, static_cast<char *>(P->state_buf), cap);
    const size_t k = P->replicas ? static_cast<size_t>(cap) : 1;
    const DevMatrix &M = P->M;
    check(hipMemsetAsync(P->D.units, 0, sizeof(UnitState) * P->n_units, P->stream), "memset units");
    if (M.n_seg_rows > 0)
        check(hipMemsetAsync(P->D.seg_count, 0, sizeof(unsigned) * k * M.n_seg_rows, P->stream), "memset");
    check(hipMemsetAsync(P->D.bm, 0, sizeof(unsigned) * k * M.nwords, P->stream), "memset");
    check(hipMemsetAsync(P->D.barrier, 0, sizeof(unsigned) * kBarrierWords, P->stream), "memset");
    check(hipMemsetAsync(P->D.global, 0, sizeof(int) * kGlobalWords, P->stream), "memset");
    check(hipStreamSynchronize(P->stream), "alloc sync");
    P->stats.n_units = P->n_units;
}

// The run state of unit 0 exists (allocated on first use: a problem that is
// only run through the scratch entry point never allocates it).
void ensure_units(Problem *P) {
    if (P->capacity < 1) alloc_units(P, 1);
}

// The instances' CSR arrays (column indices, coefficients) and variable
// types, concatenated on the device.  The copy runs on its own stream from
// its own host thread, started once every other bulk copy of the creation
// has been issued (bulk copies take the staging buffers one at a time), so
// that it overlaps the ordering of the rows on the problem's stream and the
// host's building of the work items.
class CsrUpload {
  public:
    int *col = nullptr;
    double *val = nullptr;
    int *vartype = nullptr;

    CsrUpload(const std::vector<InstanceView> &inst, const std::vector<int> &col_off) : inst_(inst), off_(col_off) {
        long long nnz = 0;
        for (const auto &v : inst) nnz += v.nnz;
        col = dev_alloc<int>(nnz, mem_);
        val = dev_alloc<double>(nnz, mem_);
        vartype = dev_alloc<int>(col_off.back(), mem_);
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    CsrUpload(const CsrUpload &) = delete;
    CsrUpload &operator=(const CsrUpload &) = delete;
    ~CsrUpload() {
        if (th_.joinable()) th_.join();
        if (stream_) (void)hipStreamDestroy(stream_);
        for (void *p : mem_) (void)cached_free(p);
    }

    void start(int device) {
        th_ = std::thread([this, device] {
            try {
                check(hipSetDevice(device), "hipSetDevice");
                long long at = 0;
                for (size_t i = 0; i < inst_.size(); i++) {
                    upload(col + at, inst_[i].col, inst_[i].nnz, stream_);
                    upload(val + at, inst_[i].val, inst_[i].nnz, stream_);
                    upload(vartype + off_[i], inst_[i].vartype, inst_[i].n, stream_);
                    at += inst_[i].nnz;
                }
                check(hipStreamSynchronize(stream_), "CSR upload");
            } catch (...) {
                err_ = std::current_exception();
            }
        });
    }
    // Waits for the copy; rethrows its error.
    void finish() {
        if (th_.joinable()) th_.join();
        if (err_) std::rethrow_exception(err_);
    }

  private:
    const std::vector<InstanceView> &inst_;
    const std::vector<int> &off_;
    std::vector<void *> mem_;
    hipStream_t stream_ = nullptr;
    std::thread th_;
    std::exception_ptr err_;
};

// The passes over the nonzeros, on the device: from the CSR arrays as
// uploaded (concatenated), the entries are moved into the padded
// layout (layout row r from source position src_start[r] of instance
// row_inst[r], at the starts of d_rp), the columns' entries are counted, the
// counts scanned into the column starts and the column-to-row index filled.
// Column indices outside an instance's range are reported here.  src_start,
// row_inst (null for one instance), row_part and row_item are device arrays
// of the mp layout rows.  Synchronous.
void device_layout_build(CsrUpload &csr, const std::vector<int> &col_off, int mp, int n_wide, int n_ids, int n_padded,
                         const long long *src_start, const int *row_inst, const int *row_part, const int *row_item,
                         const int *d_rp, const int *d_internal, int *d_col, double *d_val, int *d_cptr,
                         int *d_crows, int *d_cseg, int *d_citem, hipStream_t s, CreateTrace &trace) {
    const int NI = static_cast<int>(col_off.size()) - 1;
    std::vector<void *> tmp;
    try {
        int *t_coff = dev_alloc<int>(NI + 1, tmp);
        int *t_cnt = dev_alloc<int>(n_ids + 2, tmp);  // counts (n_ids + 1: the scan's total), the out-of-range flag
        upload(t_coff, col_off.data(), NI + 1, s);
        check(hipMemsetAsync(d_col, 0, sizeof(int) * n_padded, s), "memset layout");
        check(hipMemsetAsync(d_val, 0, sizeof(double) * n_padded, s), "memset layout");
        check(hipMemsetAsync(t_cnt, 0, sizeof(int) * (n_ids + 2), s), "memset counts");
        csr.finish();
        trace.mark("wait_csr");
        const int *t_col = csr.col;
        const double *t_val = csr.val;
        const int *t_vt = csr.vartype;
        // rows [0, n_wide) one wave per row, the short rows after them 64 per wave
        const int grid = std::max(1, std::min(2048, (n_wide + 3) / 4));
        const int grid_short = std::max(1, std::min(2048, (mp - n_wide + 255) / 256));
        if (n_wide > 0)
            hipLaunchKernelGGL(layout_rows_kernel, dim3(grid), dim3(256), 0, s, n_wide, src_start, row_inst, t_coff,
                               t_col, t_val, d_rp, d_internal, t_vt, d_col, d_val, t_cnt, t_cnt + n_ids + 1);
        if (mp > n_wide)
            hipLaunchKernelGGL(layout_rows_short_kernel, dim3(grid_short), dim3(256), 0, s, n_wide, mp, src_start,
                               row_inst, t_coff, t_col, t_val, d_rp, d_internal, t_vt, d_col, d_val, t_cnt,
                               t_cnt + n_ids + 1);
        check(hipGetLastError(), "layout launch");
        int bad = 0;
        check(hipMemcpyAsync(&bad, t_cnt + n_ids + 1, sizeof(int), hipMemcpyDeviceToHost, s), "D2H flag");
        check(hipStreamSynchronize(s), "layout");
        trace.mark("layout_rows");
        if (bad) throw std::runtime_error("column index out of range");
        order_exclusive_sum(t_cnt, d_cptr, n_ids + 1, s);
        check(hipMemcpyAsync(t_cnt, d_cptr, sizeof(int) * n_ids, hipMemcpyDeviceToDevice, s), "fill positions");
        if (n_wide > 0)
            hipLaunchKernelGGL(layout_csc_kernel, dim3(grid), dim3(256), 0, s, n_wide, d_rp, d_col, row_part, row_item,
                               t_cnt, d_crows, d_cseg, d_citem);
        if (mp > n_wide)
            hipLaunchKernelGGL(layout_csc_short_kernel, dim3(grid_short), dim3(256), 0, s, n_wide, mp, d_rp, d_col,
                               row_item, t_cnt, d_crows, d_cseg, d_citem);
        check(hipGetLastError(), "layout launch");
        check(hipStreamSynchronize(s), "layout");
        trace.mark("layout_csc");
    } catch (...) {
        for (void *p : tmp) (void)cached_free(p);
        throw;
    }
    for (void *p : tmp) check(cached_free(p), "hipFree");
    trace.mark("free_temp");
}

// Layout rows as ordered on the device (gdp_order.hip): host copies of the
// totals per bin, of the first layout row and row count of each (bin key,
// instance) run, and of the long rows' lengths; device arrays of the source
// position and instance of each layout row.
struct RowOrder {
    int mp = 0;            // nonempty rows
    int n_padded = 0;      // entries of the padded layout
    std::vector<int> start, count;  // [key * NI + i]; key 0: long rows, key k: tree-width bin 10 - k
    std::vector<int> long_len;      // lengths of layout rows [0, bin_rows[10])
    long long bin_rows[11] = {0}, bin_nnz[11] = {0};
    long long *src = nullptr;  // device, per layout row
    int *inst = nullptr;       // device, per layout row (null for one instance)
};

// Orders the rows on the device.  Allocates the row starts d_rp (mp + 1
// entries, padding count in the low bits) and sides d_sides (mp) into
// `owned`, the per-row temporaries into `tmp`.  Calls after_copies once its
// bulk copies are issued.
RowOrder device_row_order(const std::vector<InstanceView> &inst, const std::vector<int> &row_off, long long nnz_total,
                          int **d_rp, double2 **d_sides, std::vector<void *> &owned, std::vector<void *> &tmp,
                          hipStream_t s, const std::function<void()> &after_copies) {
    const int NI = static_cast<int>(inst.size());
    const int M = row_off[NI];
    RowOrder R;
    std::vector<long long> nnz_off(NI + 1, 0);
    for (int i = 0; i < NI; i++) nnz_off[i + 1] = nnz_off[i] + inst[i].nnz;
    int *t_rp = dev_alloc<int>(static_cast<size_t>(M) + NI, tmp);
    double *t_lhs = dev_alloc<double>(M, tmp);
    double *t_rhs = dev_alloc<double>(M, tmp);
    int *t_roff = dev_alloc<int>(NI + 1, tmp);
    long long *t_noff = dev_alloc<long long>(NI + 1, tmp);
    unsigned char *t_key = dev_alloc<unsigned char>(2 * static_cast<size_t>(M), tmp);  // keys, sorted keys
    int *t_row = dev_alloc<int>(2 * static_cast<size_t>(M), tmp);                     // rows, sorted rows
    unsigned long long *t_tot = dev_alloc<unsigned long long>(48, tmp);
    for (int i = 0; i < NI; i++) {
        upload(t_rp + row_off[i] + i, inst[i].row_ptr, static_cast<size_t>(inst[i].m) + 1, s);
        upload(t_lhs + row_off[i], inst[i].lhs, inst[i].m, s);
        upload(t_rhs + row_off[i], inst[i].rhs, inst[i].m, s);
    }
    upload(t_roff, row_off.data(), NI + 1, s);
    upload(t_noff, nnz_off.data(), NI + 1, s);
    after_copies();
    check(hipMemsetAsync(t_tot, 0, sizeof(unsigned long long) * 48, s), "memset totals");
    order_row_keys(M, NI, t_roff, t_rp, t_key, t_row, t_tot, s);
    order_sort_rows(t_key, t_key + M, t_row, t_row + M, M, s);
    unsigned long long tot[48];
    check(hipMemcpyAsync(tot, t_tot, sizeof(tot), hipMemcpyDeviceToHost, s), "D2H totals");
    check(hipStreamSynchronize(s), "row keys");
    long long mp = 0, pad = 0;
    for (int key = 0; key <= 10; key++) {
        R.bin_rows[10 - key] = static_cast<long long>(tot[key]);
        R.bin_nnz[10 - key] = static_cast<long long>(tot[16 + key]);
        mp += static_cast<long long>(tot[key]);
        pad += static_cast<long long>(tot[32 + key]);
    }
    if (nnz_total + pad > 0x7fffff00LL) throw std::runtime_error("too many nonzeros");
    R.mp = static_cast<int>(mp);
    const int n_long = static_cast<int>(R.bin_rows[10]);
    *d_rp = dev_alloc<int>(static_cast<size_t>(R.mp) + 1, owned);
    *d_sides = dev_alloc<double2>(R.mp, owned);
    int *t_len = dev_alloc<int>(R.mp, tmp);
    int *t_plen = dev_alloc<int>(static_cast<size_t>(R.mp) + 1, tmp);
    int *t_start = dev_alloc<int>(11 * static_cast<size_t>(NI), tmp);
    R.src = dev_alloc<long long>(R.mp, tmp);
    R.inst = NI > 1 ? dev_alloc<int>(R.mp, tmp) : nullptr;
    check(hipMemsetAsync(t_start, 0xff, sizeof(int) * 11 * NI, s), "memset starts");
    check(hipMemsetAsync(t_plen + R.mp, 0, sizeof(int), s), "memset total");
    order_place_rows(R.mp, NI, t_roff, t_noff, t_rp, t_lhs, t_rhs, t_key + M, t_row + M, t_len, t_plen, R.src,
                     R.inst, *d_sides, t_start, s);
    order_exclusive_sum(t_plen, *d_rp, R.mp + 1, s);
    order_pad_rows(R.mp, t_len, t_plen, *d_rp, s);
    R.start.resize(11 * static_cast<size_t>(NI));
    R.long_len.resize(n_long);
    check(hipMemcpyAsync(R.start.data(), t_start, sizeof(int) * 11 * NI, hipMemcpyDeviceToHost, s), "D2H starts");
    if (n_long)
        check(hipMemcpyAsync(R.long_len.data(), t_len, sizeof(int) * n_long, hipMemcpyDeviceToHost, s), "D2H lens");
    check(hipMemcpyAsync(&R.n_padded, *d_rp + R.mp, sizeof(int), hipMemcpyDeviceToHost, s), "D2H total");
    check(hipStreamSynchronize(s), "row order");
    // runs are contiguous and ordered by (key, instance): a run ends where
    // the next present one starts
    R.count.assign(R.start.size(), 0);
    int prev = -1;
    for (size_t k = 0; k < R.start.size(); k++) {
        if (R.start[k] < 0) continue;
        if (prev >= 0) R.count[prev] = R.start[k] - R.start[prev];
        prev = static_cast<int>(k);
    }
    if (prev >= 0) R.count[prev] = R.mp - R.start[prev];
    return R;
}

Problem *build(const std::vector<InstanceView> &inst, bool replicas, const Tuning *tuning) {
    if (inst.empty()) throw std::runtime_error("no instance");
    CreateTrace trace;
    // The host orders rows and columns; the nonzeros are laid out on the
    // device, which also checks the column indices.
    for (const auto &v : inst) validate(v);
    trace.mark("validate");
    auto *P = new Problem();
    staging().acquire();
    P->staging_held = true;
    try {
        P->replicas = replicas;
        P->n_inst = static_cast<int>(inst.size());
        if (tuning) P->tuning = *tuning;
        const Tuning &T = P->tuning;
        const int NI = P->n_inst;
        std::vector<int> col_off(NI + 1, 0), row_off(NI + 1, 0);
        long long nnz_total = 0;
        for (int i = 0; i < NI; i++) {
            col_off[i + 1] = col_off[i] + inst[i].n;
            row_off[i + 1] = row_off[i] + inst[i].m;
            nnz_total += inst[i].nnz;
            P->inst_ncols.push_back(inst[i].n);
        }
        if (nnz_total > 0x7fffff00LL || col_off[NI] > kIdMask) throw std::runtime_error("problem too large");
        // row marking pays while the rows it marks (about accepted candidates
        // x mean column length) stay below a quarter of the rows
        for (int i = 0; i < NI; i++) {
            const long long lim = T.mark_limit > 0 ? T.mark_limit
                                                    : static_cast<long long>(inst[i].m) * inst[i].n /
                                                          (4LL * std::max(1, inst[i].nnz));
            P->mark_limit.push_back(static_cast<int>(std::max(64LL, std::min(lim, 1LL << 30))));
        }
        if (replicas) P->mark_limit.resize(1);
        P->m = row_off[NI];
        P->n = col_off[NI];
        P->nnz = static_cast<int>(nnz_total);

        // ---- columns: internal ids by storage class, instance-major within a
        // class; blocks mode pads each (class, instance) range to 32 ids ----
        std::vector<unsigned char> cls(P->n);
        std::vector<int> cnt_ci(3 * static_cast<size_t>(NI), 0);
        for (int i = 0; i < NI; i++)
            for (int j = 0; j < inst[i].n; j++) {
                const int c =
                    T.compact_bounds ? column_class(inst[i].vartype[j], inst[i].env_lb, inst[i].env_ub, j) : 2;
                cls[col_off[i] + j] = static_cast<unsigned char>(c);
                cnt_ci[c * NI + i]++;
            }
        std::vector<int> start_ci(3 * static_cast<size_t>(NI));
        int class_start[3];
        long long n_ids_l = 0;
        for (int c = 0; c < 3; c++) {
            class_start[c] = static_cast<int>(n_ids_l);
            for (int i = 0; i < NI; i++) {
                start_ci[c * NI + i] = static_cast<int>(n_ids_l);
                n_ids_l += cnt_ci[c * NI + i];
                if (!replicas) n_ids_l = (n_ids_l + 31) / 32 * 32;
                if (n_ids_l > kIdMask) throw std::runtime_error("too many columns");
            }
        }
        const int n_ids = static_cast<int>(n_ids_l);
        // replicas mode: the binary columns (int16 class, envelope in [0, 1])
        // take the first ids of their class
        std::vector<unsigned char> binary;
        int n_bin = 0;
        if (replicas && NI == 1) {
            binary.assign(P->n, 0);
            for (int j = 0; j < inst[0].n; j++)
                if (cls[j] == 0 && inst[0].env_lb[j] >= 0.0 && inst[0].env_ub[j] <= 1.0) {
                    binary[j] = 1;
                    n_bin++;
                }
        }
        std::vector<int> perm(n_ids, -1), internal(P->n), word_unit;
        {
            std::vector<int> pos = start_ci;
            int bin_pos = start_ci[0];
            if (n_bin) pos[0] += n_bin;
            for (int i = 0; i < NI; i++)
                for (int j = 0; j < inst[i].n; j++) {
                    const int g = col_off[i] + j;
                    const int id = n_bin && binary[g] ? bin_pos++ : pos[cls[g] * NI + i]++;
                    internal[g] = id;
                    perm[id] = g;
                }
        }
        const int nwords = (n_ids + 31) / 32;
        if (!replicas) {
            // every word lies in one (class, instance) range
            word_unit.assign(nwords, 0);
            for (int c = 0; c < 3; c++)
                for (int i = 0; i < NI; i++) {
                    const int a0 = start_ci[c * NI + i], cnt = cnt_ci[c * NI + i];
                    for (int w = a0 / 32; w < (a0 + cnt + 31) / 32; w++) word_unit[w] = i;
                }
        }
        P->stats.n16 = class_start[1] - class_start[0];
        P->stats.n32 = class_start[2] - class_start[1];
        P->stats.nf = n_ids - class_start[2];

        trace.mark("columns");
        // ---- device ----
        check(hipGetDevice(&P->device), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, P->device), "hipGetDeviceProperties");
        P->cu_count = prop.multiProcessorCount;
        check(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&P->ev0), "hipEventCreate");
        check(hipEventCreate(&P->ev1), "hipEventCreate");
        hipStream_t s = P->stream;
        auto &o = P->owned;
        std::vector<void *> tmp;  // per-row temporaries of the ordering
        struct FreeTmp {
            std::vector<void *> &v;
            ~FreeTmp() {
                for (void *p : v) (void)cached_free(p);
            }
        } free_tmp{tmp};

        // ---- rows grouped by tree width: long rows first, then wide to
        // narrow; instance-major inside a bin so that items never mix units.
        // Ordered on the device; rows padded to multiples of kRowAlign entries
        // (padding: column 0, coefficient 0), the padding count in the low
        // bits of the row starts ----
        constexpr int kLongBin = 10;
        int *d_perm = dev_alloc<int>(n_ids, o);
        P->d_inv_perm = dev_alloc<int>(P->n, o);
        upload(d_perm, perm.data(), n_ids, s);
        upload(P->d_inv_perm, internal.data(), P->n, s);
        CsrUpload csr(inst, col_off);
        int *d_rp = nullptr;
        double2 *d_sides = nullptr;
        const RowOrder R =
            device_row_order(inst, row_off, nnz_total, &d_rp, &d_sides, o, tmp, s, [&] { csr.start(P->device); });
        const int mp = R.mp;
        const int n_padded = R.n_padded;
        for (int b = 0; b <= kLongBin; b++) {
            P->stats.bin_rows[b] = R.bin_rows[b];
            P->stats.bin_nnz[b] = R.bin_nnz[b];
        }

        trace.mark("rows");
        // ---- sweep items ----
        const int item_nnz = T.item_nnz > 0 ? T.item_nnz : 1024;
        // Order per instance: segments of split rows, whole very long rows,
        // then the tree-width bins from wide to narrow, so that long work
        // starts first; instance after instance, so that the items the grid
        // works on at one time (consecutive ones, dealt out with a stride)
        // gather from the bounds of few instances of a batch.
        std::vector<std::vector<int4>> inst_items(NI);
        std::vector<int4> h_items;
        std::vector<int4> h_seg_rows;
        std::vector<int> h_part_row;
        int n_parts = 0;
        {
            std::vector<std::pair<int, int>> giant;  // (layout row, instance)
            for (int i = 0; i < NI; i++) {
                const int st = R.start[i], cnt = R.count[i];  // key 0: the long rows
                for (int q = 0; q < cnt; q++) {
                    const int row = st + q;
                    const int nseg = (R.long_len[row] + kSegLen - 1) / kSegLen;
                    const int unit_bits = i << kItemUnitShift;
                    if (T.max_split_segs < 0 || nseg <= T.max_split_segs) {
                        const int lr = static_cast<int>(h_seg_rows.size());
                        h_seg_rows.push_back(make_int4(row, n_parts, nseg, i));
                        for (int sg = 0; sg < nseg; sg += kSegsPerItem)
                            inst_items[i].push_back(
                                make_int4(kItemSeg | unit_bits, lr, sg, std::min(kSegsPerItem, nseg - sg)));
                        for (int sg = 0; sg < nseg; sg++) h_part_row.push_back(lr);
                        n_parts += nseg;
                    } else {
                        giant.push_back({row, i});
                    }
                }
            }
            for (const auto &g : giant)
                inst_items[g.second].push_back(make_int4(kItemLong | (g.second << kItemUnitShift), g.first, 1, 0));
            P->stats.n_long_rows = static_cast<int>(R.bin_rows[kLongBin]);
        }
        for (int b = kLongBin - 1; b >= 0; b--) {
            const int key = kLongBin - b;
            const int width = 1 << b;
            int logg = std::min(T.group_log[b], b);
            logg = std::max(logg, b - 2);  // at most 4 entries per lane
            logg = std::min(logg, 6);
            const int logk = b - logg;
            const int per_pass = 64 >> logg;
            int rows_per_item = std::max(1, item_nnz / width);
            rows_per_item = ((rows_per_item + per_pass - 1) / per_pass) * per_pass;
            for (int i = 0; i < NI; i++) {
                const int st = R.start[key * NI + i], cnt = R.count[key * NI + i];
                for (int q = 0; q < cnt; q += rows_per_item)
                    inst_items[i].push_back(make_int4(kItemGroup | (logg << 4) | (logk << 8) | (i << kItemUnitShift),
                                                      st + q, std::min(rows_per_item, cnt - q), 0));
            }
        }
        for (const auto &v : inst_items) h_items.insert(h_items.end(), v.begin(), v.end());
        if (NI >= (1 << (31 - kItemUnitShift))) throw std::runtime_error("too many instances in a batch");

        // per item: the row whose mark a seg / long item tests (-1: group item)
        std::vector<int> h_item_row(h_items.size(), -1);
        for (size_t q = 0; q < h_items.size(); q++) {
            const int4 it = h_items[q];
            const int kind = it.x & 15;
            if (kind == kItemSeg)
                h_item_row[q] = h_seg_rows[it.y].x;
            else if (kind == kItemLong)
                h_item_row[q] = it.y;
        }

        P->stats.m = P->m;
        P->stats.n = P->n;
        P->stats.nnz = P->nnz;
        P->stats.m_nonempty = mp;
        P->stats.n_items = static_cast<int>(h_items.size());

        trace.mark("items");
        DevMatrix &M = P->M;
        M.m = mp;
        M.n = n_ids;
        M.n16 = class_start[1];
        M.n_int = class_start[2];
        M.nwords = nwords;
        M.n_items = static_cast<int>(h_items.size());
        M.replicas = replicas ? 1 : 0;
        M.mark_rows = T.row_marking ? 1 : 0;
        M.n_seg_rows = static_cast<int>(h_seg_rows.size());
        M.n_seg_parts = n_parts;
        int *d_col = dev_alloc<int>(n_padded, o);
        double *d_val = dev_alloc<double>(n_padded, o);
        int4 *d_items = dev_alloc<int4>(h_items.size(), o);
        int4 *d_seg_rows = dev_alloc<int4>(h_seg_rows.size(), o);
        int *d_part_row = dev_alloc<int>(h_part_row.size(), o);
        int *d_cptr = dev_alloc<int>(n_ids + 1, o);
        const size_t n_csc = std::max<long long>(1, nnz_total);
        int *d_crows = dev_alloc<int>(n_csc, o);
        int *d_cseg = dev_alloc<int>(n_csc, o);
        int *d_citem = dev_alloc<int>(n_csc, o);
        int *d_item_row = dev_alloc<int>(h_item_row.size(), o);
        int *d_word_unit = dev_alloc<int>(word_unit.size(), o);
        // per layout row: first segment of a split row, group item (-1: none)
        int *t_row_part = dev_alloc<int>(mp, tmp);
        int *t_row_item = dev_alloc<int>(mp, tmp);
        trace.mark("device_setup_alloc");
        upload(d_items, h_items.data(), h_items.size(), s);
        upload(d_item_row, h_item_row.data(), h_item_row.size(), s);
        upload(d_seg_rows, h_seg_rows.data(), h_seg_rows.size(), s);
        upload(d_part_row, h_part_row.data(), h_part_row.size(), s);
        upload(d_word_unit, word_unit.data(), word_unit.size(), s);
        check(hipMemsetAsync(t_row_part, 0xff, sizeof(int) * mp, s), "memset row tables");
        check(hipMemsetAsync(t_row_item, 0xff, sizeof(int) * mp, s), "memset row tables");
        order_row_tables(M.n_items, d_items, M.n_seg_rows, d_seg_rows, t_row_item, t_row_part, s);
        check(hipStreamSynchronize(s), "upload");
        trace.mark("upload_order");
        // rows wider than 64 entries: the long rows and the bins of widths 2^9..2^7
        int n_wide = 0;
        for (int key = 0; key < 4; key++)
            for (int i = 0; i < NI; i++) n_wide += R.count[key * NI + i];
        device_layout_build(csr, col_off, mp, n_wide, n_ids, n_padded, R.src, R.inst, t_row_part, t_row_item, d_rp,
                            P->d_inv_perm, d_col, d_val, d_cptr, d_crows, d_cseg, d_citem, s, trace);
        M.row_ptr = d_rp;
        M.sides = d_sides;
        M.col = d_col;
        M.val = d_val;
        M.items = d_items;
        M.seg_rows = d_seg_rows;
        M.part_row = d_part_row;
        M.col_ptr = d_cptr;
        M.col_rows = d_crows;
        M.col_seg = d_cseg;
        M.col_item = d_citem;
        M.item_row = d_item_row;
        M.perm = d_perm;
        M.word_unit = d_word_unit;

        // snapshot block of a unit: int16 records, then int32, then FP64,
        // each class from a 16-byte boundary; the binary columns' codes in
        // 16-byte groups of 64 columns
        M.n_bin = n_bin;
        M.sb_words = (n_bin + 63) / 64 * 4;
        M.snap_off32 = (M.n16 * 4 + 15) / 16 * 16;
        M.snap_offF = M.snap_off32 + ((M.n_int - M.n16) * 8 + 15) / 16 * 16;
        M.snap_stride = M.snap_offF + 16LL * (M.n - M.n_int) + 16;
        check(hipStreamSynchronize(s), "upload");
        {
            // device bytes of one more bound vector's state (the run state
            // itself is allocated on first use)
            DevState D{};
            UnitTables T;
            if (replicas)
                P->stats.unit_state_bytes =
                    static_cast<long long>(carve_state(M, true, NI, 2, nullptr, D, T) -
                                           carve_state(M, true, NI, 1, nullptr, D, T));
            P->stats.n_units = replicas ? 0 : NI;
        }

        P->mega_threads = T.mega_threads == 256 || T.mega_threads == 1024 ? T.mega_threads : 512;
        P->mega_fn = P->mega_threads == 256    ? reinterpret_cast<const void *>(mega_kernel<256, false>)
                     : P->mega_threads == 1024 ? reinterpret_cast<const void *>(mega_kernel<1024, false>)
                                               : reinterpret_cast<const void *>(mega_kernel<512, false>);
        P->mega_fn_lds = P->mega_threads == 256    ? reinterpret_cast<const void *>(mega_kernel<256, true>)
                         : P->mega_threads == 1024 ? reinterpret_cast<const void *>(mega_kernel<1024, true>)
                                                   : reinterpret_cast<const void *>(mega_kernel<512, true>);
        auto mega_per_cu = [&](const void *fn, size_t lds_bytes) {
            int per_cu = 0;
            check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, P->mega_threads, lds_bytes),
                  "occupancy");
            if (T.mega_blocks_per_cu > 0) per_cu = std::min(per_cu, T.mega_blocks_per_cu);
            return std::max(1, per_cu);
        };
        // the grid with the unit filter's LDS (launches of several units);
        // launches of one unit hold as many binary columns' codes in LDS as
        // keep that many workgroups resident
        const int per_cu = mega_per_cu(P->mega_fn, kFilterLdsBytes);
        P->mega_grid = per_cu * P->cu_count;
        int n_lds = std::min(n_bin, kMaxLdsBinBytes * 4) / 64 * 64;
        while (n_lds > 0 && mega_per_cu(P->mega_fn_lds, static_cast<size_t>(n_lds) / 4) < per_cu)
            n_lds = n_lds / 2 / 64 * 64;
        P->n_lds = n_lds;
        P->stats.n_bin = n_bin;
        P->stats.n_lds = n_lds;
        int sweep_per_cu = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sweep_per_cu, reinterpret_cast<const void *>(sweep_kernel),
                                                           256, 0),
              "occupancy");
        P->sweep_grid = std::max(1, sweep_per_cu) * P->cu_count;
        P->stats.mega_threads = P->mega_threads;
        P->stats.mega_grid = P->mega_grid;
        trace.mark("finish");
    } catch (...) {
        delete P;
        throw;
    }
    return P;
}

// Hint buffers for `count` entries (device).
void reserve_hints(Problem *P, size_t count) {
    if (count <= P->hints_cap) return;
    check(hipStreamSynchronize(P->stream), "sync");
    for (void *p : P->hint_owned) (void)cached_free(p);
    P->hint_owned.clear();
    P->hints_cap = std::max<size_t>(count, 1024);
    P->d_hints = dev_alloc<int>(P->hints_cap, P->hint_owned);
}

// Loads the listed units' bound vectors for a fresh run (status ring
// cleared, round 1 in step `start`) or for units joining a running server
// (`fresh` false).
void load_units(Problem *P, int n_list, const int *list, double *const *lbp, double *const *ubp,
                const int *hint_ptr, const int *hints, const int *max_rounds, int start, bool fresh) {
    hipStream_t s = P->stream;
    DevMatrix &M = P->M;
    DevState &D = P->D;
    // a few units' parameters on the stack: a single-vector run touches no heap
    int4 few[8];
    std::vector<int4> many;
    if (n_list > 8) many.resize(n_list);
    int4 *params = n_list > 8 ? many.data() : few;
    for (int k = 0; k < n_list; k++) {
        const int u = list[k];
        const bool hinted = hint_ptr != nullptr && hint_ptr[k + 1] > hint_ptr[k] && M.mark_rows;
        const int limit = P->mark_limit[P->replicas ? 0 : u];
        params[k] = make_int4(u, limit, hinted ? 0 : 1, max_rounds[k]);
    }
    upload(P->d_list, list, n_list, s);
    upload(P->d_params, params, n_list, s);
    const int n_tables = P->replicas ? n_list : 1;
    upload(P->d_lbp, lbp, n_tables, s);
    upload(P->d_ubp, ubp, n_tables, s);
    hipLaunchKernelGGL(unit_reset_kernel, dim3(std::max(1, std::min(64, (n_list + 255) / 256))), dim3(256), 0, s, D,
                       n_list, P->d_params, start, fresh ? 1 : 0);
    const int gx = std::max(1, std::min(1024, (M.n + 255) / 256));
    const int gy = P->replicas ? std::max(1, std::min(n_list, 1024)) : 1;
    hipLaunchKernelGGL(load_kernel, dim3(gx, gy), dim3(256), 0, s, M, D, n_list, P->d_list, P->d_lbp, P->d_ubp);
    check(hipGetLastError(), "load launch");
    if (hint_ptr != nullptr && M.mark_rows && hint_ptr[n_list] > 0) {
        reserve_hints(P, hint_ptr[n_list]);
        upload(P->d_hint_ptr, hint_ptr, n_list + 1, s);
        upload(P->d_hints, hints, hint_ptr[n_list], s);
        hipLaunchKernelGGL(hint_kernel, dim3(16, std::max(1, std::min(n_list, 1024))), dim3(256), 0, s, M, D, n_list,
                           P->d_list, P->d_hint_ptr, P->d_hints, P->d_inv_perm, start);
        check(hipGetLastError(), "hint launch");
    }
}

// One step of the listed units (already uploaded as the active list).
void launch_step(Problem *P, int step, int guard, int first_step) {
    const DevMatrix &M = P->M;
    const DevState &D = P->D;
    hipStream_t s = P->stream;
    const long long n_work = P->replicas ? static_cast<long long>(M.n_items) * D.n_active : M.n_items;
    const int apply_grid = std::max(1, std::min(1024, grid_for(static_cast<long long>(M.n_seg_parts) *
                                                                 (P->replicas ? D.n_active : 1))));
    const int sweep_grid = static_cast<int>(std::min<long long>(grid_for(n_work), P->sweep_grid));
    hipLaunchKernelGGL(sweep_kernel, dim3(sweep_grid), dim3(256), 0, s, M, D, step, guard, first_step);
    if (M.n_seg_parts > 0)
        hipLaunchKernelGGL(apply_split_kernel, dim3(apply_grid), dim3(256), 0, s, M, D, step, guard, first_step);
    const long long pub_waves = static_cast<long long>(P->replicas ? D.n_active : 1) * ((M.nwords + 7) / 8);
    hipLaunchKernelGGL(publish_kernel, dim3(std::max(1, std::min(8192, grid_for(pub_waves)))), dim3(256), 0, s, M, D,
                       step, guard, first_step, P->n_units);
}

// Runs steps 1.. for the listed units (loaded), at most max_rounds steps.
// Returns the device time of the step loop.
double run_steps(Problem *P, int sync_type, int max_rounds) {
    const DevMatrix &M = P->M;
    DevState &D = P->D;
    hipStream_t s = P->stream;
    check(hipEventRecord(P->ev0, s), "event");
    if (sync_type == CPU_LOOP_ID) {
        // one set of launches per step; the host reads the step's flags back
        for (int step = 1; step <= max_rounds; step++) {
            launch_step(P, step, 0, 1);
            check(hipGetLastError(), "step launch");
            int flags[kStatusWords];
            int err = 0;
            check(hipMemcpyAsync(flags, D.status + kStatusWords * (step % kStatusRows), sizeof(flags),
                                 hipMemcpyDeviceToHost, s),
                  "flags D2H");
            check(hipMemcpyAsync(&err, D.global + kGlobalError, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "step sync");
            if (err) throw std::runtime_error("device wait timed out");
            if (!flags[kStatusChanged]) break;
        }
    } else if (sync_type == GPU_LOOP_ID && !P->graphs_enabled) {
        // the same chunks of self-stopping launches, issued without capture
        // (no graph is created, so nothing is allocated)
        const int chunk = std::max(1, P->tuning.graph_rounds);
        for (int first = 1; first <= max_rounds; first += chunk) {
            const int last = std::min(max_rounds, first + chunk - 1);
            for (int step = first; step <= last; step++) launch_step(P, step, 1, 1);
            check(hipGetLastError(), "step launch");
            int flags[kStatusWords];
            int err = 0;
            check(hipMemcpyAsync(flags, D.status + kStatusWords * (last % kStatusRows), sizeof(flags),
                                 hipMemcpyDeviceToHost, s),
                  "flags D2H");
            check(hipMemcpyAsync(&err, D.global + kGlobalError, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "step sync");
            if (err) throw std::runtime_error("device wait timed out");
            if (!flags[kStatusChanged]) break;
        }
    } else if (sync_type == GPU_LOOP_ID) {
        // captured chunks of steps whose launches stop themselves on the device
        const int chunk = std::max(1, P->tuning.graph_rounds);
        const long long key = static_cast<long long>(D.n_active) * 1000 + max_rounds;
        if (key != P->graph_key) {
            P->drop_graphs();
            P->graph_key = key;
        }
        for (int first = 1, ci = 0; first <= max_rounds; first += chunk, ci++) {
            const int last = std::min(max_rounds, first + chunk - 1);
            if (static_cast<int>(P->graphs.size()) <= ci) {
                P->graphs.resize(ci + 1, nullptr);
                P->graph_last.resize(ci + 1, 0);
            }
            if (!P->graphs[ci]) {
                hipGraph_t graph;
                check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
                for (int step = first; step <= last; step++) launch_step(P, step, 1, 1);
                check(hipStreamEndCapture(s, &graph), "end capture");
                hipGraphExec_t exec;
                check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
                (void)hipGraphDestroy(graph);
                P->graphs[ci] = exec;
                P->graph_last[ci] = last;
            }
            check(hipGraphLaunch(P->graphs[ci], s), "graph launch");
            int flags[kStatusWords];
            int err = 0;
            check(hipMemcpyAsync(flags, D.status + kStatusWords * (last % kStatusRows), sizeof(flags),
                                 hipMemcpyDeviceToHost, s),
                  "flags D2H");
            check(hipMemcpyAsync(&err, D.global + kGlobalError, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "graph sync");
            if (err) throw std::runtime_error("device wait timed out");
            // the next chunk is needed only if the chunk's last step still changed
            if (!flags[kStatusChanged]) break;
        }
    } else {
        int first = 1, steps = max_rounds, n_units = P->n_units;
        // dynamic LDS: the unit filter's arrays for several units, else the
        // binary columns' codes
        DevState Dm = D;
        size_t lds_bytes = kFilterLdsBytes;
        const void *fn = P->mega_fn;
        if ((P->replicas ? D.n_active : n_units) == 1) {
            lds_bytes = static_cast<size_t>(P->n_lds) / 4;
            Dm.n_lds = P->n_lds;
            if (P->n_lds > 0) fn = P->mega_fn_lds;
        }
        void *args[] = {const_cast<DevMatrix *>(&M), &Dm, &first, &steps, &n_units};
        // as many workgroups as the largest phase has waves of work (sweep
        // items, bitmap chunks, split-row segments), up to the resident
        // maximum: a small problem then pays its grid barriers over a few
        // workgroups instead of the whole chip
        const long long units = P->replicas ? D.n_active : 1;
        const long long need = units * std::max<long long>({static_cast<long long>(M.n_items),
                                                            static_cast<long long>((M.nwords + 7) / 8),
                                                            static_cast<long long>(M.n_seg_parts)});
        const int waves_per_group = P->mega_threads / 64;
        const int grid = static_cast<int>(
            std::max(1LL, std::min<long long>(P->mega_grid, (need + waves_per_group - 1) / waves_per_group)));
        check(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(P->mega_threads), args, lds_bytes, s),
              "mega_kernel launch");
        int g[kGlobalWords];
        check(hipMemcpyAsync(g, D.global, sizeof(g), hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "mega sync");
        if (g[kGlobalError] || g[kGlobalSteps] < 0) throw std::runtime_error("device wait timed out");
    }
    check(hipEventRecord(P->ev1, s), "event");
    check(hipEventSynchronize(P->ev1), "event sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, P->ev0, P->ev1), "elapsed");
    return ms;
}

std::vector<UnitState> read_units(Problem *P, int n_list, const int *list) {
    std::vector<UnitState> all(P->n_units);
    check(hipMemcpyAsync(all.data(), P->D.units, sizeof(UnitState) * P->n_units, hipMemcpyDeviceToHost, P->stream),
          "units D2H");
    check(hipStreamSynchronize(P->stream), "units sync");
    std::vector<UnitState> out(n_list);
    for (int k = 0; k < n_list; k++) out[k] = all[list[k]];
    return out;
}

// State words of unit u (synchronous).
UnitState read_unit(Problem *P, int u) {
    UnitState st;
    check(hipMemcpyAsync(&st, P->D.units + u, sizeof(UnitState), hipMemcpyDeviceToHost, P->stream), "unit D2H");
    check(hipStreamSynchronize(P->stream), "unit sync");
    return st;
}

// A complete run of one bound vector (device pointers) as unit 0, from
// round 1.  Host buffers stay on the stack.
RunResult run_single(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds) {
    if (sync_type < CPU_LOOP_ID || sync_type > MEGAKERNEL_ID) throw std::runtime_error("Unknown sync type");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    const int list[1] = {0};
    const int caps[1] = {max_rounds};
    P->D.n_active = 1;
    load_units(P, 1, list, &d_lbs, &d_ubs, nullptr, nullptr, caps, 1, true);
    UnitState st = read_unit(P, 0);
    if (st.envelope) throw std::runtime_error("bounds outside the problem's bound envelope");
    double ms = 0.0;
    if (!st.crossed) {
        ms = run_steps(P, sync_type, max_rounds);
        st = read_unit(P, 0);
    }
    RunResult r = outcome(st);
    r.kernel_ms = ms;
    const int gx = std::max(1, std::min(1024, (P->M.n + 255) / 256));
    hipLaunchKernelGGL(unpack_kernel, dim3(gx, 1), dim3(256), 0, P->stream, P->M, P->D, 1, P->d_list, P->d_lbp,
                       P->d_ubp);
    check(hipGetLastError(), "unpack launch");
    check(hipStreamSynchronize(P->stream), "unpack sync");
    return r;
}

// A complete run of the listed units from round 1.
void run_units(Problem *P, int n_list, const int *list, double *const *lbp, double *const *ubp, const int *hint_ptr,
               const int *hints, int sync_type, int max_rounds, RunResult *out) {
    if (sync_type < CPU_LOOP_ID || sync_type > MEGAKERNEL_ID) throw std::runtime_error("Unknown sync type");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    std::vector<int> caps(n_list, max_rounds);
    P->D.n_active = P->replicas ? n_list : 0;
    load_units(P, n_list, list, lbp, ubp, hint_ptr, hints, caps.data(), 1, true);
    // bounds outside the envelope are an error before any round
    std::vector<UnitState> st = read_units(P, n_list, list);
    bool any_active = false;
    for (const auto &u : st) {
        if (u.envelope) throw std::runtime_error("bounds outside the problem's bound envelope");
        any_active |= !u.crossed;
    }
    double ms = 0.0;
    if (any_active) {
        ms = run_steps(P, sync_type, max_rounds);
        st = read_units(P, n_list, list);
    }
    for (int k = 0; k < n_list; k++) {
        out[k] = outcome(st[k]);
        out[k].kernel_ms = ms;
    }
    const int gx = std::max(1, std::min(1024, (P->M.n + 255) / 256));
    const int gy = P->replicas ? std::max(1, std::min(n_list, 1024)) : 1;
    hipLaunchKernelGGL(unpack_kernel, dim3(gx, gy), dim3(256), 0, P->stream, P->M, P->D, n_list, P->d_list,
                       P->d_lbp, P->d_ubp);
    check(hipGetLastError(), "unpack launch");
    check(hipStreamSynchronize(P->stream), "unpack sync");
}

}  // namespace

Problem *problem_create(int m, int n, int nnz, const int *col, const int *row_ptr, const double *val,
                        const double *lhs, const double *rhs, const int *vartype, const double *env_lb,
                        const double *env_ub, const Tuning *tuning) {
    if (n > kIdMask) throw std::runtime_error("too many columns");
    InstanceView v;
    v.m = m;
    v.n = n;
    v.nnz = nnz;
    v.col = col;
    v.row_ptr = row_ptr;
    v.val = val;
    v.lhs = lhs;
    v.rhs = rhs;
    v.vartype = vartype;
    v.env_lb = env_lb;
    v.env_ub = env_ub;
    return build({v}, true, tuning);
}

Problem *batch_create(const std::vector<InstanceView> &instances, const Tuning *tuning) {
    return build(instances, false, tuning);
}

void problem_destroy(Problem *p) {
    {
        std::lock_guard<std::mutex> lock(p->run_mutex);
        if (p->servers > 0) throw std::runtime_error("a server is attached to this problem: destroy the server first");
    }
    delete p;
}

ProblemStats problem_stats(const Problem *p) { return p->stats; }

namespace {
// With P->run_mutex held.  A server owns the problem's unit slots (slot k
// is unit k) for its whole life: nothing else may run units or reallocate
// their state meanwhile.
void check_no_server(const Problem *P) {
    if (P->servers > 0)
        throw std::runtime_error("a server is attached to this problem: its unit slots belong to the server");
}

// With P->run_mutex held: one bound vector (device pointers), unit 0.
RunResult propagate_device_locked(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds) {
    if (!P->replicas) throw std::runtime_error("use the batch entry points for a batch of instances");
    check_no_server(P);
    ensure_units(P);
    return run_single(P, d_lbs, d_ubs, sync_type, max_rounds);
}
}  // namespace

void problem_reserve(Problem *P, int capacity) {
    if (!P->replicas) throw std::runtime_error("a batch of instances has no node capacity");
    if (capacity < 1) throw std::runtime_error("capacity must be positive");
    std::lock_guard<std::mutex> lock(P->run_mutex);
    check_no_server(P);
    if (capacity != P->capacity) alloc_units(P, capacity);
}

RunResult problem_propagate_device(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds) {
    std::lock_guard<std::mutex> lock(P->run_mutex);
    return propagate_device_locked(P, d_lbs, d_ubs, sync_type, max_rounds);
}

// The whole call holds the problem's lock: the bound vectors travel on the
// problem's stream, which another thread's call may otherwise be capturing
// into a graph (GPU_LOOP) or waiting on.
RunResult problem_propagate(Problem *P, double *lbs, double *ubs, int sync_type, int max_rounds) {
    std::lock_guard<std::mutex> lock(P->run_mutex);
    check_no_server(P);
    const int n = P->n;
    double *d_b = nullptr;
    check(cached_malloc(&d_b, 2 * sizeof(double) * static_cast<size_t>(n)), "hipMalloc bounds");
    RunResult res;
    try {
        upload(d_b, lbs, n, P->stream);
        upload(d_b + n, ubs, n, P->stream);
        res = propagate_device_locked(P, d_b, d_b + n, sync_type, max_rounds);
        download(lbs, d_b, n, P->stream);
        download(ubs, d_b + n, n, P->stream);
    } catch (...) {
        (void)hipStreamSynchronize(P->stream);
        (void)cached_free(d_b);
        throw;
    }
    check(cached_free(d_b), "hipFree");
    return res;
}

RunResult problem_work_profile(Problem *P, double *lbs, double *ubs, int max_rounds, std::vector<RoundWork> &work) {
    if (!P->replicas) throw std::runtime_error("work profiles are for a single instance");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    std::lock_guard<std::mutex> lock(P->run_mutex);
    check_no_server(P);
    ensure_units(P);
    const DevMatrix &M = P->M;
    hipStream_t s = P->stream;
    const int n = P->n;
    // layout rows and columns, host side
    std::vector<int> rp(M.m + 1);
    check(hipMemcpy(rp.data(), M.row_ptr, sizeof(int) * (M.m + 1), hipMemcpyDeviceToHost), "D2H row_ptr");
    const int n_padded = rp[M.m] & ~(kRowAlign - 1);
    std::vector<int> cols(std::max(1, n_padded));
    if (n_padded > 0)
        check(hipMemcpy(cols.data(), M.col, sizeof(int) * n_padded, hipMemcpyDeviceToHost), "D2H col");
    std::vector<int> marks(M.m), stamp(M.n, 0);
    // split rows: a marked row recomputes only some segments (the kernel's
    // rule: every row processed in a full sweep; otherwise segments not yet
    // computed in the run or marked for the step)
    std::vector<int4> seg_rows(M.n_seg_rows);
    if (M.n_seg_rows > 0)
        check(hipMemcpy(seg_rows.data(), M.seg_rows, sizeof(int4) * M.n_seg_rows, hipMemcpyDeviceToHost),
              "D2H seg_rows");
    std::vector<int> split_of(M.m, -1);
    for (int q = 0; q < M.n_seg_rows; q++) split_of[seg_rows[q].x] = q;
    std::vector<int> seg_marks(std::max(1, M.n_seg_parts)), seg_stamp(std::max(1, M.n_seg_parts), 0);
    double *d_b = nullptr;
    check(cached_malloc(&d_b, 2 * sizeof(double) * static_cast<size_t>(n)), "hipMalloc bounds");
    RunResult res;
    work.clear();
    try {
        check(hipMemcpy(d_b, lbs, n * sizeof(double), hipMemcpyHostToDevice), "H2D lbs");
        check(hipMemcpy(d_b + n, ubs, n * sizeof(double), hipMemcpyHostToDevice), "H2D ubs");
        const int list[1] = {0};
        double *lbp = d_b, *ubp = d_b + n;
        const int caps[1] = {max_rounds};
        P->D.n_active = 1;
        load_units(P, 1, list, &lbp, &ubp, nullptr, nullptr, caps, 1, true);
        std::vector<UnitState> st = read_units(P, 1, list);
        if (st[0].envelope) throw std::runtime_error("bounds outside the problem's bound envelope");
        for (int step = 1; step <= max_rounds; step++) {
            if (step == 1 ? st[0].crossed != 0 : !active_after(st[0], step - 1)) break;
            RoundWork w;
            const bool all_rows = !M.mark_rows || st[0].markall == step;
            if (!all_rows) {
                check(hipMemcpy(marks.data(), P->D.marks, sizeof(int) * M.m, hipMemcpyDeviceToHost), "D2H marks");
                if (M.n_seg_parts > 0)
                    check(hipMemcpy(seg_marks.data(), P->D.seg_marks, sizeof(int) * M.n_seg_parts,
                                    hipMemcpyDeviceToHost),
                          "D2H seg_marks");
            }
            auto count_entries = [&](int start, int from, int to) {
                w.nnz += to - from;
                for (int e = from; e < to; e++) {
                    const int id = cols[start + e] & kIdMask;
                    if (stamp[id] != step) {
                        stamp[id] = step;
                        w.cols++;
                    }
                }
            };
            for (int r = 0; r < M.m; r++) {
                if (!all_rows && marks[r] != step) continue;
                const int a = rp[r], b = rp[r + 1];
                const int start = a & ~(kRowAlign - 1);
                const int len = (b & ~(kRowAlign - 1)) - start - (a & (kRowAlign - 1));
                w.rows++;
                const int q = split_of[r];
                if (q < 0) {
                    count_entries(start, 0, len);
                    continue;
                }
                for (int g = 0; g < seg_rows[q].z; g++) {
                    const int part = seg_rows[q].y + g;
                    if (!all_rows && seg_stamp[part] >= 1 && seg_marks[part] != step) continue;
                    seg_stamp[part] = step;
                    count_entries(start, g * kSegLen, std::min(len, (g + 1) * kSegLen));
                }
            }
            launch_step(P, step, 0, 1);
            check(hipGetLastError(), "step launch");
            int cnt[kMoveLanes * kMoveStride];
            check(hipMemcpyAsync(cnt, P->D.moved + moved_index(0, step % kMoveSlots, 0), sizeof(cnt),
                                 hipMemcpyDeviceToHost, s),
                  "D2H moved");
            check(hipStreamSynchronize(s), "step sync");
            for (int j = 0; j < kMoveLanes; j++) w.accepted += cnt[j * kMoveStride];
            work.push_back(w);
            st = read_units(P, 1, list);
        }
        res = outcome(st[0]);
        const int gx = std::max(1, std::min(1024, (M.n + 255) / 256));
        hipLaunchKernelGGL(unpack_kernel, dim3(gx, 1), dim3(256), 0, s, P->M, P->D, 1, P->d_list, P->d_lbp,
                           P->d_ubp);
        check(hipGetLastError(), "unpack launch");
        check(hipStreamSynchronize(s), "unpack sync");
        check(hipMemcpy(lbs, d_b, n * sizeof(double), hipMemcpyDeviceToHost), "D2H lbs");
        check(hipMemcpy(ubs, d_b + n, n * sizeof(double), hipMemcpyDeviceToHost), "D2H ubs");
    } catch (...) {
        (void)cached_free(d_b);
        throw;
    }
    check(cached_free(d_b), "hipFree");
    return res;
}

RunResult problem_propagate_measured(Problem *P, double *lbs, double *ubs, int max_rounds, const double *wl,
                                     const double *wu, const double *ll, const double *lu, std::vector<double> &fin,
                                     std::vector<int> &inf) {
    if (!P->replicas) throw std::runtime_error("progress measures are for a single instance");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    std::lock_guard<std::mutex> lock(P->run_mutex);
    check_no_server(P);
    ensure_units(P);
    hipStream_t s = P->stream;
    const int n = P->n;
    const size_t nb = sizeof(double) * static_cast<size_t>(n);
    const int gx = std::max(1, std::min(1024, (n + 255) / 256));
    // bounds of the round | previous round | weakest | limit, each lb then ub;
    // then the workgroup partials and the per-round sums
    const size_t bytes = 8 * nb + gx * (sizeof(double) + sizeof(int)) + kMaxRounds * (sizeof(double) + sizeof(int));
    char *buf = nullptr;
    check(cached_malloc(&buf, bytes), "hipMalloc measure buffers");
    double *d_b = reinterpret_cast<double *>(buf);
    double *d_prev = d_b + 2 * static_cast<size_t>(n);
    double *d_w = d_prev + 2 * static_cast<size_t>(n);
    double *d_l = d_w + 2 * static_cast<size_t>(n);
    double *part_f = d_l + 2 * static_cast<size_t>(n);
    double *out_f = part_f + gx;
    int *part_i = reinterpret_cast<int *>(out_f + kMaxRounds);
    int *out_i = part_i + gx;
    RunResult res;
    fin.clear();
    inf.clear();
    try {
        const double *src[8] = {lbs, ubs, lbs, ubs, wl, wu, ll, lu};
        for (int q = 0; q < 8; q++)
            check(hipMemcpy(d_b + q * static_cast<size_t>(n), src[q], nb, hipMemcpyHostToDevice), "H2D measure");
        const int list[1] = {0};
        double *lbp = d_b, *ubp = d_b + n;
        const int caps[1] = {max_rounds};
        P->D.n_active = 1;
        load_units(P, 1, list, &lbp, &ubp, nullptr, nullptr, caps, 1, true);
        std::vector<UnitState> st = read_units(P, 1, list);
        if (st[0].envelope) throw std::runtime_error("bounds outside the problem's bound envelope");
        int rounds = 0;
        for (int step = 1; step <= max_rounds; step++) {
            if (step == 1 ? st[0].crossed != 0 : !active_after(st[0], step - 1)) break;
            launch_step(P, step, 0, 1);
            check(hipGetLastError(), "step launch");
            hipLaunchKernelGGL(unpack_kernel, dim3(gx, 1), dim3(256), 0, s, P->M, P->D, 1, P->d_list, P->d_lbp,
                               P->d_ubp);
            hipLaunchKernelGGL(measure_kernel, dim3(gx), dim3(256), 0, s, n, d_b, d_b + n, d_w, d_w + n, d_l,
                               d_l + n, d_prev, d_prev + n, part_f, part_i);
            hipLaunchKernelGGL(measure_total_kernel, dim3(1), dim3(64), 0, s, gx, part_f, part_i, out_f + step - 1,
                               out_i + step - 1);
            check(hipGetLastError(), "measure launch");
            st = read_units(P, 1, list);
            rounds = step;
        }
        res = outcome(st[0]);
        hipLaunchKernelGGL(unpack_kernel, dim3(gx, 1), dim3(256), 0, s, P->M, P->D, 1, P->d_list, P->d_lbp,
                           P->d_ubp);
        check(hipGetLastError(), "unpack launch");
        check(hipStreamSynchronize(s), "unpack sync");
        fin.resize(rounds);
        inf.resize(rounds);
        if (rounds > 0) {
            check(hipMemcpy(fin.data(), out_f, sizeof(double) * rounds, hipMemcpyDeviceToHost), "D2H measure");
            check(hipMemcpy(inf.data(), out_i, sizeof(int) * rounds, hipMemcpyDeviceToHost), "D2H measure");
        }
        check(hipMemcpy(lbs, d_b, nb, hipMemcpyDeviceToHost), "D2H lbs");
        check(hipMemcpy(ubs, d_b + n, nb, hipMemcpyDeviceToHost), "D2H ubs");
    } catch (...) {
        (void)cached_free(buf);
        throw;
    }
    check(cached_free(buf), "hipFree");
    return res;
}

void nodes_propagate_device(Problem *P, int count, double *const *d_lbs, double *const *d_ubs, const int *hint_ptr,
                            const int *hints, int sync_type, int max_rounds, RunResult *out) {
    if (!P->replicas) throw std::runtime_error("use the batch entry points for a batch of instances");
    if (count < 1) throw std::runtime_error("no bound vector");
    if (hint_ptr) {
        if (hint_ptr[0] != 0) throw std::runtime_error("hint_ptr[0] must be 0");
        for (int k = 0; k < count; k++)
            if (hint_ptr[k + 1] < hint_ptr[k]) throw std::runtime_error("hint_ptr not monotone");
        for (int h = 0; h < hint_ptr[count]; h++)
            if (!hints || hints[h] < 0 || hints[h] >= P->n) throw std::runtime_error("hint column out of range");
    }
    std::lock_guard<std::mutex> lock(P->run_mutex);
    check_no_server(P);
    if (count > P->capacity) alloc_units(P, count);
    std::vector<int> list(count);
    for (int k = 0; k < count; k++) list[k] = k;
    run_units(P, count, list.data(), d_lbs, d_ubs, hint_ptr, hints, sync_type, max_rounds, out);
}

void batch_propagate_device(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds,
                            RunResult *out) {
    if (P->replicas) throw std::runtime_error("not a batch of instances");
    std::lock_guard<std::mutex> lock(P->run_mutex);
    ensure_units(P);
    std::vector<int> list(P->n_inst);
    for (int k = 0; k < P->n_inst; k++) list[k] = k;
    run_units(P, P->n_inst, list.data(), &d_lbs, &d_ubs, nullptr, nullptr, sync_type, max_rounds, out);
}

// ---------------------------------------------------------------------------
// propagation that allocates nothing
// ---------------------------------------------------------------------------

// The run state of the allocation-free entry point lives in this fixed array
// of device memory, which every device the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device array: its device address comes from hipGetSymbolAddress (once per
// device, with that device current), and only that address is handed to
// kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                 // the address table
char *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device
std::mutex g_scratch_use[kMaxDevices];   // one call at a time uses a device's array

// Device address of `scratch` on device `dev` (the current device): looked
// up on the first call on that device.
char *scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = static_cast<char *>(ptr);
    }
    return g_scratch_ptr[dev];
}

// Runs the problem on a borrowed run state for the lifetime of the object:
// the owned state (if any), its launch tables and its captured graphs are
// set aside and restored afterwards.
class BorrowedState {
  public:
    BorrowedState(Problem *P, char *base)
        : P_(P), D_(P->D), list_(P->d_list), params_(P->d_params), lbp_(P->d_lbp), ubp_(P->d_ubp),
          hint_ptr_(P->d_hint_ptr), capacity_(P->capacity), n_units_(P->n_units), graphs_(P->graphs_enabled) {
        bind_state(P, base, 1);
        P->graphs_enabled = false;
    }
    ~BorrowedState() {
        P_->D = D_;
        P_->d_list = list_;
        P_->d_params = params_;
        P_->d_lbp = lbp_;
        P_->d_ubp = ubp_;
        P_->d_hint_ptr = hint_ptr_;
        P_->capacity = capacity_;
        P_->n_units = n_units_;
        P_->graphs_enabled = graphs_;
    }
    BorrowedState(const BorrowedState &) = delete;
    BorrowedState &operator=(const BorrowedState &) = delete;

  private:
    Problem *P_;
    DevState D_;
    int *list_;
    int4 *params_;
    double **lbp_, **ubp_;
    int *hint_ptr_;
    int capacity_, n_units_;
    bool graphs_;
};
}  // namespace

size_t problem_run_state_bytes(const Problem *P) {
    DevState D{};
    UnitTables T;
    return carve_state(P->M, P->replicas, P->n_inst, 1, nullptr, D, T);
}

RunResult problem_propagate_no_alloc(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds) {
    if (!P->replicas) throw std::runtime_error("use the batch entry points for a batch of instances");
    const size_t need = problem_run_state_bytes(P);
    if (need > kScratchBytes)
        throw std::runtime_error("the problem needs " + std::to_string(need) +
                                 " bytes of run state, more than the " + std::to_string(kScratchBytes) +
                                 " bytes of the allocation-free entry point's device array");
    int dev = -1;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev != P->device) throw std::runtime_error("the problem's device is not the current device");
    char *base = scratch_address(dev);
    std::lock_guard<std::mutex> use(g_scratch_use[dev]);
    std::lock_guard<std::mutex> lock(P->run_mutex);
    BorrowedState borrowed(P, base);
    // The array keeps the previous call's contents.  Everything a run reads
    // before writing is reset by its loading launches: the unit's state words
    // and candidate counters, the status ring, global words and barrier
    // (unit_reset_kernel); the snapshot, accumulators, bitmap, row marks,
    // split-row counters and partial stamps (load_kernel).
    return run_single(P, d_lbs, d_ubs, sync_type, max_rounds);
}

// ---------------------------------------------------------------------------
// autotuning
// ---------------------------------------------------------------------------

Tuning autotune(const std::vector<InstanceView> &instances, const double *lbs, const double *ubs, int sync_type,
                const Tuning *start, std::vector<AutotuneTrial> *trials) {
    if (instances.empty()) throw std::runtime_error("no instance");
    const bool batch = instances.size() > 1;
    long long n_total = 0;
    for (const auto &v : instances) n_total += v.n;
    // device copies of the bounds every trial starts from
    double *d = nullptr;
    check(cached_malloc(&d, sizeof(double) * 4 * n_total), "hipMalloc");
    double *root_l = d, *root_u = d + n_total, *work_l = d + 2 * n_total, *work_u = d + 3 * n_total;
    Tuning best = start ? *start : Tuning();
    double best_ms = 1e300;
    auto measure = [&](const Tuning &t) {
        Problem *P = batch ? batch_create(instances, &t) : build({instances[0]}, true, &t);
        double ms = 1e300;
        try {
            std::vector<RunResult> out(batch ? instances.size() : 1);
            for (int rep = 0; rep < 3; rep++) {
                check(hipMemcpy(work_l, root_l, sizeof(double) * n_total, hipMemcpyDeviceToDevice), "D2D");
                check(hipMemcpy(work_u, root_u, sizeof(double) * n_total, hipMemcpyDeviceToDevice), "D2D");
                check(hipDeviceSynchronize(), "sync");
                if (batch) batch_propagate_device(P, work_l, work_u, sync_type, kMaxRounds, out.data());
                else out[0] = problem_propagate_device(P, work_l, work_u, sync_type, kMaxRounds);
                if (rep > 0) ms = std::min(ms, out[0].kernel_ms);  // the first run warms up
            }
        } catch (...) {
            problem_destroy(P);
            throw;
        }
        problem_destroy(P);
        if (trials) {
            AutotuneTrial tr;
            tr.tuning = t;
            tr.kernel_ms = ms;
            trials->push_back(tr);
        }
        return ms;
    };
    try {
        check(hipMemcpy(root_l, lbs, sizeof(double) * n_total, hipMemcpyHostToDevice), "H2D");
        check(hipMemcpy(root_u, ubs, sizeof(double) * n_total, hipMemcpyHostToDevice), "H2D");
        best_ms = measure(best);
        // coordinate search: work-item size, lanes per row, persistent grid, row marking
        const int items[] = {512, 1024, 2048, 4096};
        const int lanes[3][10] = {{0, 0, 0, 1, 2, 3, 4, 5, 6, 6},   // 4 entries per lane
                                  {0, 0, 1, 2, 3, 4, 5, 6, 6, 6},   // 2 entries per lane
                                  {0, 1, 2, 3, 4, 5, 6, 6, 6, 6}};  // 1 entry per lane
        auto try_one = [&](const Tuning &t) {
            const double ms = measure(t);
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        };
        for (int v : items)
            if (v != best.item_nnz) {
                Tuning t = best;
                t.item_nnz = v;
                try_one(t);
            }
        for (const auto &g : lanes) {
            Tuning t = best;
            std::copy(g, g + 10, t.group_log);
            if (!std::equal(g, g + 10, best.group_log)) try_one(t);
        }
        if (sync_type == MEGAKERNEL_ID)
            for (int w : {256, 512, 1024})
                if (w != best.mega_threads) {
                    Tuning t = best;
                    t.mega_threads = w;
                    try_one(t);
                }
        {
            Tuning t = best;
            t.row_marking = best.row_marking ? 0 : 1;
            try_one(t);
        }
        if (best.row_marking) {
            // the full-sweep threshold: the default and larger ones (marks
            // pushed in more rounds)
            const auto &v = instances[0];
            const long long def = std::max(64LL, static_cast<long long>(v.m) * v.n / (4LL * std::max(1, v.nnz)));
            for (long long f : {4LL, 16LL}) {
                Tuning t = best;
                t.mark_limit = static_cast<int>(std::min(def * f, 1LL << 30));
                if (t.mark_limit != best.mark_limit) try_one(t);
            }
        }
    } catch (...) {
        (void)cached_free(d);
        throw;
    }
    check(cached_free(d), "hipFree");
    return best;
}

// ---------------------------------------------------------------------------
// serving
// ---------------------------------------------------------------------------

struct Server {
    struct Request {
        int64_t ticket = 0;
        int inbox = -1;                  // pinned inbox entry holding the vectors, or -1
        std::vector<double> lb, ub;      // the vectors when no inbox entry was free
        bool delta = false;              // base vector + changed columns
        std::vector<int> cols;           // changed columns (delta) / hints (full)
        std::vector<double> dlb, dub;    // new bounds of the changed columns (delta)
        int max_rounds = kMaxRounds;
        // caller-owned pinned buffers (zero-copy requests): the vectors are
        // transferred from in_* at admission and the result into out_*
        const double *pin_lb = nullptr, *pin_ub = nullptr;
        double *out_lb = nullptr, *out_ub = nullptr;
        int stage = -1;                  // device staging entry the vectors were prefetched into
    };
    struct Done {
        int outbox = -1;                 // pinned outbox entry holding the result, or -1
        int park = -1;                   // device parking entry holding the result, or -1
        std::vector<double> lb, ub;      // the result (full vector), or empty
        bool changes = false;            // the result as columns that differ from the base
        std::vector<int> cols;
        std::vector<double> clb, cub;
        bool in_place = false;           // the result went to the caller's pinned buffers
        RunResult r;
    };
    static constexpr int kDiffCap = 4096;  // changed columns returned without the full vector
    Problem *P = nullptr;
    bool attached = false;               // counted in P->servers
    int slots = 0;
    size_t n = 0;
    std::deque<Request> queue;
    // Request-side state shared between the caller's threads and the
    // stepping thread: the queue, tickets, the free lists of the pinned and
    // device staging and of the spare vectors, finished requests, busy slot
    // tickets.  The device work of a step runs without it.
    std::mutex mu;
    std::condition_variable cv;
    // Back-pressure while the stepping thread runs: submitters wait here for
    // room in the queue (queue_cap requests) and, for copied requests, for a
    // free pinned inbox entry; woken when a step admits requests or frees
    // entries.
    std::condition_variable space_cv;
    size_t queue_cap = 0;
    std::thread worker;
    bool running = false, stop = false;
    std::string worker_error;
    int busy = 0;                         // slots holding a request
    long long parked_results = 0;         // results retired into device parking (no free outbox entry)
    bool trace = std::getenv("GDP_TRACE_SERVER") != nullptr;
    double slowest_ms = 0.0, slowest_admit_ms = 0.0, slowest_copies_ms = 0.0, slowest_run_ms = 0.0;
    int slowest_step = 0;
    long long pageable_results = 0;       // results retired into host vectors (parking full too)
    std::deque<int64_t> completed_acc;    // completed by the stepping thread, not yet taken
    std::vector<int64_t> slot_ticket;    // -1: free
    std::unordered_map<int64_t, Done> done;
    std::vector<int64_t> last_completed;
    int64_t next_ticket = 1;
    int step = 1;
    double *d_buf = nullptr;             // slots x (lb, ub) bound vectors
    // Device parking for finished results that find no free pinned outbox
    // entry (the callers collect more slowly than the server finishes): a
    // device-to-device copy in the step, the copy to the host in the caller's
    // result call, so that the stepping thread neither waits for a copy into
    // pageable memory nor allocates host memory for it.  One entry per slot.
    double *d_park = nullptr;
    std::vector<int> park_free;
    double *d_base = nullptr;            // base vector of delta requests
    bool has_base = false;
    std::vector<double> base_lb, base_ub;  // host copy of the base
    std::vector<char> slot_delta;        // slot holds a delta request
    std::vector<double *> slot_out_lb, slot_out_ub;  // caller's pinned result buffers, or null
    std::vector<char> slot_inplace;      // zero-copy request whose result replaces its input:
                                         // only the changed columns come back
    double *d_orig = nullptr;            // slots x (lb, ub): inputs of in-place requests
    double **d_blp = nullptr, **d_bup = nullptr;  // base vector of each diffed slot
    double **d_cp = nullptr;             // admission copies: sources, then destinations
    // changed columns of finished delta requests (device and pinned host)
    int *d_ccnt = nullptr, *d_ccols = nullptr;
    double *d_cl = nullptr, *d_cu = nullptr;
    int *h_ccnt = nullptr, *h_ccols = nullptr;
    // the changes packed for the copy back: columns and bounds, each
    // vector's start (device, then pinned host)
    int *d_kcols = nullptr, *d_off = nullptr, *h_off = nullptr;
    double *d_kl = nullptr, *d_ku = nullptr;
    double *h_cl = nullptr, *h_cu = nullptr;
    int diff_slots = 0;
    std::vector<double *> slot_lb, slot_ub;
    double *inbox = nullptr, *outbox = nullptr;  // pinned host memory
    std::vector<int> inbox_free, outbox_free;
    // prefetch: request vectors from pinned memory are copied at submission
    // into device staging entries on a copy stream, overlapping the steps
    hipStream_t cs = nullptr;
    hipEvent_t cs_tail = nullptr;        // recorded at the copy stream's tail by a step
    double *d_stage = nullptr;
    std::vector<hipEvent_t> stage_ev;
    std::vector<int> stage_free;
    double *stage_lb(int e) { return d_stage + 2 * n * e; }
    // host vectors of n doubles for requests and results that found no
    // pinned entry, kept for reuse (up to 1 GiB): a fresh allocation each
    // time pays page faults on every page, which stalls the submitting thread
    std::vector<std::vector<double>> spare;
    std::vector<double> take_vec() {
        if (spare.empty()) return std::vector<double>(n);
        std::vector<double> v = std::move(spare.back());
        spare.pop_back();
        return v;
    }
    void give_vec(std::vector<double> &&v) {
        if (v.size() == n && (spare.size() + 1) * n * sizeof(double) <= (1ull << 30)) spare.push_back(std::move(v));
    }
    // device buffers for the changes of delta requests
    int *d_dptr = nullptr, *d_dcols = nullptr;
    double *d_dlb = nullptr, *d_dub = nullptr;
    double **d_dl = nullptr, **d_du = nullptr;
    size_t dcap = 0;
    ~Server() {
        if (running) {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv.notify_all();
            worker.join();
        }
        // copies on either stream may still read the pinned entries or
        // write the device buffers freed below
        if (cs) (void)hipStreamSynchronize(cs);
        if (P) (void)hipStreamSynchronize(P->stream);
        if (d_buf) (void)cached_free(d_buf);
        if (d_park) (void)cached_free(d_park);
        if (d_base) (void)cached_free(d_base);
        if (inbox) (void)hipHostFree(inbox);
        if (outbox) (void)hipHostFree(outbox);
        for (void *q : {static_cast<void *>(d_kcols), static_cast<void *>(d_off), static_cast<void *>(d_kl),
                        static_cast<void *>(d_ku)})
            if (q) (void)cached_free(q);
        if (h_off) (void)hipHostFree(h_off);
        for (void *q : {static_cast<void *>(h_ccnt), static_cast<void *>(h_ccols), static_cast<void *>(h_cl),
                        static_cast<void *>(h_cu)})
            if (q) (void)hipHostFree(q);
        for (void *q : {static_cast<void *>(d_ccnt), static_cast<void *>(d_ccols), static_cast<void *>(d_cl),
                        static_cast<void *>(d_cu)})
            if (q) (void)cached_free(q);
        for (void *q : {static_cast<void *>(d_dptr), static_cast<void *>(d_dcols), static_cast<void *>(d_dlb),
                        static_cast<void *>(d_dub), static_cast<void *>(d_dl), static_cast<void *>(d_du)})
            if (q) (void)cached_free(q);
        for (void *q : {static_cast<void *>(d_orig), static_cast<void *>(d_blp), static_cast<void *>(d_bup),
                        static_cast<void *>(d_cp)})
            if (q) (void)cached_free(q);
        for (hipEvent_t e : stage_ev) (void)hipEventDestroy(e);
        if (cs_tail) (void)hipEventDestroy(cs_tail);
        if (d_stage) (void)cached_free(d_stage);
        if (cs) (void)hipStreamDestroy(cs);
        if (attached) {
            std::lock_guard<std::mutex> lock(P->run_mutex);
            P->servers--;
        }
    }
    double *orig_lb(int k) { return d_orig + 2 * n * k; }
    double *orig_ub(int k) { return d_orig + 2 * n * k + n; }
    double *in_lb(int e) { return inbox + 2 * n * e; }
    double *in_ub(int e) { return inbox + 2 * n * e + n; }
    double *out_lb(int e) { return outbox + 2 * n * e; }
    double *out_ub(int e) { return outbox + 2 * n * e + n; }
    double *park_lb(int e) { return d_park + 2 * n * e; }
    double *park_ub(int e) { return d_park + 2 * n * e + n; }
};

// Writes zeros to every page of a fresh host allocation, split between up to
// 8 threads.
void prefault(void *p, size_t bytes) {
    const int k = std::max(1, std::min(8, static_cast<int>(std::thread::hardware_concurrency())));
    const size_t part = (bytes + k - 1) / k;
    std::vector<std::thread> th;
    int done_by_caller = 0;  // parts this thread takes when no more threads start
    try {
        for (int t = 1; t < k; t++) {
            const size_t off = std::min(bytes, part * t), len = std::min(part, bytes - off);
            th.emplace_back([p, off, len] { std::memset(static_cast<char *>(p) + off, 0, len); });
        }
    } catch (...) {
        done_by_caller = k - 1 - static_cast<int>(th.size());
    }
    for (int t = k - done_by_caller; t < k; t++) {
        const size_t off = std::min(bytes, part * t);
        std::memset(static_cast<char *>(p) + off, 0, std::min(part, bytes - off));
    }
    std::memset(p, 0, std::min(part, bytes));
    for (auto &x : th) x.join();
}

Server *server_create(Problem *P, int64_t hbm_budget, int max_slots) {
    if (!P->replicas) throw std::runtime_error("serving needs a single matrix");
    if (max_slots < 1) throw std::runtime_error("max_slots must be positive");
    // state, bound vectors, inputs of in-place requests, device staging,
    // result parking
    const int64_t per_slot = P->stats.unit_state_bytes + 64LL * P->n;
    int slots = static_cast<int>(std::min<int64_t>(max_slots, hbm_budget / std::max<int64_t>(1, per_slot)));
    if (slots < 1) throw std::runtime_error("HBM budget below one node's state");
    auto *S = new Server();
    try {
        S->P = P;
        S->slots = slots;
        S->queue_cap = static_cast<size_t>(slots);
        S->n = static_cast<size_t>(P->n);
        check(cached_malloc(&S->d_buf, sizeof(double) * 2 * S->n * slots), "hipMalloc slots");
        check(cached_malloc(&S->d_base, sizeof(double) * 2 * S->n), "hipMalloc base");
        check(cached_malloc(&S->d_park, sizeof(double) * 2 * S->n * slots), "hipMalloc parking");
        for (int e = slots - 1; e >= 0; e--) S->park_free.push_back(e);
        for (int k = 0; k < slots; k++) {
            S->slot_lb.push_back(S->d_buf + 2 * k * S->n);
            S->slot_ub.push_back(S->d_buf + (2 * k + 1) * S->n);
        }
        S->slot_ticket.assign(slots, -1);
        S->slot_delta.assign(slots, 0);
        S->slot_out_lb.assign(slots, nullptr);
        S->slot_out_ub.assign(slots, nullptr);
        S->slot_inplace.assign(slots, 0);
        check(cached_malloc(&S->d_orig, sizeof(double) * 2 * S->n * slots), "hipMalloc inputs");
        check(cached_malloc(&S->d_blp, sizeof(double *) * slots), "hipMalloc");
        check(cached_malloc(&S->d_bup, sizeof(double *) * slots), "hipMalloc");
        check(cached_malloc(&S->d_cp, sizeof(double *) * 4 * slots), "hipMalloc");
        // changed-column regions for finished delta requests (kDiffCap each)
        S->diff_slots = slots;
        const size_t dc = static_cast<size_t>(slots) * Server::kDiffCap;
        check(cached_malloc(&S->d_ccnt, sizeof(int) * slots), "hipMalloc");
        check(cached_malloc(&S->d_ccols, sizeof(int) * dc), "hipMalloc");
        check(cached_malloc(&S->d_cl, sizeof(double) * dc), "hipMalloc");
        check(cached_malloc(&S->d_cu, sizeof(double) * dc), "hipMalloc");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->h_ccnt), sizeof(int) * slots, 0), "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->h_ccols), sizeof(int) * dc, 0), "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->h_cl), sizeof(double) * dc, 0), "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->h_cu), sizeof(double) * dc, 0), "hipHostMalloc");
        check(cached_malloc(&S->d_kcols, sizeof(int) * dc), "hipMalloc");
        check(cached_malloc(&S->d_kl, sizeof(double) * dc), "hipMalloc");
        check(cached_malloc(&S->d_ku, sizeof(double) * dc), "hipMalloc");
        check(cached_malloc(&S->d_off, sizeof(int) * (slots + 1)), "hipMalloc");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->h_off), sizeof(int) * (slots + 1), 0), "hipHostMalloc");
        // pinned staging for request and result vectors (up to 1 GiB each)
        const int entries = static_cast<int>(
            std::max<int64_t>(1, std::min<int64_t>(slots, (1LL << 30) / static_cast<int64_t>(16 * S->n))));
        check(hipHostMalloc(reinterpret_cast<void **>(&S->inbox), sizeof(double) * 2 * S->n * entries, 0),
              "hipHostMalloc inbox");
        check(hipHostMalloc(reinterpret_cast<void **>(&S->outbox), sizeof(double) * 2 * S->n * entries, 0),
              "hipHostMalloc outbox");
        for (int e = entries - 1; e >= 0; e--) {
            S->inbox_free.push_back(e);
            S->outbox_free.push_back(e);
        }
        // The host maps pinned pages on their first access, about 0.4 s for
        // these 2 x 1 GiB at serving sizes; taken here, not by the first
        // burst of requests (it stalled their submitters and collectors).
        prefault(S->inbox, sizeof(double) * 2 * S->n * entries);
        prefault(S->outbox, sizeof(double) * 2 * S->n * entries);
        // device staging for prefetched request vectors: one per slot, up to
        // 4 GiB (a small share of the device), so that requests from pinned
        // memory cross PCIe on the copy stream while steps run
        const int n_stage = static_cast<int>(
            std::max<int64_t>(1, std::min<int64_t>(slots, (4LL << 30) / static_cast<int64_t>(16 * S->n))));
        check(hipStreamCreateWithFlags(&S->cs, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreateWithFlags(&S->cs_tail, hipEventDisableTiming), "hipEventCreate");
        check(cached_malloc(&S->d_stage, sizeof(double) * 2 * S->n * n_stage), "hipMalloc staging");
        for (int e = n_stage - 1; e >= 0; e--) {
            hipEvent_t ev;
            check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
            S->stage_ev.push_back(ev);
            S->stage_free.push_back(e);
        }
        std::reverse(S->stage_ev.begin(), S->stage_ev.end());
        // the problem's unit slots become the server's until it is destroyed
        std::lock_guard<std::mutex> lock(P->run_mutex);
        check_no_server(P);
        if (P->capacity != slots) alloc_units(P, slots);
        P->servers++;
        S->attached = true;
        // fresh status ring for the server's steps
        hipLaunchKernelGGL(unit_reset_kernel, dim3(1), dim3(256), 0, P->stream, P->D, 0, P->d_params, 0, 1);
        check(hipStreamSynchronize(P->stream), "server init");
    } catch (...) {
        delete S;
        throw;
    }
    return S;
}

void server_destroy(Server *s) {
    // GDP_TRACE_SERVER: where results that found no free pinned outbox entry went
    if (std::getenv("GDP_TRACE_SERVER"))
        std::fprintf(stderr,
                     "server: %d steps, results retired into device parking %lld, into host vectors %lld; slowest "
                     "step %d: %.2f ms (admission %.2f of which copies %.2f, step %.2f, retirement %.2f)\n",
                     s->step, s->parked_results, s->pageable_results, s->slowest_step, s->slowest_ms,
                     s->slowest_admit_ms, s->slowest_copies_ms, s->slowest_run_ms,
                     s->slowest_ms - s->slowest_admit_ms - s->slowest_run_ms);
    delete s;
}

int server_slots(const Server *s) { return s->slots; }

void server_set_base(Server *S, const double *lbs, const double *ubs) {
    // between steps: a step reads the device base while it admits requests
    std::lock_guard<std::mutex> lock(S->P->run_mutex);
    check(hipMemcpy(S->d_base, lbs, sizeof(double) * S->n, hipMemcpyHostToDevice), "H2D base");
    check(hipMemcpy(S->d_base + S->n, ubs, sizeof(double) * S->n, hipMemcpyHostToDevice), "H2D base");
    S->base_lb.assign(lbs, lbs + S->n);
    S->base_ub.assign(ubs, ubs + S->n);
    S->has_base = true;
}

namespace {
// With the server's lock held: once the stepping thread has failed, no
// request it would have to run is accepted.
void check_worker(const Server *S) {
    if (!S->worker_error.empty()) throw std::runtime_error("stepping thread: " + S->worker_error);
}

// With the server's lock held (lk): while the stepping thread runs, waits
// until the queue has room and, for a copied request (inbox), a pinned
// inbox entry is free.  Bounded admission: an overloaded server holds its
// submitters back instead of queueing without limit.  Returns at once when
// the thread is not running (a caller that steps by hand must not wait for
// itself) or has stopped or failed.
void wait_for_room(Server *S, std::unique_lock<std::mutex> &lk, bool inbox) {
    S->space_cv.wait(lk, [S, inbox] {
        return !S->running || S->stop || !S->worker_error.empty() ||
               (S->queue.size() < S->queue_cap && (!inbox || !S->inbox_free.empty()));
    });
}

// A free device staging entry (-1 if none); with the server's lock held.
int take_stage(Server *S) {
    if (S->stage_free.empty()) return -1;
    const int e = S->stage_free.back();
    S->stage_free.pop_back();
    return e;
}

// Starts copying a request's vectors (pinned host memory) into device
// staging entry e on the copy stream.
void prefetch(Server *S, int e, const double *lb, const double *ub) {
    check(hipMemcpyAsync(S->stage_lb(e), lb, S->n * sizeof(double), hipMemcpyHostToDevice, S->cs), "H2D");
    check(hipMemcpyAsync(S->stage_lb(e) + S->n, ub, S->n * sizeof(double), hipMemcpyHostToDevice, S->cs), "H2D");
    check(hipEventRecord(S->stage_ev[e], S->cs), "event");
}

// Queues a prepared request and wakes the stepping thread.
int64_t enqueue(Server *S, Server::Request &&q) {
    const int64_t t = q.ticket;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->queue.push_back(std::move(q));
    }
    S->cv.notify_one();
    return t;
}
}  // namespace

int64_t server_submit(Server *S, const double *lbs, const double *ubs, const int *hints, int n_hints,
                      int max_rounds) {
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    Server::Request q;
    if (hints && n_hints > 0) {
        for (int k = 0; k < n_hints; k++)
            if (hints[k] < 0 || hints[k] >= static_cast<int>(S->n)) throw std::runtime_error("hint column out of range");
        q.cols.assign(hints, hints + n_hints);
    }
    q.max_rounds = max_rounds;
    int stage = -1;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        check_worker(S);
        // while the stepping thread runs, wait for room instead of holding the
        // vectors in pageable memory, whose copies would stall the steps
        wait_for_room(S, lk, true);
        check_worker(S);
        q.ticket = S->next_ticket++;
        if (!S->inbox_free.empty()) {
            q.inbox = S->inbox_free.back();
            S->inbox_free.pop_back();
            stage = take_stage(S);
        } else {
            q.lb = S->take_vec();
            q.ub = S->take_vec();
        }
    }
    // the entries taken above belong to this request: filled without the lock
    if (q.inbox >= 0) {
        std::memcpy(S->in_lb(q.inbox), lbs, sizeof(double) * S->n);
        std::memcpy(S->in_ub(q.inbox), ubs, sizeof(double) * S->n);
        if (stage >= 0) {
            prefetch(S, stage, S->in_lb(q.inbox), S->in_ub(q.inbox));
            q.stage = stage;
        }
    } else {
        std::memcpy(q.lb.data(), lbs, sizeof(double) * S->n);
        std::memcpy(q.ub.data(), ubs, sizeof(double) * S->n);
    }
    return enqueue(S, std::move(q));
}

int64_t server_submit_pinned(Server *S, const double *lbs, const double *ubs, double *out_lbs, double *out_ubs,
                             const int *hints, int n_hints, int max_rounds) {
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    if (!lbs || !ubs || !out_lbs || !out_ubs) throw std::runtime_error("null buffer");
    Server::Request q;
    q.pin_lb = lbs;
    q.pin_ub = ubs;
    q.out_lb = out_lbs;
    q.out_ub = out_ubs;
    if (hints && n_hints > 0) {
        for (int k = 0; k < n_hints; k++)
            if (hints[k] < 0 || hints[k] >= static_cast<int>(S->n)) throw std::runtime_error("hint column out of range");
        q.cols.assign(hints, hints + n_hints);
    }
    q.max_rounds = max_rounds;
    int stage;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        check_worker(S);
        wait_for_room(S, lk, false);
        check_worker(S);
        q.ticket = S->next_ticket++;
        stage = take_stage(S);
    }
    if (stage >= 0) {
        prefetch(S, stage, lbs, ubs);
        q.stage = stage;
    }
    return enqueue(S, std::move(q));
}

int64_t server_submit_delta(Server *S, int n_changes, const int *cols, const double *lbs, const double *ubs,
                            int max_rounds) {
    if (!S->has_base) throw std::runtime_error("no base vector set");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    if (n_changes < 0 || (n_changes > 0 && (!cols || !lbs || !ubs))) throw std::runtime_error("bad changes");
    Server::Request q;
    q.delta = true;
    for (int k = 0; k < n_changes; k++)
        if (cols[k] < 0 || cols[k] >= static_cast<int>(S->n)) throw std::runtime_error("changed column out of range");
    q.cols.assign(cols, cols + n_changes);
    q.dlb.assign(lbs, lbs + n_changes);
    q.dub.assign(ubs, ubs + n_changes);
    q.max_rounds = max_rounds;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        check_worker(S);
        wait_for_room(S, lk, false);
        check_worker(S);
        q.ticket = S->next_ticket++;
    }
    return enqueue(S, std::move(q));
}

int server_pending(Server *S) {
    std::lock_guard<std::mutex> lk(S->mu);
    return S->busy + static_cast<int>(S->queue.size());
}

const std::vector<int64_t> &server_completed(const Server *S) { return S->last_completed; }

int server_take_completed(Server *S, int64_t *tickets, int cap) {
    std::lock_guard<std::mutex> lk(S->mu);
    int k = 0;
    while (k < cap && !S->completed_acc.empty()) {
        tickets[k++] = S->completed_acc.front();
        S->completed_acc.pop_front();
    }
    // the requests it finished are handed out first; then its failure
    if (k == 0) check_worker(S);
    return k;
}

bool server_running(Server *S) {
    std::lock_guard<std::mutex> lk(S->mu);
    return S->running;
}

namespace {

template <typename T>
void grow(T *&p, size_t count) {
    if (p) (void)cached_free(p);
    p = nullptr;
    check(cached_malloc(&p, sizeof(T) * std::max<size_t>(1, count)), "hipMalloc");
}

}  // namespace

namespace {
int step_impl(Server *S);

// Copies bound-vector pairs (2n doubles each, lb then ub) from src[k] to
// dst[k] on the problem's stream in one launch; the pointer table goes to
// d_cp from entry `at`.
void copy_pairs(Server *S, const std::vector<double *> &src, const std::vector<double *> &dst, int at) {
    const int count = static_cast<int>(src.size());
    if (!count) return;
    std::vector<double *> table(src);
    table.insert(table.end(), dst.begin(), dst.end());
    hipStream_t s = S->P->stream;
    upload(S->d_cp + at, table.data(), table.size(), s);
    const long long n2 = 2LL * static_cast<long long>(S->n);
    const int gx = static_cast<int>(std::max(1LL, std::min(256LL, (n2 / 2 + 255) / 256)));
    hipLaunchKernelGGL(copy_pairs_kernel, dim3(gx, std::min(count, 1024)), dim3(256), 0, s, n2, count, S->d_cp + at);
    check(hipGetLastError(), "copy launch");
}
}  // namespace

int server_step(Server *S) {
    if (server_running(S)) throw std::runtime_error("the server steps on its own thread");
    return step_impl(S);
}

namespace {
// The stepping thread: a step whenever requests are queued or running.
void worker_loop(Server *S) {
    try {
        check(hipSetDevice(S->P->device), "hipSetDevice");
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(S->mu);
                S->cv.wait(lk, [S] { return S->stop || !S->queue.empty() || S->busy > 0; });
                if (S->stop) return;
            }
            step_impl(S);
        }
    } catch (const std::exception &e) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->worker_error = e.what();
            S->stop = true;
        }
        S->space_cv.notify_all();
    }
}
}  // namespace

void server_start(Server *S) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->running) throw std::runtime_error("the server already runs its stepping thread");
    S->stop = false;
    S->worker_error.clear();
    S->completed_acc.clear();
    S->running = true;
    S->worker = std::thread(worker_loop, S);
}

void server_stop(Server *S) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->running) return;
        S->stop = true;
    }
    S->cv.notify_all();
    S->space_cv.notify_all();
    S->worker.join();
    std::string err;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->running = false;
        err = S->worker_error;
    }
    if (!err.empty()) throw std::runtime_error("stepping thread: " + err);
}

namespace {
int step_impl(Server *S) {
    using Clock = std::chrono::steady_clock;
    const Clock::time_point t_begin = Clock::now();
    Clock::time_point t_run = t_begin, t_retire = t_begin, t_copies = t_begin;
    // GDP_TRACE_SERVER: the slowest step and where its time went
    struct StepClock {
        Server *S;
        const Clock::time_point &b, &r, &t, &c;
        ~StepClock() {
            if (!S->trace) return;
            const Clock::time_point e = Clock::now();
            const double ms = std::chrono::duration<double, std::milli>(e - b).count();
            if (ms > S->slowest_ms) {
                S->slowest_ms = ms;
                S->slowest_step = S->step;
                S->slowest_admit_ms = std::chrono::duration<double, std::milli>(r - b).count();
                S->slowest_copies_ms = std::chrono::duration<double, std::milli>(c - b).count();
                S->slowest_run_ms = std::chrono::duration<double, std::milli>(t - r).count();
            }
        }
    } step_clock{S, t_begin, t_run, t_retire, t_copies};
    Problem *P = S->P;
    std::lock_guard<std::mutex> lock(P->run_mutex);
    S->last_completed.clear();
    hipStream_t s = P->stream;
    const size_t n = S->n;
    // ---- admit queued requests into free slots; they start in this step ----
    std::vector<int> adm, caps, hint_ptr(1, 0), hints, used_inbox, used_stage;
    std::vector<double *> adm_lb, adm_ub;
    std::vector<double *> del_lb, del_ub;
    std::vector<int> dptr(1, 0), dcols;
    std::vector<double> dlv, duv;
    std::vector<std::vector<double>> keep;  // pageable request vectors in flight
    std::vector<std::pair<int, Server::Request>> taken;  // (slot, request) admitted in this step
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (int k = 0; k < S->slots && !S->queue.empty(); k++) {
            if (S->slot_ticket[k] >= 0) continue;
            S->slot_ticket[k] = S->queue.front().ticket;
            S->busy++;
            taken.emplace_back(k, std::move(S->queue.front()));
            S->queue.pop_front();
        }
    }
    if (!taken.empty()) S->space_cv.notify_all();  // room in the queue
    // Device-side fills of the admitted slots go into two batched copies:
    // (1) the base vector into delta requests' slots and prefetched staging
    // entries into their slots, (2) in-place requests' inputs, kept for the
    // diff at retirement, once every slot is filled.
    std::vector<double *> cp1_src, cp1_dst, cp2_src, cp2_dst;
    bool staged = false;
    for (auto &kq : taken) {
        const int k = kq.first;
        Server::Request &q = kq.second;
        const bool in_place = q.pin_lb && q.out_lb == q.pin_lb && q.out_ub == q.pin_ub;
        if (q.delta) {
            cp1_src.push_back(S->d_base);
            cp1_dst.push_back(S->slot_lb[k]);
            del_lb.push_back(S->slot_lb[k]);
            del_ub.push_back(S->slot_ub[k]);
            dcols.insert(dcols.end(), q.cols.begin(), q.cols.end());
            dlv.insert(dlv.end(), q.dlb.begin(), q.dlb.end());
            duv.insert(duv.end(), q.dub.begin(), q.dub.end());
            dptr.push_back(static_cast<int>(dcols.size()));
        } else if (q.stage >= 0) {
            // prefetched: the slot's (lb, ub) pair is contiguous, like the entry's
            cp1_src.push_back(S->stage_lb(q.stage));
            cp1_dst.push_back(S->slot_lb[k]);
            staged = true;
            used_stage.push_back(q.stage);
            if (q.inbox >= 0) used_inbox.push_back(q.inbox);
        } else if (q.pin_lb) {
            check(hipMemcpyAsync(S->slot_lb[k], q.pin_lb, n * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
            check(hipMemcpyAsync(S->slot_ub[k], q.pin_ub, n * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
        } else if (q.inbox >= 0) {
            check(hipMemcpyAsync(S->slot_lb[k], S->in_lb(q.inbox), n * sizeof(double), hipMemcpyHostToDevice, s),
                  "H2D");
            check(hipMemcpyAsync(S->slot_ub[k], S->in_ub(q.inbox), n * sizeof(double), hipMemcpyHostToDevice, s),
                  "H2D");
            used_inbox.push_back(q.inbox);
        } else {
            check(hipMemcpyAsync(S->slot_lb[k], q.lb.data(), n * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
            check(hipMemcpyAsync(S->slot_ub[k], q.ub.data(), n * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
            keep.push_back(std::move(q.lb));
            keep.push_back(std::move(q.ub));
        }
        if (in_place) {
            cp2_src.push_back(S->slot_lb[k]);
            cp2_dst.push_back(S->orig_lb(k));
        }
        adm.push_back(k);
        adm_lb.push_back(S->slot_lb[k]);
        adm_ub.push_back(S->slot_ub[k]);
        hints.insert(hints.end(), q.cols.begin(), q.cols.end());
        hint_ptr.push_back(static_cast<int>(hints.size()));
        caps.push_back(q.max_rounds);
        S->slot_delta[k] = q.delta ? 1 : 0;
        S->slot_out_lb[k] = q.out_lb;
        S->slot_out_ub[k] = q.out_ub;
        S->slot_inplace[k] = in_place ? 1 : 0;
    }
    // every prefetch of an admitted request was issued on the copy stream
    // before the request was queued: one event at the copy stream's tail
    if (staged) {
        check(hipEventRecord(S->cs_tail, S->cs), "event");
        check(hipStreamWaitEvent(s, S->cs_tail, 0), "wait");
    }
    copy_pairs(S, cp1_src, cp1_dst, 0);
    if (!del_lb.empty()) {
        const size_t nd = del_lb.size(), nc = dcols.size();
        if (nc > S->dcap || !S->d_dl) {
            S->dcap = std::max<size_t>(nc, 4096);
            grow(S->d_dcols, S->dcap);
            grow(S->d_dlb, S->dcap);
            grow(S->d_dub, S->dcap);
            grow(S->d_dptr, static_cast<size_t>(S->slots) + 1);
            grow(S->d_dl, static_cast<size_t>(S->slots));
            grow(S->d_du, static_cast<size_t>(S->slots));
        }
        upload(S->d_dptr, dptr.data(), nd + 1, s);
        upload(S->d_dcols, dcols.data(), nc, s);
        upload(S->d_dlb, dlv.data(), nc, s);
        upload(S->d_dub, duv.data(), nc, s);
        upload(S->d_dl, del_lb.data(), nd, s);
        upload(S->d_du, del_ub.data(), nd, s);
        hipLaunchKernelGGL(delta_kernel, dim3(1, std::max(1, std::min(static_cast<int>(nd), 1024))), dim3(256), 0, s,
                           static_cast<int>(nd), S->d_dl, S->d_du, S->d_dptr, S->d_dcols, S->d_dlb, S->d_dub);
        check(hipGetLastError(), "delta launch");
    }
    copy_pairs(S, cp2_src, cp2_dst, 2 * S->slots);
    t_copies = Clock::now();
    if (!adm.empty())
        load_units(P, static_cast<int>(adm.size()), adm.data(), adm_lb.data(), adm_ub.data(), hint_ptr.data(),
                   hints.data(), caps.data(), S->step, false);
    t_run = Clock::now();
    t_retire = t_run;
    // ---- every active slot advances one round ----
    std::vector<int> active;
    for (int k = 0; k < S->slots; k++)
        if (S->slot_ticket[k] >= 0) active.push_back(k);
    if (active.empty()) return 0;
    upload(P->d_list, active.data(), active.size(), s);
    P->D.n_active = static_cast<int>(active.size());
    launch_step(P, S->step, 0, S->step);
    check(hipGetLastError(), "step launch");
    std::vector<UnitState> st = read_units(P, static_cast<int>(active.size()), active.data());  // synchronizes
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (int e : used_inbox) S->inbox_free.push_back(e);
        for (int e : used_stage) S->stage_free.push_back(e);
        for (auto &v : keep) S->give_vec(std::move(v));
    }
    if (!used_inbox.empty()) S->space_cv.notify_all();
    keep.clear();
    int err = 0;
    check(hipMemcpy(&err, P->D.global + kGlobalError, sizeof(int), hipMemcpyDeviceToHost), "D2H");
    if (err) throw std::runtime_error("device wait timed out");
    t_retire = Clock::now();
    // ---- retire finished slots: unpack and copy the results out ----
    std::vector<int> fin;
    std::vector<RunResult> fin_r;
    for (size_t k = 0; k < active.size(); k++) {
        if (active_after(st[k], S->step)) continue;
        RunResult r;
        if (st[k].envelope) r.rounds = -1;
        else r = outcome(st[k]);
        fin.push_back(active[k]);
        fin_r.push_back(r);
    }
    if (!fin.empty()) {
        std::vector<double *> lbp, ubp;
        for (int k : fin) {
            lbp.push_back(S->slot_lb[k]);
            ubp.push_back(S->slot_ub[k]);
        }
        upload(P->d_list, fin.data(), fin.size(), s);
        upload(P->d_lbp, lbp.data(), fin.size(), s);
        upload(P->d_ubp, ubp.data(), fin.size(), s);
        const int gx = std::max(1, std::min(1024, (P->M.n + 255) / 256));
        hipLaunchKernelGGL(unpack_kernel, dim3(gx, std::max(1, std::min(static_cast<int>(fin.size()), 1024))),
                           dim3(256), 0, s, P->M, P->D, static_cast<int>(fin.size()), P->d_list, P->d_lbp, P->d_ubp);
        check(hipGetLastError(), "unpack launch");
        std::vector<Server::Done> out(fin.size());
        // delta requests (against the base vector) and in-place zero-copy
        // requests (against their own input): the columns that differ, found
        // on the device, so that only they cross PCIe
        std::vector<int> dq;  // indices into fin of diffed requests
        std::vector<double *> dl, du;
        std::vector<const double *> bl, bu;
        for (size_t q = 0; q < fin.size(); q++)
            if (S->slot_delta[fin[q]] || S->slot_inplace[fin[q]]) {
                const int k = fin[q];
                dq.push_back(static_cast<int>(q));
                dl.push_back(S->slot_lb[k]);
                du.push_back(S->slot_ub[k]);
                bl.push_back(S->slot_delta[k] ? S->d_base : S->orig_lb(k));
                bu.push_back(S->slot_delta[k] ? S->d_base + n : S->orig_ub(k));
            }
        if (!dq.empty()) {
            const int nd = static_cast<int>(dq.size());
            upload(P->d_lbp, dl.data(), nd, s);
            upload(P->d_ubp, du.data(), nd, s);
            upload(S->d_blp, const_cast<double **>(bl.data()), nd, s);
            upload(S->d_bup, const_cast<double **>(bu.data()), nd, s);
            check(hipMemsetAsync(S->d_ccnt, 0, sizeof(int) * nd, s), "memset");
            hipLaunchKernelGGL(diff_kernel, dim3(gx, std::max(1, std::min(nd, 1024))), dim3(256), 0, s,
                               static_cast<int>(n), nd, P->d_lbp, P->d_ubp, S->d_blp, S->d_bup,
                               Server::kDiffCap, S->d_ccnt, S->d_ccols, S->d_cl, S->d_cu);
            check(hipGetLastError(), "diff launch");
            // packed on the device, so that only the changes cross PCIe
            hipLaunchKernelGGL(compact_changes_kernel, dim3(std::max(1, std::min(nd, 1024))), dim3(256), 0, s, nd,
                               Server::kDiffCap, S->d_ccnt, S->d_ccols, S->d_cl, S->d_cu, S->d_off, S->d_kcols,
                               S->d_kl, S->d_ku);
            check(hipGetLastError(), "compact launch");
            check(hipMemcpyAsync(S->h_ccnt, S->d_ccnt, sizeof(int) * nd, hipMemcpyDeviceToHost, s), "D2H");
            check(hipMemcpyAsync(S->h_off, S->d_off, sizeof(int) * (nd + 1), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "diff sync");
            const size_t total = static_cast<size_t>(S->h_off[nd]);
            if (total > 0) {
                check(hipMemcpyAsync(S->h_ccols, S->d_kcols, sizeof(int) * total, hipMemcpyDeviceToHost, s), "D2H");
                check(hipMemcpyAsync(S->h_cl, S->d_kl, sizeof(double) * total, hipMemcpyDeviceToHost, s), "D2H");
                check(hipMemcpyAsync(S->h_cu, S->d_ku, sizeof(double) * total, hipMemcpyDeviceToHost, s), "D2H");
                check(hipStreamSynchronize(s), "diff sync");
            }
            for (int j = 0; j < nd; j++) {
                const int c = S->h_ccnt[j];
                if (c > Server::kDiffCap) continue;  // too many: the full vector below
                const int k = fin[dq[j]];
                Server::Done &d = out[dq[j]];
                const size_t o = static_cast<size_t>(S->h_off[j]);
                if (S->slot_inplace[k]) {
                    // the caller's buffers hold the input: write the changes
                    for (int q = 0; q < c; q++) {
                        S->slot_out_lb[k][S->h_ccols[o + q]] = S->h_cl[o + q];
                        S->slot_out_ub[k][S->h_ccols[o + q]] = S->h_cu[o + q];
                    }
                    d.in_place = true;
                    continue;
                }
                d.changes = true;
                d.cols.assign(S->h_ccols + o, S->h_ccols + o + c);
                d.clb.assign(S->h_cl + o, S->h_cl + o + c);
                d.cub.assign(S->h_cu + o, S->h_cu + o + c);
            }
        }
        for (size_t q = 0; q < fin.size(); q++) {
            const int k = fin[q];
            Server::Done &d = out[q];
            d.r = fin_r[q];
            if (d.changes || d.in_place) continue;
            double *hl, *hu;
            if (S->slot_out_lb[k]) {
                d.in_place = true;
                hl = S->slot_out_lb[k];
                hu = S->slot_out_ub[k];
            } else {
                std::lock_guard<std::mutex> lk(S->mu);
                if (!S->outbox_free.empty()) {
                    d.outbox = S->outbox_free.back();
                    S->outbox_free.pop_back();
                    hl = S->out_lb(d.outbox);
                    hu = S->out_ub(d.outbox);
                } else if (!S->park_free.empty()) {
                    d.park = S->park_free.back();
                    S->park_free.pop_back();
                    S->parked_results++;
                    hl = S->park_lb(d.park);
                    hu = S->park_ub(d.park);
                } else {
                    d.lb = S->take_vec();
                    d.ub = S->take_vec();
                    S->pageable_results++;
                    hl = d.lb.data();
                    hu = d.ub.data();
                }
            }
            const hipMemcpyKind kind = d.park >= 0 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
            check(hipMemcpyAsync(hl, S->slot_lb[k], n * sizeof(double), kind, s), "result copy");
            check(hipMemcpyAsync(hu, S->slot_ub[k], n * sizeof(double), kind, s), "result copy");
        }
        check(hipStreamSynchronize(s), "retire sync");
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t q = 0; q < fin.size(); q++) {
            const int k = fin[q];
            S->done[S->slot_ticket[k]] = std::move(out[q]);
            S->last_completed.push_back(S->slot_ticket[k]);
            if (S->running) S->completed_acc.push_back(S->slot_ticket[k]);
            S->slot_ticket[k] = -1;
            S->busy--;
            S->slot_out_lb[k] = nullptr;
            S->slot_out_ub[k] = nullptr;
            S->slot_inplace[k] = 0;
        }
    }
    S->step++;
    return static_cast<int>(fin.size());
}
}  // namespace

namespace {
// Takes a finished request out of the server (false if not finished).
bool take_done(Server *S, int64_t ticket, Server::Done &d) {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->done.find(ticket);
    if (it == S->done.end()) return false;
    d = std::move(it->second);
    S->done.erase(it);
    return true;
}

// Returns a finished request's host buffers to the server.
void release_done(Server *S, Server::Done &d) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (d.outbox >= 0) S->outbox_free.push_back(d.outbox);
    if (d.park >= 0) S->park_free.push_back(d.park);
    S->give_vec(std::move(d.lb));
    S->give_vec(std::move(d.ub));
}

// Copies a parked result to the caller's host arrays on the server's copy
// stream (the steps run on the problem's stream meanwhile), each vector
// through the library's pinned staging: a DMA into a staging buffer, then
// the copy threads into the caller's memory.  On failure the entry is
// returned to the server before the error is passed on.
void download_parked(Server *S, Server::Done &d, double *lbs, double *ubs) {
    try {
        download_bulk(lbs, S->park_lb(d.park), sizeof(double) * S->n, S->cs);
        download_bulk(ubs, S->park_ub(d.park), sizeof(double) * S->n, S->cs);
    } catch (...) {
        release_done(S, d);
        throw;
    }
}
}  // namespace

bool server_result(Server *S, int64_t ticket, double *lbs, double *ubs, RunResult *r) {
    Server::Done d;
    if (!take_done(S, ticket, d)) return false;
    if (d.in_place || (!lbs && !ubs)) {
        // the bounds are in the caller's pinned buffers (or not wanted)
    } else if (d.changes) {
        std::memcpy(lbs, S->base_lb.data(), sizeof(double) * S->n);
        std::memcpy(ubs, S->base_ub.data(), sizeof(double) * S->n);
        for (size_t q = 0; q < d.cols.size(); q++) {
            lbs[d.cols[q]] = d.clb[q];
            ubs[d.cols[q]] = d.cub[q];
        }
    } else if (d.outbox >= 0) {
        std::memcpy(lbs, S->out_lb(d.outbox), sizeof(double) * S->n);
        std::memcpy(ubs, S->out_ub(d.outbox), sizeof(double) * S->n);
    } else if (d.park >= 0) {
        download_parked(S, d, lbs, ubs);
    } else {
        std::memcpy(lbs, d.lb.data(), sizeof(double) * S->n);
        std::memcpy(ubs, d.ub.data(), sizeof(double) * S->n);
    }
    *r = d.r;
    release_done(S, d);
    return true;
}

int server_result_changes(Server *S, int64_t ticket, std::vector<int> &cols, std::vector<double> &lbs,
                          std::vector<double> &ubs, RunResult *r) {
    if (!S->has_base) throw std::runtime_error("no base vector set");
    Server::Done d;
    if (!take_done(S, ticket, d)) return 0;
    cols.clear();
    lbs.clear();
    ubs.clear();
    if (d.changes) {
        cols = std::move(d.cols);
        lbs = std::move(d.clb);
        ubs = std::move(d.cub);
    } else {
        if (d.park >= 0) {
            d.lb = S->take_vec();
            d.ub = S->take_vec();
            download_parked(S, d, d.lb.data(), d.ub.data());
        }
        const double *fl = d.outbox >= 0 ? S->out_lb(d.outbox) : d.lb.data();
        const double *fu = d.outbox >= 0 ? S->out_ub(d.outbox) : d.ub.data();
        for (size_t j = 0; j < S->n; j++)
            if (std::memcmp(&fl[j], &S->base_lb[j], sizeof(double)) || std::memcmp(&fu[j], &S->base_ub[j], sizeof(double))) {
                cols.push_back(static_cast<int>(j));
                lbs.push_back(fl[j]);
                ubs.push_back(fu[j]);
            }
    }
    *r = d.r;
    release_done(S, d);
    return 1;
}

}  // namespace gdp
