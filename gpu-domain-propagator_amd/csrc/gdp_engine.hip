// gdp_engine.hip - host driver of the MI355X propagation engine.
//
// Reference counterpart: propagateConstraintsGPUAtomic<T>
// (src/propagators/GPU_propagator.cuh:98-268) and the row blocking of
// fill_row_blocks (src/kernels/util_kernels.cu:12-62).  Differences by
// design: the matrix is laid out once per Problem (rows grouped by tree
// width, columns relabelled by bound storage class, integrality folded into
// the column index), rounds are round-synchronous (sweep, apply of split rows, publish), and
// GPU_LOOP is a captured HIP graph of round launches that stop themselves on
// the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "gdp_arith.h"
#include "gdp_device.h"
#include "gdp_engine.h"

namespace gdp {

__global__ void sweep_kernel(DevMatrix M, DevState D, int round, int guard);
__global__ void apply_split_kernel(DevMatrix M, DevState D, int round, int guard);
__global__ void publish_kernel(DevMatrix M, DevState D, int round, int guard);
__global__ void mega_kernel(DevMatrix M, DevState D, int max_rounds);
__global__ void init_kernel(DevMatrix M, DevState D, const double *lbs, const double *ubs);
__global__ void clear_envelope_kernel(DevState D);
__global__ void unpack_kernel(DevMatrix M, DevState D, double *lbs, double *ubs);

namespace {

thread_local std::string g_last_error;

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int next_pow2_log(int len) {
    int lg = 0;
    while ((1 << lg) < len) lg++;
    return lg;
}

template <typename T>
T *dev_alloc(size_t count, std::vector<void *> &owned) {
    void *p = nullptr;
    if (count == 0) count = 1;
    check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    owned.push_back(p);
    return static_cast<T *>(p);
}

// Storage class of a column from its bound envelope: 0 = 2 x int16,
// 1 = 2 x int32, 2 = 2 x double.
int column_class(int vartype, const double *env_lb, const double *env_ub, int j) {
    if (vartype == kVarContinuous || !env_lb || !env_ub) return 2;
    const double l = env_lb[j], u = env_ub[j];
    if (!(std::isfinite(l) && std::isfinite(u)) || std::floor(l) != l || std::floor(u) != u) return 2;
    if (is_neg_inf(l) || is_pos_inf(u)) return 2;
    if (l >= -32767.0 && u <= 32767.0 && l <= 32767.0 && u >= -32767.0) return 0;
    if (l >= -2147483647.0 && u <= 2147483647.0 && l <= 2147483647.0 && u >= -2147483647.0) return 1;
    return 2;
}

}  // namespace

std::string last_error() { return g_last_error; }
void set_last_error(const std::string &msg) { g_last_error = msg; }

struct Problem {
    int m = 0, n = 0, nnz = 0;
    Tuning tuning;
    ProblemStats stats;
    DevMatrix M{};
    DevState D{};
    std::vector<void *> owned;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int device = 0;
    int cu_count = 0;
    int mega_grid = 0;
    // captured GPU_LOOP graphs, one per chunk of rounds
    std::vector<hipGraphExec_t> graphs;
    std::vector<int> graph_last;
    std::mutex run_mutex;

    ~Problem() {
        for (auto g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        for (void *p : owned) (void)hipFree(p);
    }
};

Problem *problem_create(int m, int n, int nnz, const int *col, const int *row_ptr, const double *val,
                        const double *lhs, const double *rhs, const int *vartype, const double *env_lb,
                        const double *env_ub, const Tuning *tuning) {
    if (m <= 0 || n <= 0 || nnz <= 0) throw std::runtime_error("empty problem");
    if (n > kIdMask) throw std::runtime_error("too many columns");
    if (!col || !row_ptr || !val || !lhs || !rhs || !vartype) throw std::runtime_error("null input array");
    if (row_ptr[0] != 0 || row_ptr[m] != nnz) throw std::runtime_error("row_ptrs do not span nnz");
    for (int i = 0; i < m; i++)
        if (row_ptr[i + 1] < row_ptr[i]) throw std::runtime_error("row_ptrs not monotone");
    for (int k = 0; k < nnz; k++)
        if (col[k] < 0 || col[k] >= n) throw std::runtime_error("column index out of range");

    auto *P = new Problem();
    try {
        P->m = m;
        P->n = n;
        P->nnz = nnz;
        if (tuning) P->tuning = *tuning;
        const Tuning &T = P->tuning;

        // ---- columns: internal ids grouped by bound storage class ----
        std::vector<int> cls(n);
        int ncls[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            cls[j] = T.compact_bounds ? column_class(vartype[j], env_lb, env_ub, j) : 2;
            ncls[cls[j]]++;
        }
        std::vector<int> perm(n), internal(n);
        {
            int next[3] = {0, ncls[0], ncls[0] + ncls[1]};
            for (int j = 0; j < n; j++) {
                const int id = next[cls[j]]++;
                perm[id] = j;
                internal[j] = id;
            }
        }
        P->stats.n16 = ncls[0];
        P->stats.n32 = ncls[1];
        P->stats.nf = ncls[2];

        // ---- rows grouped by tree width: long rows first, then wide to narrow ----
        // bin b < kLongBin holds rows with next_pow2(len) == 2^b; bin kLongBin
        // rows longer than kSegLen.
        constexpr int kLongBin = 10;
        std::vector<std::vector<int>> rows_in_bin(11);
        for (int i = 0; i < m; i++) {
            const int len = row_ptr[i + 1] - row_ptr[i];
            if (len == 0) continue;
            const int b = len > kSegLen ? kLongBin : next_pow2_log(len);
            rows_in_bin[b].push_back(i);
            P->stats.bin_rows[b]++;
            P->stats.bin_nnz[b] += len;
        }
        std::vector<int> order;
        order.reserve(m);
        for (int b = kLongBin; b >= 0; b--)
            for (int i : rows_in_bin[b]) order.push_back(i);
        const int mp = static_cast<int>(order.size());

        // rows padded to multiples of kRowAlign entries (padding: column 0,
        // coefficient 0); the padding count sits in the low bits of h_rp
        std::vector<int> h_len(mp);
        long long padded = 0;
        for (int r = 0; r < mp; r++) {
            const int i = order[r];
            h_len[r] = row_ptr[i + 1] - row_ptr[i];
            padded += (h_len[r] + kRowAlign - 1) / kRowAlign * kRowAlign;
        }
        if (padded > 0x7fffff00LL) throw std::runtime_error("too many nonzeros");
        std::vector<int> h_rp(mp + 1);
        std::vector<double2> h_sides(mp);
        std::vector<int> h_col(padded, 0);
        std::vector<double> h_val(padded, 0.0);
        int pos = 0;
        for (int r = 0; r < mp; r++) {
            const int i = order[r];
            const int pad = (kRowAlign - h_len[r] % kRowAlign) % kRowAlign;
            h_rp[r] = pos | pad;
            h_sides[r] = make_double2(lhs[i], rhs[i]);
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
                const int j = col[k];
                h_col[pos] = internal[j] | (vartype[j] == kVarContinuous ? 0 : kIntFlag);
                h_val[pos] = val[k];
                pos++;
            }
            pos += pad;
        }
        h_rp[mp] = pos;
        const int n_padded = pos;

        // ---- sweep items ----
        // Order: segments of split rows, whole very long rows, then the
        // tree-width bins from wide to narrow, so that long work starts first.
        std::vector<int4> h_items;
        std::vector<int4> h_seg_rows;
        std::vector<int> h_part_row;
        int n_parts = 0;
        int r = 0;
        {
            const int cnt = static_cast<int>(rows_in_bin[kLongBin].size());
            std::vector<int> giant;
            for (int q = 0; q < cnt; q++) {
                const int len = h_len[r + q];
                const int nseg = (len + kSegLen - 1) / kSegLen;
                if (T.max_split_segs < 0 || nseg <= T.max_split_segs) {
                    const int lr = static_cast<int>(h_seg_rows.size());
                    h_seg_rows.push_back(make_int4(r + q, n_parts, nseg, 0));
                    for (int sg = 0; sg < nseg; sg += kSegsPerItem)
                        h_items.push_back(make_int4(kItemSeg, lr, sg, std::min(kSegsPerItem, nseg - sg)));
                    for (int sg = 0; sg < nseg; sg++) h_part_row.push_back(lr);
                    n_parts += nseg;
                } else {
                    giant.push_back(r + q);
                }
            }
            for (int row : giant) h_items.push_back(make_int4(kItemLong, row, 1, 0));
            P->stats.n_long_rows = cnt;
            r += cnt;
        }
        for (int b = kLongBin - 1; b >= 0; b--) {
            const int cnt = static_cast<int>(rows_in_bin[b].size());
            if (cnt > 0) {
                const int width = 1 << b;
                int logg = std::min(T.group_log[b], b);
                logg = std::max(logg, b - 2);  // at most 4 entries per lane
                logg = std::min(logg, 6);
                const int logk = b - logg;
                const int per_pass = 64 >> logg;
                int rows_per_item = std::max(1, T.item_nnz / width);
                rows_per_item = ((rows_per_item + per_pass - 1) / per_pass) * per_pass;
                for (int q = 0; q < cnt; q += rows_per_item) {
                    const int nr = std::min(rows_per_item, cnt - q);
                    h_items.push_back(make_int4(kItemGroup | (logg << 4) | (logk << 8), r + q, nr, 0));
                }
            }
            r += cnt;
        }

        P->stats.m = m;
        P->stats.n = n;
        P->stats.nnz = nnz;
        P->stats.m_nonempty = mp;
        P->stats.n_items = static_cast<int>(h_items.size());

        // ---- device ----
        check(hipGetDevice(&P->device), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, P->device), "hipGetDeviceProperties");
        P->cu_count = prop.multiProcessorCount;
        check(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&P->ev0), "hipEventCreate");
        check(hipEventCreate(&P->ev1), "hipEventCreate");
        hipStream_t s = P->stream;

        const int nwords = (n + 31) / 32;
        DevMatrix &M = P->M;
        M.m = mp;
        M.n = n;
        M.nnz = nnz;
        M.n16 = ncls[0];
        M.n_int = ncls[0] + ncls[1];
        M.nwords = nwords;
        M.n_items = static_cast<int>(h_items.size());
        M.n_seg_rows = static_cast<int>(h_seg_rows.size());
        M.n_seg_parts = n_parts;
        int *d_rp = dev_alloc<int>(mp + 1, P->owned);
        double2 *d_sides = dev_alloc<double2>(mp, P->owned);
        int *d_col = dev_alloc<int>(n_padded, P->owned);
        double *d_val = dev_alloc<double>(n_padded, P->owned);
        int4 *d_items = dev_alloc<int4>(h_items.size(), P->owned);
        int4 *d_seg_rows = dev_alloc<int4>(h_seg_rows.size(), P->owned);
        int *d_part_row = dev_alloc<int>(h_part_row.size(), P->owned);
        int *d_perm = dev_alloc<int>(n, P->owned);
        check(hipMemcpyAsync(d_rp, h_rp.data(), (mp + 1) * sizeof(int), hipMemcpyHostToDevice, s), "H2D");
        check(hipMemcpyAsync(d_sides, h_sides.data(), mp * sizeof(double2), hipMemcpyHostToDevice, s), "H2D");
        check(hipMemcpyAsync(d_col, h_col.data(), n_padded * sizeof(int), hipMemcpyHostToDevice, s), "H2D");
        check(hipMemcpyAsync(d_val, h_val.data(), n_padded * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
        if (!h_items.empty())
            check(hipMemcpyAsync(d_items, h_items.data(), h_items.size() * sizeof(int4), hipMemcpyHostToDevice, s),
                  "H2D");
        if (!h_seg_rows.empty())
            check(hipMemcpyAsync(d_seg_rows, h_seg_rows.data(), h_seg_rows.size() * sizeof(int4),
                                 hipMemcpyHostToDevice, s),
                  "H2D");
        if (!h_part_row.empty())
            check(hipMemcpyAsync(d_part_row, h_part_row.data(), h_part_row.size() * sizeof(int),
                                 hipMemcpyHostToDevice, s),
                  "H2D");
        check(hipMemcpyAsync(d_perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, s), "H2D");
        M.row_ptr = d_rp;
        M.sides = d_sides;
        M.col = d_col;
        M.val = d_val;
        M.items = d_items;
        M.seg_rows = d_seg_rows;
        M.part_row = d_part_row;
        M.perm = d_perm;

        DevState &D = P->D;
        D.S16 = dev_alloc<unsigned>(ncls[0], P->owned);
        D.S32 = dev_alloc<int2>(ncls[1], P->owned);
        D.SF = dev_alloc<double2>(ncls[2], P->owned);
        D.AI = dev_alloc<int2>(M.n_int, P->owned);
        D.AX = dev_alloc<double2>(M.n_int, P->owned);
        D.AF = dev_alloc<double2>(ncls[2], P->owned);
        D.bm = dev_alloc<unsigned>(nwords, P->owned);
        D.status = dev_alloc<int>(kStatusRows * kStatusWords, P->owned);
        D.barrier = dev_alloc<unsigned>(kBarrierWords, P->owned);
        D.seg_parts = dev_alloc<SegPart>(n_parts, P->owned);
        D.seg_state = dev_alloc<SegState>(h_seg_rows.size(), P->owned);
        D.seg_count = dev_alloc<unsigned>(h_seg_rows.size(), P->owned);
        D.apply_tasks = dev_alloc<int>(n_parts, P->owned);

        int per_cu = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(mega_kernel), 256,
                                                           0),
              "occupancy");
        if (T.mega_blocks_per_cu > 0) per_cu = std::min(per_cu, T.mega_blocks_per_cu);
        P->mega_grid = std::max(1, per_cu) * P->cu_count;
        check(hipStreamSynchronize(s), "upload");
    } catch (...) {
        delete P;
        throw;
    }
    return P;
}

void problem_destroy(Problem *p) { delete p; }

ProblemStats problem_stats(const Problem *p) { return p->stats; }

namespace {

int grid_for(int items) { return std::max(1, (items + 3) / 4); }

// Walks the per-round status rows exactly as the persistent kernel does.
void outcome_from_status(const std::vector<int> &st, int max_rounds, RunResult &res) {
    res.rounds = 0;
    res.infeasible = 0;
    res.capped = 0;
    for (int round = 1; round <= max_rounds; round++) {
        const int *row = st.data() + kStatusWords * round;
        res.rounds = round;
        if (row[kStatusInfeasible]) { res.infeasible = 1; return; }
        if (!row[kStatusChanged]) return;
        if (round == max_rounds) res.capped = 1;
    }
}

}  // namespace

RunResult problem_propagate_device(Problem *P, double *d_lbs, double *d_ubs, int sync_type, int max_rounds) {
    if (sync_type < CPU_LOOP_ID || sync_type > MEGAKERNEL_ID) throw std::runtime_error("Unknown sync type");
    if (max_rounds < 1 || max_rounds > kMaxRounds) throw std::runtime_error("max_rounds out of range");
    std::lock_guard<std::mutex> lock(P->run_mutex);
    const DevMatrix &M = P->M;
    const DevState &D = P->D;
    hipStream_t s = P->stream;
    RunResult res;

    const int fill_grid = std::min(4096, std::max(1, (P->n + 255) / 256));
    hipLaunchKernelGGL(clear_envelope_kernel, dim3(1), dim3(1), 0, s, D);
    hipLaunchKernelGGL(init_kernel, dim3(fill_grid), dim3(256), 0, s, M, D, d_lbs, d_ubs);
    check(hipGetLastError(), "init_kernel launch");

    const int sweep_grid = grid_for(M.n_items);
    // publish: one thread per column; apply: a small grid that takes the
    // queued segments in turn (an empty queue costs one short launch)
    const int pub_grid = std::max(1, std::min(4096, (M.n + 255) / 256));
    const int apply_grid = std::max(1, std::min(1024, grid_for(M.n_seg_parts)));
    std::vector<int> st(kStatusRows * kStatusWords, 0);
    int head[kStatusWords];
    check(hipMemcpyAsync(head, D.status, sizeof(head), hipMemcpyDeviceToHost, s), "flags D2H");
    check(hipStreamSynchronize(s), "init sync");
    if (head[kGlobalEnvelope]) throw std::runtime_error("bounds outside the problem's bound envelope");
    if (head[kGlobalCrossed]) {
        // infeasible before any round: bounds are returned unchanged
        res.infeasible = 1;
        return res;
    }

    check(hipEventRecord(P->ev0, s), "event");
    if (sync_type == CPU_LOOP_ID) {
        // two launches per round; the host reads the round's flags back
        for (int round = 1; round <= max_rounds; round++) {
            hipLaunchKernelGGL(sweep_kernel, dim3(sweep_grid), dim3(256), 0, s, M, D, round, 0);
            if (M.n_seg_parts > 0)
                hipLaunchKernelGGL(apply_split_kernel, dim3(apply_grid), dim3(256), 0, s, M, D, round, 0);
            hipLaunchKernelGGL(publish_kernel, dim3(pub_grid), dim3(256), 0, s, M, D, round, 0);
            check(hipGetLastError(), "round launch");
            int flags[kStatusWords];
            check(hipMemcpyAsync(flags, D.status + kStatusWords * round, sizeof(flags), hipMemcpyDeviceToHost, s),
                  "flags D2H");
            check(hipMemcpyAsync(head, D.status, sizeof(head), hipMemcpyDeviceToHost, s), "flags D2H");
            check(hipStreamSynchronize(s), "round sync");
            if (head[kGlobalError]) throw std::runtime_error("device wait timed out");
            std::memcpy(st.data() + kStatusWords * round, flags, sizeof(flags));
            if (flags[kStatusInfeasible] || !flags[kStatusChanged]) break;
        }
    } else if (sync_type == GPU_LOOP_ID) {
        // captured chunks of rounds whose launches stop themselves on the device
        const int chunk = std::max(1, P->tuning.graph_rounds);
        for (int first = 1, ci = 0; first <= max_rounds; first += chunk, ci++) {
            const int last = std::min(max_rounds, first + chunk - 1);
            if (static_cast<int>(P->graphs.size()) <= ci) {
                P->graphs.resize(ci + 1, nullptr);
                P->graph_last.resize(ci + 1, 0);
            }
            if (P->graphs[ci] && P->graph_last[ci] != last) {
                (void)hipGraphExecDestroy(P->graphs[ci]);
                P->graphs[ci] = nullptr;
            }
            if (!P->graphs[ci]) {
                hipGraph_t graph;
                check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
                for (int round = first; round <= last; round++) {
                    hipLaunchKernelGGL(sweep_kernel, dim3(sweep_grid), dim3(256), 0, s, M, D, round, 1);
                    if (M.n_seg_parts > 0)
                        hipLaunchKernelGGL(apply_split_kernel, dim3(apply_grid), dim3(256), 0, s, M, D, round, 1);
                    hipLaunchKernelGGL(publish_kernel, dim3(pub_grid), dim3(256), 0, s, M, D, round, 1);
                }
                check(hipStreamEndCapture(s, &graph), "end capture");
                hipGraphExec_t exec;
                check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
                (void)hipGraphDestroy(graph);
                P->graphs[ci] = exec;
                P->graph_last[ci] = last;
            }
            check(hipGraphLaunch(P->graphs[ci], s), "graph launch");
            int flags[kStatusWords];
            check(hipMemcpyAsync(flags, D.status + kStatusWords * last, sizeof(flags), hipMemcpyDeviceToHost, s),
                  "flags D2H");
            check(hipStreamSynchronize(s), "graph sync");
            // the next chunk is needed only if the chunk's last round still changed
            if (flags[kStatusInfeasible] || !flags[kStatusChanged]) break;
        }
        check(hipMemcpyAsync(st.data(), D.status, st.size() * sizeof(int), hipMemcpyDeviceToHost, s), "status D2H");
        check(hipStreamSynchronize(s), "status sync");
        if (st[kGlobalError]) throw std::runtime_error("device wait timed out");
    } else {
        void *args[] = {const_cast<DevMatrix *>(&M), const_cast<DevState *>(&D), &max_rounds};
        check(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(mega_kernel), dim3(P->mega_grid), dim3(256),
                                         args, 0, s),
              "mega_kernel launch");
    }
    check(hipEventRecord(P->ev1, s), "event");

    if (sync_type == MEGAKERNEL_ID) {
        check(hipMemcpyAsync(st.data(), D.status, kStatusWords * sizeof(int), hipMemcpyDeviceToHost, s),
              "outcome D2H");
        check(hipStreamSynchronize(s), "mega sync");
        if (st[kGlobalError] || st[kOutRounds] < 0) throw std::runtime_error("device wait timed out");
        res.rounds = st[kOutRounds];
        res.infeasible = st[kOutInfeasible];
        res.capped = st[kOutCapped];
    } else {
        outcome_from_status(st, max_rounds, res);
    }
    float ms = 0.f;
    check(hipEventSynchronize(P->ev1), "event sync");
    check(hipEventElapsedTime(&ms, P->ev0, P->ev1), "elapsed");
    res.kernel_ms = ms;

    hipLaunchKernelGGL(unpack_kernel, dim3(fill_grid), dim3(256), 0, s, M, D, d_lbs, d_ubs);
    check(hipGetLastError(), "unpack_kernel launch");
    check(hipStreamSynchronize(s), "unpack sync");
    return res;
}

RunResult problem_propagate(Problem *P, double *lbs, double *ubs, int sync_type, int max_rounds) {
    const int n = P->n;
    double *d_b = nullptr;
    check(hipMalloc(&d_b, 2 * sizeof(double) * static_cast<size_t>(n)), "hipMalloc bounds");
    RunResult res;
    try {
        check(hipMemcpy(d_b, lbs, n * sizeof(double), hipMemcpyHostToDevice), "H2D lbs");
        check(hipMemcpy(d_b + n, ubs, n * sizeof(double), hipMemcpyHostToDevice), "H2D ubs");
        res = problem_propagate_device(P, d_b, d_b + n, sync_type, max_rounds);
        check(hipMemcpy(lbs, d_b, n * sizeof(double), hipMemcpyDeviceToHost), "D2H lbs");
        check(hipMemcpy(ubs, d_b + n, n * sizeof(double), hipMemcpyDeviceToHost), "D2H ubs");
    } catch (...) {
        (void)hipFree(d_b);
        throw;
    }
    (void)hipFree(d_b);
    return res;
}

}  // namespace gdp
