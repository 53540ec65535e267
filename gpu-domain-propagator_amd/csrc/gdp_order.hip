// gdp_order.hip - the order of a problem's rows, computed on the device.
//
// Layout rows are the nonempty rows of all instances, ordered by tree-width
// bin (rows longer than kSegLen first, then the widths 2^9 down to 2^0), by
// instance inside a bin and by row number inside an instance.  That is a
// stable sort of the global row numbers (instance-major) by a 4-bit bin key:
//   keys     one thread per global row: its length from the row starts, its
//            key (0: longer than kSegLen, 10 - log2(tree width) for the
//            width bins, 15: empty), per-key row, nonzero and padding totals
//   sort     stable counting sort of (key, row) pairs: per-tile key counts,
//            a scan of the counts (key-major: all tiles of key 0, then key
//            1, ...), and a scatter in which each thread's rank among the
//            tile's elements of its key comes from a scan in LDS, so equal
//            keys keep their order
//   place    one thread per layout row: its sides, source position in the
//            concatenated CSR arrays, instance, length and padded length, and
//            the first layout row of each (key, instance) run
//   scan     exclusive sum of the padded lengths: the row starts, with the
//            padding count put into their low bits afterwards
// The host reads back only the totals, the (key, instance) starts and the
// lengths of the long rows, from which it builds the sweep items, and then
// the rows' item and segment tables are filled here from the items.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "gdp_arith.h"
#include "gdp_device.h"
#include "gdp_devmem.h"
#include "gdp_order.h"

namespace gdp {

namespace {

void check_order(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Instance of global row g: the largest i with row_off[i] <= g.
__device__ __forceinline__ int instance_of(int g, int NI, const int *row_off) {
    int lo = 0, hi = NI - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_off[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int key_of_length(int len) {
    if (len == 0) return kRowKeyEmpty;
    if (len > kSegLen) return 0;
    const int lg = len == 1 ? 0 : 32 - __clz(len - 1);  // ceil(log2(len))
    return 10 - lg;
}

__device__ __forceinline__ int padded_length(int len) { return (len + kRowAlign - 1) & ~(kRowAlign - 1); }

// Per-key totals: [key] rows, [16 + key] nonzeros, [32 + key] padding
// entries, summed per workgroup in LDS.
__global__ __launch_bounds__(256) void row_keys_kernel(int M, int NI, const int *row_off, const int *rp_cat,
                                                       unsigned char *keys, int *rows,
                                                       unsigned long long *totals) {
    __shared__ unsigned long long s_tot[48];
    if (threadIdx.x < 48) s_tot[threadIdx.x] = 0;
    __syncthreads();
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < M; g += gridDim.x * blockDim.x) {
        const int i = NI > 1 ? instance_of(g, NI, row_off) : 0;
        const int len = rp_cat[g + i + 1] - rp_cat[g + i];
        const int key = key_of_length(len);
        keys[g] = static_cast<unsigned char>(key);
        rows[g] = g;
        if (key != kRowKeyEmpty) {
            atomicAdd(&s_tot[key], 1ull);
            atomicAdd(&s_tot[16 + key], static_cast<unsigned long long>(len));
            atomicAdd(&s_tot[32 + key], static_cast<unsigned long long>(padded_length(len) - len));
        }
    }
    __syncthreads();
    if (threadIdx.x < 48 && s_tot[threadIdx.x]) atomicAdd(&totals[threadIdx.x], s_tot[threadIdx.x]);
}

__global__ __launch_bounds__(256) void place_rows_kernel(int mp, int NI, const int *row_off,
                                                         const long long *nnz_off, const int *rp_cat,
                                                         const double *lhs, const double *rhs,
                                                         const unsigned char *keys, const int *order, int *len_out,
                                                         int *plen_out, long long *src_out, int *inst_out,
                                                         double2 *sides_out, int *run_start) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < mp; r += gridDim.x * blockDim.x) {
        const int g = order[r];
        const int i = NI > 1 ? instance_of(g, NI, row_off) : 0;
        const int a = rp_cat[g + i];
        const int len = rp_cat[g + i + 1] - a;
        len_out[r] = len;
        plen_out[r] = padded_length(len);
        src_out[r] = nnz_off[i] + a;
        if (inst_out) inst_out[r] = i;
        sides_out[r] = make_double2(lhs[g], rhs[g]);
        const int key = keys[r];
        bool first = r == 0 || keys[r - 1] != key;
        if (!first && NI > 1) first = instance_of(order[r - 1], NI, row_off) != i;
        if (first) run_start[key * NI + i] = r;
    }
}

__global__ __launch_bounds__(256) void pad_rows_kernel(int mp, const int *len, const int *plen, int *row_ptr) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < mp; r += gridDim.x * blockDim.x)
        row_ptr[r] |= plen[r] - len[r];
}

// row_item[r] = the group item holding layout row r; row_part[r] = the first
// segment of split row r.  Both start at -1 (set by the caller).  One wave
// per item, then one thread per split row.
__global__ __launch_bounds__(256) void row_tables_kernel(int n_items, const int4 *items, int n_seg_rows,
                                                         const int4 *seg_rows, int *row_item, int *row_part) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < n_items; q += waves) {
        const int4 it = items[q];
        if ((it.x & 15) != kItemGroup) continue;
        for (int t = lane; t < it.z; t += 64) row_item[it.y + t] = q;
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_seg_rows; k += gridDim.x * blockDim.x)
        row_part[seg_rows[k].x] = seg_rows[k].y;
}

// ---- counting sort and scan ----
// A tile is 4096 elements: 256 threads with 16 consecutive elements each.
constexpr int kTileThreads = 256;
constexpr int kTileItems = 16;
constexpr int kTile = kTileThreads * kTileItems;
constexpr int kSortKeys = 16;

// Exclusive scan of the 4096 ints in LDS `v` (thread t owns entries
// 16 t .. 16 t + 15); returns the total.  Every thread of the block calls it.
__device__ int block_exclusive_scan(int *v) {
    __shared__ int wave_sum[kTileThreads / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int own = 0;
#pragma unroll
    for (int i = 0; i < kTileItems; i++) own += v[t * kTileItems + i];
    // inclusive scan of the thread sums within the wave
    int inc = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int x = __shfl_up(inc, d, 64);
        if (lane >= d) inc += x;
    }
    if (lane == 63) wave_sum[w] = inc;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kTileThreads / 64; k++) {
        if (k < w) before += wave_sum[k];
        total += wave_sum[k];
    }
    int run = before + inc - own;
#pragma unroll
    for (int i = 0; i < kTileItems; i++) {
        const int x = v[t * kTileItems + i];
        v[t * kTileItems + i] = run;
        run += x;
    }
    __syncthreads();
    return total;
}

// hist[key * tiles + tile] = elements of the tile with that key.
__global__ __launch_bounds__(kTileThreads) void key_hist_kernel(long long M, const unsigned char *keys, int tiles,
                                                                int *hist) {
    __shared__ int cnt[kSortKeys];
    if (threadIdx.x < kSortKeys) cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    for (int i = threadIdx.x; i < kTile; i += kTileThreads) {
        const long long g = base + i;
        if (g < M) atomicAdd(&cnt[keys[g] & (kSortKeys - 1)], 1);
    }
    __syncthreads();
    if (threadIdx.x < kSortKeys) hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Moves every (key, row) pair to off[key * tiles + tile] + its rank among
// the tile's elements of that key, ranks in element order (stable).
__global__ __launch_bounds__(kTileThreads) void key_scatter_kernel(long long M, const unsigned char *keys,
                                                                   const int *rows, int tiles, const int *off,
                                                                   unsigned char *keys_out, int *rows_out) {
    __shared__ int pos[kSortKeys * kTileThreads];  // [key][thread]
    __shared__ int key_first[kSortKeys];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSortKeys; k++) pos[k * kTileThreads + t] = 0;
    const long long base = static_cast<long long>(blockIdx.x) * kTile + static_cast<long long>(t) * kTileItems;
    unsigned char kk[kTileItems];
#pragma unroll
    for (int i = 0; i < kTileItems; i++) {
        kk[i] = base + i < M ? static_cast<unsigned char>(keys[base + i] & (kSortKeys - 1)) : 0xff;
        if (kk[i] != 0xff) pos[kk[i] * kTileThreads + t]++;  // the thread's own entry
    }
    __syncthreads();
    block_exclusive_scan(pos);  // key-major: elements of smaller keys, then of earlier threads
    if (t < kSortKeys) key_first[t] = pos[t * kTileThreads];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortKeys; k++)
        pos[k * kTileThreads + t] += off[k * tiles + blockIdx.x] - key_first[k];
#pragma unroll
    for (int i = 0; i < kTileItems; i++) {
        if (kk[i] == 0xff) continue;
        const int d = pos[kk[i] * kTileThreads + t]++;
        keys_out[d] = kk[i];
        rows_out[d] = rows[base + i];
    }
}

// Exclusive scan of each tile of `in` into `out`; the tile totals go to
// partial[tile].
__global__ __launch_bounds__(kTileThreads) void scan_tiles_kernel(long long n, const int *in, int *out, int *partial) {
    __shared__ int v[kTile];
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    for (int i = threadIdx.x; i < kTile; i += kTileThreads) v[i] = base + i < n ? in[base + i] : 0;
    __syncthreads();
    const int total = block_exclusive_scan(v);
    for (int i = threadIdx.x; i < kTile; i += kTileThreads)
        if (base + i < n) out[base + i] = v[i];
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan of the tile totals in place (one workgroup, 4096 at a time
// with a running carry).
__global__ __launch_bounds__(kTileThreads) void scan_partials_kernel(int tiles, int *partial) {
    __shared__ int v[kTile];
    int carry = 0;
    for (int c0 = 0; c0 < tiles; c0 += kTile) {
        for (int i = threadIdx.x; i < kTile; i += kTileThreads) v[i] = c0 + i < tiles ? partial[c0 + i] : 0;
        __syncthreads();
        const int total = block_exclusive_scan(v);
        for (int i = threadIdx.x; i < kTile; i += kTileThreads)
            if (c0 + i < tiles) partial[c0 + i] = v[i] + carry;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kTileThreads) void scan_add_kernel(long long n, int *out, const int *partial) {
    const long long base = static_cast<long long>(blockIdx.x) * kTile;
    const int add = partial[blockIdx.x];
    for (int i = threadIdx.x; i < kTile; i += kTileThreads)
        if (base + i < n) out[base + i] += add;
}

// Exclusive scan of n ints on the device into out (may equal in); `partial`
// holds at least one int per tile.  Asynchronous on s.
void exclusive_scan_async(const int *in, int *out, long long n, int *partial, hipStream_t s) {
    const int tiles = static_cast<int>((n + kTile - 1) / kTile);
    if (tiles == 0) return;
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(tiles), dim3(kTileThreads), 0, s, n, in, out, partial);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kTileThreads), 0, s, tiles, partial);
    hipLaunchKernelGGL(scan_add_kernel, dim3(tiles), dim3(kTileThreads), 0, s, n, out, partial);
    check_order(hipGetLastError(), "scan launch");
}

int grid_of(long long n) { return static_cast<int>(std::max(1LL, std::min(4096LL, (n + 255) / 256))); }

}  // namespace

void order_row_keys(int M, int NI, const int *row_off, const int *rp_cat, unsigned char *keys, int *rows,
                    unsigned long long *totals, hipStream_t s) {
    hipLaunchKernelGGL(row_keys_kernel, dim3(grid_of(M)), dim3(256), 0, s, M, NI, row_off, rp_cat, keys, rows,
                       totals);
    check_order(hipGetLastError(), "row keys launch");
}

void order_sort_rows(const unsigned char *keys, unsigned char *keys_sorted, const int *rows, int *rows_sorted, int M,
                     hipStream_t s) {
    const int tiles = (M + kTile - 1) / kTile;
    if (tiles == 0) return;
    const int nh = kSortKeys * tiles;
    int *tmp = nullptr;  // histogram, then the scan's tile totals
    check_order(cached_malloc(&tmp, sizeof(int) * (static_cast<size_t>(nh) + (nh + kTile - 1) / kTile)), "hipMalloc");
    hipError_t e = hipSuccess;
    try {
        hipLaunchKernelGGL(key_hist_kernel, dim3(tiles), dim3(kTileThreads), 0, s, static_cast<long long>(M), keys,
                           tiles, tmp);
        check_order(hipGetLastError(), "sort launch");
        exclusive_scan_async(tmp, tmp, nh, tmp + nh, s);
        hipLaunchKernelGGL(key_scatter_kernel, dim3(tiles), dim3(kTileThreads), 0, s, static_cast<long long>(M), keys,
                           rows, tiles, tmp, keys_sorted, rows_sorted);
        check_order(hipGetLastError(), "sort launch");
        e = hipStreamSynchronize(s);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)cached_free(tmp);
        throw;
    }
    (void)cached_free(tmp);
    check_order(e, "row sort");
}

void order_place_rows(int mp, int NI, const int *row_off, const long long *nnz_off, const int *rp_cat,
                      const double *lhs, const double *rhs, const unsigned char *keys_sorted, const int *order,
                      int *len, int *plen, long long *src, int *inst, double2 *sides, int *run_start, hipStream_t s) {
    hipLaunchKernelGGL(place_rows_kernel, dim3(grid_of(mp)), dim3(256), 0, s, mp, NI, row_off, nnz_off, rp_cat, lhs,
                       rhs, keys_sorted, order, len, plen, src, inst, sides, run_start);
    check_order(hipGetLastError(), "place rows launch");
}

void order_exclusive_sum(const int *in, int *out, int n, hipStream_t s) {
    const int tiles = (n + kTile - 1) / kTile;
    if (tiles == 0) return;
    int *partial = nullptr;
    check_order(cached_malloc(&partial, sizeof(int) * tiles), "hipMalloc");
    hipError_t e = hipSuccess;
    try {
        exclusive_scan_async(in, out, n, partial, s);
        e = hipStreamSynchronize(s);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)cached_free(partial);
        throw;
    }
    (void)cached_free(partial);
    check_order(e, "scan");
}

void order_pad_rows(int mp, const int *len, const int *plen, int *row_ptr, hipStream_t s) {
    hipLaunchKernelGGL(pad_rows_kernel, dim3(grid_of(mp)), dim3(256), 0, s, mp, len, plen, row_ptr);
    check_order(hipGetLastError(), "pad rows launch");
}

void order_row_tables(int n_items, const int4 *items, int n_seg_rows, const int4 *seg_rows, int *row_item,
                      int *row_part, hipStream_t s) {
    const int grid = std::max(1, std::min(4096, std::max((n_items + 3) / 4, (n_seg_rows + 255) / 256)));
    hipLaunchKernelGGL(row_tables_kernel, dim3(grid), dim3(256), 0, s, n_items, items, n_seg_rows, seg_rows,
                       row_item, row_part);
    check_order(hipGetLastError(), "row tables launch");
}

}  // namespace gdp
