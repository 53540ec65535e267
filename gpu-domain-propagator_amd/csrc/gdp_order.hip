// gdp_order.hip - the order of a problem's rows, computed on the device.
//
// Layout rows are the nonempty rows of all instances, ordered by tree-width
// bin (rows longer than kSegLen first, then the widths 2^9 down to 2^0), by
// instance inside a bin and by row number inside an instance.  That is a
// stable sort of the global row numbers (instance-major) by a 4-bit bin key,
// done here with rocPRIM's radix sort (through hipCUB):
//   keys     one thread per global row: its length from the row starts, its
//            key (0: longer than kSegLen, 10 - log2(tree width) for the
//            width bins, 15: empty), per-key row, nonzero and padding totals
//   sort     stable radix sort of (key, row) over the 4 key bits
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
#include <hipcub/hipcub.hpp>
#include <stdexcept>
#include <string>

#include "gdp_arith.h"
#include "gdp_device.h"
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
    size_t bytes = 0;
    check_order(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys_sorted, rows, rows_sorted, M, 0, 4, s),
                "sort size");
    void *tmp = nullptr;
    check_order(hipMalloc(&tmp, std::max<size_t>(bytes, 16)), "hipMalloc");
    const hipError_t e =
        hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, keys_sorted, rows, rows_sorted, M, 0, 4, s);
    const hipError_t f = e == hipSuccess ? hipStreamSynchronize(s) : e;
    (void)hipFree(tmp);
    check_order(f, "row sort");
}

void order_place_rows(int mp, int NI, const int *row_off, const long long *nnz_off, const int *rp_cat,
                      const double *lhs, const double *rhs, const unsigned char *keys_sorted, const int *order,
                      int *len, int *plen, long long *src, int *inst, double2 *sides, int *run_start, hipStream_t s) {
    hipLaunchKernelGGL(place_rows_kernel, dim3(grid_of(mp)), dim3(256), 0, s, mp, NI, row_off, nnz_off, rp_cat, lhs,
                       rhs, keys_sorted, order, len, plen, src, inst, sides, run_start);
    check_order(hipGetLastError(), "place rows launch");
}

void order_exclusive_sum(const int *in, int *out, int n, hipStream_t s) {
    size_t bytes = 0;
    check_order(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s), "scan size");
    void *tmp = nullptr;
    check_order(hipMalloc(&tmp, std::max<size_t>(bytes, 16)), "hipMalloc");
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
    const hipError_t f = e == hipSuccess ? hipStreamSynchronize(s) : e;
    (void)hipFree(tmp);
    check_order(f, "scan");
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
