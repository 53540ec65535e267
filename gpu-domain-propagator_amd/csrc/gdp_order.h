// gdp_order.h - device ordering of a problem's rows (gdp_order.hip).
//
// Inputs on the device: the instances' row starts concatenated (instance i's
// m_i + 1 entries from position row_off[i] + i), their sides concatenated
// (global row g = row_off[i] + local row), row_off (NI + 1 entries) and the
// instances' nonzero offsets in the concatenated CSR arrays (NI + 1).
// All calls are asynchronous on stream s except the sort and the scan,
// which return after completing (their temporary storage is freed).
#pragma once

#include <hip/hip_runtime.h>

namespace gdp {

// bin key of an empty row (sorted after every other key)
constexpr int kRowKeyEmpty = 15;

// keys[g], rows[g] = g for the M global rows; totals (48 entries, zeroed by
// the caller): [key] rows, [16 + key] nonzeros, [32 + key] padding entries.
void order_row_keys(int M, int NI, const int *row_off, const int *rp_cat, unsigned char *keys, int *rows,
                    unsigned long long *totals, hipStream_t s);

// Stable sort of the (key, row) pairs by the 4 key bits.
void order_sort_rows(const unsigned char *keys, unsigned char *keys_sorted, const int *rows, int *rows_sorted, int M,
                     hipStream_t s);

// Per layout row r < mp (the nonempty rows, in sorted order): length,
// padded length, source position, instance (inst may be null for one
// instance), sides; run_start[key * NI + i] = first layout row of the run of
// (key, instance) (entries of absent runs are left as they are).
void order_place_rows(int mp, int NI, const int *row_off, const long long *nnz_off, const int *rp_cat,
                      const double *lhs, const double *rhs, const unsigned char *keys_sorted, const int *order,
                      int *len, int *plen, long long *src, int *inst, double2 *sides, int *run_start, hipStream_t s);

// out[k] = in[0] + ... + in[k - 1] for k < n (int).
void order_exclusive_sum(const int *in, int *out, int n, hipStream_t s);

// row_ptr[r] |= plen[r] - len[r] for r < mp (the padding count).
void order_pad_rows(int mp, const int *len, const int *plen, int *row_ptr, hipStream_t s);

// row_item[r] = group item of layout row r; row_part[seg_rows[k].x] =
// seg_rows[k].y.  Entries not written keep their values (the caller sets -1).
void order_row_tables(int n_items, const int4 *items, int n_seg_rows, const int4 *seg_rows, int *row_item,
                      int *row_part, hipStream_t s);

}  // namespace gdp
