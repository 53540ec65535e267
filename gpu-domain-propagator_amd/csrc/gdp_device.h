// gdp_device.h - device-side layout of a propagation problem.
//
// HBM layout (all arrays 256-byte aligned, rows grouped by length):
//   row_ptr[m'+1]  int32   CSR row pointers of the m' nonempty rows
//   sides[m']      double2 (lhs, rhs) per row, one 16-byte load
//   col[nnz]       int32   column index; bit 31 set when the column is an
//                          integer variable, so the rounding rule needs no
//                          extra gather
//   val[nnz]       double  coefficients
//   items[]        int4    work items of a round (see gdp_kernels.hip)
//   bnd0/bnd1[n]   double2 (lb, ub) per column, snapshot / accumulator
//   bm0/bm1[nw]    uint32  bitmaps of the columns changed in a round
//   status[]       int32   per-round flags and the outcome of a run
#pragma once

#include <hip/hip_runtime.h>

namespace gdp {

constexpr int kColMask = 0x7fffffff;
constexpr int kItemGroup = 0;
constexpr int kItemLong = 1;   // one row, one wave, all segments (very long rows)
constexpr int kItemSeg = 2;    // one segment of a split row

// Rows longer than kSegLen are split into one item per segment when they
// have at most this many segments; longer rows go to single-wave items.
// Segment waves of one row wait for each other, which is deadlock-free as
// long as this bound is below the number of co-resident waves.
constexpr int kMaxSplitSegs = 512;

// status words: row 0 holds the outcome, row r >= 1 the flags of round r
constexpr int kStatusWords = 8;
constexpr int kOutRounds = 0;
constexpr int kOutResultBuf = 1;
constexpr int kOutInfeasible = 2;
constexpr int kOutCapped = 3;
constexpr int kGlobalError = 4;
constexpr int kGlobalDone = 5;
constexpr int kStatusChanged = 0;
constexpr int kStatusInfeasPrev = 1;
constexpr int kStatusRows = 128;  // rounds 1..101 fit

// wall_clock64 runs at 100 MHz on CDNA; give up a barrier wait after 1 s
constexpr long long kBarrierTimeoutTicks = 100000000LL;

// Partial activity of one segment of a split row.
struct SegPart {
    double smin, smax, dmax;
    int mininf, maxinf;
};

// Activity and gate of a split row, published by the last segment wave.
struct SegState {
    double minact, maxact;
    int mininf, maxinf, gate, pad;
};

struct DevMatrix {
    int m;            // nonempty rows
    int n;            // columns
    int nnz;
    int nwords;       // bitmap words, ceil(n / 32)
    int n_items;      // row items
    int n_fix_items;  // fixup items, ceil(nwords / 64)
    const int *row_ptr;
    const double2 *sides;
    const int *col;
    const double *val;
    const int4 *items;
    const int4 *seg_rows;  // per split row: {row, first partial, segments, 0}
    int n_seg_rows;
    int n_seg_parts;
};

struct DevState {
    double2 *bnd0;
    double2 *bnd1;
    unsigned *bm0;
    unsigned *bm1;
    int *status;        // kStatusRows * kStatusWords
    unsigned *barrier;  // [count, generation]
    SegPart *seg_parts;    // [n_seg_parts]
    SegState *seg_state;   // [n_seg_rows]
    unsigned *seg_count;   // [n_seg_rows] arrivals in the current round
    int *seg_flag;         // [n_seg_rows] round whose state is published
};

}  // namespace gdp
