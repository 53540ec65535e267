// gdp_device.h - device-side layout of a propagation problem.
//
// Columns are relabelled internally ("internal ids") and grouped by how their
// bounds are stored:
//   ids [0, n16)      integer columns whose bounds stay integral within
//                     [-32767, 32767]: snapshot record 2 x int16 = 4 bytes
//   ids [n16, n_int)  integer columns whose bounds stay integral within the
//                     int32 range: snapshot record 2 x int32 = 8 bytes
//   ids [n_int, n)    all other columns: snapshot record 2 x double = 16 bytes
// The entry order inside every row is unchanged, so the relabelling does not
// alter any arithmetic.  Small records keep the randomly gathered working set
// of a round inside the per-XCD L2.
//
// HBM layout (rows grouped by length):
//   row_ptr[m'+1]  int32   start of each of the m' nonempty rows (multiple of
//                          kRowAlign) | padding entries after the row
//   sides[m']      double2 (lhs, rhs) per row
//   col[nnz]       int32   internal column id; bit 31 set for an integer column
//   val[nnz]       double  coefficients
//   items[]        int4    work items of the sweep
//   snapshot       S16[n16] uint32, S32[n32] int2, SF[nf] double2: the bounds
//                  every row of a round reads
//   accumulator    AI[n_int] int2 (integer atomics), AF[nf] double2 (FP64
//                  atomics), AX[n_int] double2 (exact values of integer
//                  candidates beyond the int32 range; they only arise when
//                  bounds cross)
//   bm[nw]         uint32  columns changed in the current round
//   status[]       int32   per-round flags and the outcome of a run
#pragma once

#include <hip/hip_runtime.h>

namespace gdp {

constexpr int kIdMask = 0x3fffffff;
constexpr int kIntFlag = static_cast<int>(0x80000000u);
constexpr int kItemGroup = 0;
constexpr int kItemLong = 1;   // one row, one wave, all segments (rows not split)
constexpr int kItemSeg = 2;    // up to kSegsPerItem consecutive segments of a split row
constexpr int kSegsPerItem = 4;

// Rows start at multiples of kRowAlign entries (padding entries hold column
// 0 and coefficient 0 and are never used); the low bits of row_ptr[r] hold
// the number of padding entries after row r.  Aligned starts let a lane load
// its 4 consecutive entries with one 16-byte column load and two 16-byte
// coefficient loads.
constexpr int kRowAlign = 4;

// status words: row 0 holds the outcome, row r >= 1 the flags of round r
constexpr int kStatusWords = 8;
constexpr int kOutRounds = 0;
constexpr int kOutInfeasible = 2;
constexpr int kOutCapped = 3;
constexpr int kGlobalError = 4;
constexpr int kGlobalDone = 5;
constexpr int kGlobalEnvelope = 6;  // input bounds outside the class envelope
constexpr int kGlobalCrossed = 7;   // input bounds already crossed
constexpr int kStatusChanged = 0;
constexpr int kStatusInfeasible = 1;
constexpr int kStatusApply = 2;   // segments queued for the apply phase
constexpr int kStatusRows = 128;  // rounds 1..100 fit

// grid barrier words, each counter on its own 128-byte line
constexpr int kBarrierTop = 0;
constexpr int kBarrierGen = 32;
constexpr int kBarrierGroup = 64;  // + 32 * group, 8 groups
constexpr int kBarrierWords = 64 + 32 * 8;

// wall_clock64 runs at 100 MHz on CDNA; give up a wait after 1 s
constexpr long long kWaitTimeoutTicks = 100000000LL;

// Partial activity of one segment of a split row.
struct SegPart {
    double smin, smax, dmax;
    int mininf, maxinf;
};

// Activity of a split row, written by the last segment wave of a round
// when the row gate passes (read by the apply phase of that round).
struct SegState {
    double minact, maxact;
    int mininf, maxinf;
};

struct DevMatrix {
    int m;            // nonempty rows
    int n;            // columns
    int nnz;
    int n16, n_int;   // class boundaries of the internal ids
    int nwords;       // bitmap words, ceil(n / 32)
    int n_items;      // sweep items
    const int *row_ptr;
    const double2 *sides;
    const int *col;
    const double *val;
    const int4 *items;
    const int4 *seg_rows;  // per split row: {row, first partial, segments, 0}
    const int *part_row;   // per segment of a split row: its split-row index
    int n_seg_rows;
    int n_seg_parts;
    const int *perm;       // internal id -> original column
};

struct DevState {
    unsigned *S16;
    int2 *S32;
    double2 *SF;
    int2 *AI;
    double2 *AX;
    double2 *AF;
    unsigned *bm;
    int *status;        // kStatusRows * kStatusWords
    unsigned *barrier;  // kBarrierWords: top counter, generation, 8 group counters
    SegPart *seg_parts;    // [n_seg_parts]
    SegState *seg_state;   // [n_seg_rows]
    unsigned *seg_count;   // [n_seg_rows] segments reduced in the current round
    int *apply_tasks;      // [n_seg_parts] segments whose row passed the gate
};

}  // namespace gdp
