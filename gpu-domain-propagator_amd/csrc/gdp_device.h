// gdp_device.h - device-side layout of a propagation problem.
//
// A problem holds one matrix layout and the state of one or more propagation
// units.  A unit is one bound vector being propagated with its own round
// count, change and infeasibility flags.  Two unit modes:
//   blocks    the matrix is a block-diagonal batch of independent instances;
//             unit u owns the rows and columns of instance u, all units share
//             the state arrays (disjoint parts of them)
//   replicas  all units share one matrix; unit u has its own copy of every
//             state array (offset u * per-unit size); a single propagation is
//             the replicas mode with one unit
// Every launch works on a step: unit u runs its round (step - start_u + 1)
// in step `step`.  Units that have converged, crossed or reached their round
// cap skip their work, so units of different ages share a step (continuous
// batching) and finished units cost nothing.
//
// Columns are relabelled internally ("internal ids") and grouped by how their
// bounds are stored:
//   ids [0, n16)      integer columns whose bounds stay integral within
//                     [-32767, 32767]: snapshot record 2 x int16 = 4 bytes
//   ids [n16, n_int)  integer columns whose bounds stay integral within the
//                     int32 range: snapshot record 2 x int32 = 8 bytes
//   ids [n_int, n)    all other columns: snapshot record 2 x double = 16 bytes
// In blocks mode each (class, instance) range is padded to a multiple of 32
// ids, so that every 32-column bitmap word belongs to one unit.  The entry
// order inside every row is unchanged, so the relabelling does not alter any
// arithmetic.
//
// HBM layout (rows grouped by length):
//   row_ptr[m'+1]  int32   start of each of the m' nonempty rows (multiple of
//                          kRowAlign) | padding entries after the row
//   sides[m']      double2 (lhs, rhs) per row
//   col[nnz']      int32   internal column id; bit 31 set for an integer column
//   val[nnz']      double  coefficients
//   items[]        int4    work items of the sweep
//   csc            col_ptr[n+1], col_rows[nnz]: rows of each column (marking)
// per unit:
//   snapshot       S16[n16] uint32, S32[n32] int2, SF[nf] double2: the bounds
//                  every row of a round reads, in one block of 16-byte chunks
//                  (S32 from byte snap_off32, SF from byte snap_offF, both
//                  multiples of 16; blocks of consecutive units snap_stride
//                  bytes apart) so that any record is read with one aligned
//                  16-byte load
//   accumulator    AI[n_int] int2 (integer atomics), AF[nf] double2 (FP64
//                  atomics), AX[n_int] double2 (exact values of integer
//                  candidates beyond the int32 range; they only arise when
//                  bounds cross)
//   bm[nwords]     uint32  columns moved in the current step (scanned by publish)
//   marks[m']      int32   step in which the row must be processed
//   item_mark[]    int32   per group item: step in which one of its rows must be
//                          processed (set with the row marks), so that a
//                          step with few marked rows visits few items
//   split rows     seg_parts, seg_state, seg_count, seg_marks: a marked split
//                  row recomputes only the segments holding a column that
//                  moved in the previous step (and segments not yet computed
//                  in the run); the others keep their stored partials
#pragma once

#include <hip/hip_runtime.h>

namespace gdp {

constexpr int kIdMask = 0x3fffffff;
constexpr int kIntFlag = static_cast<int>(0x80000000u);
constexpr int kItemGroup = 0;
constexpr int kItemLong = 1;   // one row, one wave, all segments (rows not split)
constexpr int kItemSeg = 2;    // up to kSegsPerItem consecutive segments of a split row
constexpr int kSegsPerItem = 4;
constexpr int kItemUnitShift = 12;  // items[].x >> 12: unit of the item (blocks mode)
// Most units of one persistent launch whose activity its workgroups keep in LDS.
constexpr int kMaxFilterUnits = 4096;
// Bytes of the persistent kernel's dynamic LDS in a launch of several units:
// the working / all-rows bits and first steps of up to kMaxFilterUnits units.
constexpr int kFilterLdsBytes = (kMaxFilterUnits / 32) * 4 * 2 + kMaxFilterUnits * 4;
// Binary columns (integer columns whose bound envelope lies in [0, 1]) have
// internal ids [0, n_bin) in replicas mode, and each unit keeps their bounds
// also as 2-bit codes, 16 columns to a 32-bit word: bit 0 = (lb >= 1), bit
// 1 = (ub < 1).  As bounds only tighten during a run, a code's bits are only
// ever set, so the publish of a step ORs new bits in.  A persistent launch of
// one unit copies the codes of the first n_lds binary columns into LDS at
// the start of every step and reads their bounds there instead of gathering
// their records.  At most kMaxLdsBinBytes of LDS per workgroup.
constexpr int kMaxLdsBinBytes = 64 * 1024;

// Rows start at multiples of kRowAlign entries (padding entries hold column
// 0 and coefficient 0 and are never used); the low bits of row_ptr[r] hold
// the number of padding entries after row r.  Aligned starts let a lane load
// its 4 consecutive entries with one 16-byte column load and two 16-byte
// coefficient loads.
constexpr int kRowAlign = 4;

// Per-step status words, in a ring of kStatusRows rows (row step % kStatusRows).
constexpr int kStatusWords = 8;
constexpr int kStatusRows = 128;
constexpr int kStatusChanged = 0;  // some unit moved a bound in the step
constexpr int kStatusApply = 2;    // split-row segments queued for the apply phase
// Words of the problem-wide status block (separate from the ring).
constexpr int kGlobalError = 0;    // a bounded device wait timed out
constexpr int kGlobalSteps = 1;    // steps run by the persistent kernel
constexpr int kGlobalWords = 8;

// State of one unit (int32 words, read as 64-bit pairs).  A step's changes
// go to slot step & 1 of `changed`, and step s asks slot (s - 1) & 1 whether
// the unit moved in its previous round: writers and readers of one step
// never share a slot.
struct UnitState {
    int start;        // step of the unit's round 1; 0: no bound vector loaded
    int infeasible;   // step in which bounds crossed (0: none)
    int markall;      // step in which every row of the unit is processed
    int max_rounds;   // round cap
    int crossed;      // input bounds already crossed
    int envelope;     // input bounds outside the storage-class envelope
    int changed[2];   // last step of each parity in which the unit moved a bound
    int mark_limit;   // accepted candidates of a step from which all rows are processed
    int pad[7];
};

// Accepted candidates counted per unit and step (ring of kMoveSlots steps,
// kMoveLanes counters per slot to spread the atomics); a unit whose count
// reaches its threshold processes all rows next step instead of pushing row
// marks (the count bounds the number of moved columns from above).
constexpr int kMoveSlots = 4;
constexpr int kMoveLanes = 16;
// Each counter on its own 128-byte line (ints between two counters).
constexpr int kMoveStride = 32;
// Index of counter j of unit u's slot `slot`.
__host__ __device__ inline int moved_index(int u, int slot, int j) {
    return ((u * kMoveSlots + slot) * kMoveLanes + j) * kMoveStride;
}

// Grid barrier words, each counter on its own 128-byte line.
constexpr int kBarrierTop = 0;
constexpr int kBarrierGen = 32;
constexpr int kBarrierGroup = 64;  // + 32 * group, 8 groups
constexpr int kBarrierWords = 64 + 32 * 8;

// wall_clock64 runs at 100 MHz on CDNA; give up a wait after 1 s
constexpr long long kWaitTimeoutTicks = 100000000LL;

// Partial activity of one segment of a split row.
struct SegPart {
    double smin, smax, dmax;
    int infs;    // infinite contributions: min side in the low 16 bits, max side in the high
                 // (at most kSegLen each)
    int stamp;   // step in which the partial was computed (0: not in this run); written
                 // with infs as one 8-byte word
};
static_assert(kSegLen < (1 << 15), "a segment's infinite contributions fit in 16 bits");

// Activity of a split row, written by the last segment wave of a step
// when the row gate passes (read by the apply phase of that step).
struct SegState {
    double minact, maxact;
    int mininf, maxinf;  // written and read as one 8-byte word
};

struct DevMatrix {
    int m;            // nonempty rows
    int n;            // internal column ids (including padding ids)
    int n16, n_int;   // class boundaries of the internal ids
    int nwords;       // bitmap words per unit, n / 32
    int n_items;      // sweep items (of one unit in replicas mode)
    int replicas;     // 1: replicas mode, 0: blocks mode
    int mark_rows;    // 0: every step processes every row of an active unit
    const int *row_ptr;
    const double2 *sides;
    const int *col;
    const double *val;
    const int4 *items;
    const int4 *seg_rows;  // per split row: {row, first partial, segments, unit}
    const int *part_row;   // per segment of a split row: its split-row index
    int n_seg_rows;
    int n_seg_parts;
    const int *col_ptr;    // csc of the internal ids (marking)
    const int *col_rows;
    const int *col_seg;    // per csc entry: its segment of a split row (-1: other rows)
    const int *col_item;   // per csc entry: the group item holding its row (-1: a split or
                           // unsplit long row, whose items test the row mark itself)
    const int *item_row;   // per item: the row whose mark a seg / long item tests (-1: group
                           // item, which tests its own item mark)
    const int *perm;       // internal id -> caller's column (-1: padding id)
    const int *word_unit;  // blocks mode: unit of each bitmap word
    int snap_off32;        // byte offset of S32 in a unit's snapshot block
    int snap_offF;         // byte offset of SF
    long long snap_stride; // bytes between the snapshot blocks of two units
    int n_bin;             // binary columns: internal ids [0, n_bin) (replicas mode; blocks: 0)
    int sb_words;          // words of a unit's binary codes: ceil(n_bin / 64) * 4
};

struct DevState {
    unsigned *S16;
    int2 *S32;
    double2 *SF;
    int2 *AI;
    double2 *AX;
    double2 *AF;
    unsigned *bm;
    int *marks;
    int *item_mark;        // per group item: step in which one of its rows must be processed
    SegPart *seg_parts;
    int *seg_marks;        // step in which a split row's segment must be recomputed
    SegState *seg_state;
    unsigned *seg_count;   // segments reduced in the current step
    // shared by all units
    UnitState *units;
    int *moved;            // [units][kMoveSlots][kMoveLanes] accepted candidates, kMoveStride ints apart
    int *status;           // kStatusRows * kStatusWords ring
    int *global;           // kGlobalWords
    unsigned *barrier;     // kBarrierWords
    int *apply_tasks;      // unit * n_seg_parts + segment, per step
    const int *active;     // units of the launch (replicas mode)
    int n_active;
    int n_lds;             // binary columns read from LDS (persistent launch of one unit; else 0)
    unsigned *SB;          // 2-bit codes of the binary columns (sb_words per unit)
};

// State arrays of unit u (replicas mode); the shared arrays are unchanged.
__device__ __forceinline__ DevState unit_view(const DevMatrix &M, const DevState &D, int u) {
    if (!M.replicas) return D;
    DevState V = D;
    const size_t k = static_cast<size_t>(u);
    const size_t snap = k * static_cast<size_t>(M.snap_stride);
    V.S16 = reinterpret_cast<unsigned *>(reinterpret_cast<char *>(D.S16) + snap);
    V.S32 = reinterpret_cast<int2 *>(reinterpret_cast<char *>(D.S32) + snap);
    V.SF = reinterpret_cast<double2 *>(reinterpret_cast<char *>(D.SF) + snap);
    V.AI += k * M.n_int;
    V.AX += k * M.n_int;
    V.AF += k * (M.n - M.n_int);
    V.bm += k * M.nwords;
    V.marks += k * M.m;
    V.item_mark += k * M.n_items;
    V.seg_parts += k * M.n_seg_parts;
    V.seg_marks += k * M.n_seg_parts;
    V.seg_state += k * M.n_seg_rows;
    V.seg_count += k * M.n_seg_rows;
    V.SB += k * M.sb_words;
    return V;
}

}  // namespace gdp
