/*
 * gdp_def.h - enums, tolerances and epsilon comparison macros of the
 * MI355X domain propagator's C interface.
 *
 * Source compatible with the reference's src/def.h:15-79 and
 * src/params.h:1-19 (both are pulled into clients by the reference's
 * include/interface.h:5): the same enum names and values, the same
 * GDP_INF / GDP_EPS / MAX_NUM_ROUNDS constants and the same EPS* / IS*INF
 * comparison macros, so code written against the reference header compiles
 * and behaves the same against this one.
 */
#ifndef GDP_AMD_DEF_H
#define GDP_AMD_DEF_H

#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes: non-positive values are errors (reference src/def.h:15-20). */
enum GDP_Retcode {
    GDP_OKAY = +1,
    GDP_ERROR = 0,
    GDP_NOTIMPLEMENTED = -18
};
typedef enum GDP_Retcode GDP_RETCODE;

/* Variable types; only "== GDP_CONTINUOUS" changes the arithmetic
 * (reference src/def.h:26-31). */
enum GDP_Vartype {
    GDP_BINARY = 0,
    GDP_INTEGER = 1,
    GDP_CONTINUOUS = 3
};
typedef enum GDP_Vartype GDP_VARTYPE;

/* How propagation rounds are driven (reference src/def.h:36-41).
 *   CPU_LOOP   - one kernel launch per round, host reads the change flag.
 *   GPU_LOOP   - rounds run without host round-trips from a captured HIP
 *                graph of round launches that exit early on the device once
 *                the fixpoint is reached (HIP has no dynamic parallelism).
 *   MEGAKERNEL - one persistent launch iterates all rounds with device-wide
 *                barriers. */
enum GDP_Synctype {
    CPU_LOOP = 0,
    GPU_LOOP = 1,
    MEGAKERNEL = 2
};
typedef enum GDP_Synctype GDP_SYNCTYPE;

enum GDP_algorithm {
    CPU_SEQ_DOUBLE = 0,
    CPU_OMP_DOUBLE = 1,
    GPU_ATOMIC_DOUBLE = 2,
    GPU_REDUCTION_DOUBLE = 3,
    CPU_SEQ_DISJOINT_DOUBLE = 4
};
typedef enum GDP_algorithm GDP_ALGORITHM;

#ifdef __cplusplus
}
#endif

/* Tunables kept for source compatibility (reference src/params.h). */
#ifndef NNZ_PER_WG
#define NNZ_PER_WG 256
#endif
#define MAX_NUM_ROUNDS 100
#define SHARED_MEM_THREADS 8
#define GDP_INF 1e20
#define GDP_EPS 1e-9

/* Absolute-epsilon comparisons; identical meaning to reference
 * src/def.h:68-79.  Every one is evaluated as "difference, then compare",
 * which is what makes the arithmetic reproducible bit for bit. */
#define REALABS(x)     (fabs(x))
#define EPSLE(x, y)    ((x) - (y) <= (GDP_EPS))
#define EPSLT(x, y)    ((x) - (y) < -(GDP_EPS))
#define EPSGE(x, y)    ((x) - (y) >= -(GDP_EPS))
#define EPSGT(x, y)    ((x) - (y) > (GDP_EPS))
#define EPSEQ(x, y)    (REALABS((x) - (y)) <= (GDP_EPS))
#ifndef MAX
#define MAX(x, y)      ((x) >= (y) ? (x) : (y))
#endif
#ifndef MIN
#define MIN(x, y)      ((x) <= (y) ? (x) : (y))
#endif
#define EPSFLOOR(x)    (floor((x) + (GDP_EPS)))
#define EPSCEIL(x)     (ceil((x) - (GDP_EPS)))
#define ISPOSINF(x)    (EPSGE(x, GDP_INF))
#define ISNEGINF(x)    (EPSLE(x, -GDP_INF))

#endif /* GDP_AMD_DEF_H */
