/*
 * chameleon_amd.h -- entry points of libchameleon_amd.so that have no
 * counterpart in the reference Chameleon API.
 */
#ifndef CHAMELEON_AMD_EXTRAS_H
#define CHAMELEON_AMD_EXTRAS_H

#include "chameleon.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes timed while CHAMELEON_GENERATE_STATS is enabled (the
 * reference's GENERATE_STATS switches on StarPU per-kernel statistics,
 * runtime/starpu/control/runtime_profiling.c:120-150; here each launch of
 * the class is bracketed by HIP events on its own stream). */
enum {
    CHAMAMD_KERNEL_UPDATE = 0,   /* trailing SYRK+GEMM update of the factorisation */
    CHAMAMD_KERNEL_POTRF  = 1,   /* diagonal tile POTRF                             */
    CHAMAMD_KERNEL_TRSM   = 2,   /* panel TRSM                                      */
    CHAMAMD_KERNEL_SOLVE  = 3,   /* triangular-solve GEMMs (forward + backward)     */
    CHAMAMD_KERNEL_NCLASS = 4
};

/* Accumulated launches, summed device time (ms) and algorithmic flops
 * (Chameleon's flop model, include/chameleon/flops.h) of one kernel class
 * since the last reset.  Results of a sequence are folded in when it is
 * waited on.  Returns CHAMELEON_SUCCESS or CHAMELEON_ERR_ILLEGAL_VALUE. */
int CHAMAMD_Kernel_Stats(int kernel_class, long long *launches, double *total_ms, double *flops);
int CHAMAMD_Kernel_Stats_Reset(void);

#ifdef __cplusplus
}
#endif

#endif
