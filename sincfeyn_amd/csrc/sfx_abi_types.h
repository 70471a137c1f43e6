/* sfx_abi_types.h -- plain-C types shared by the C ABI (include/sfx.h), the device runtime and the
 * emitted per-graph integrand translation units. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One lattice point or cell as stored in HBM: 16 bytes, one dwordx4 load.  k[0..D-1] are the integer
 * coordinates (z_i = k_i h, |k_i| <= K <= 127), fl holds the reference's PFLAG byte (reference
 * driver/sq.h:78-82, sum.c:69-73): bit 0 "children may skip the |.| sums", bits 1-2 the shadow tier. */
#define SFX_MAXD 14
typedef struct { int8_t k[SFX_MAXD]; uint8_t pad; uint8_t fl; } sfx_cell_t;

/* Per-job constants of the point evaluation (reference driver/eval.c:7-24, lattice.c:8-34). */
typedef struct {
    const double *EX;        /* device: e^{k h}, k = -K..K (2K+1 doubles, computed on the host with exp()) */
    int K;                   /* lattice half-width */
    int abskip;              /* ABSKIP: honour the skip bit (eval.c:18) */
    double H;                /* Sinc step h */
    double abserr;           /* fallback threshold ABSERR (eval.c:7, sum.c:651) */
    double nearf;            /* NEARF (eval.c:18) */
    double ddslack;          /* DDSLACK (eval.c:17) */
    int dropunres;           /* DROPUNRES (eval.c:24) */
    int abcheck;             /* ABCHECK: skip-class points evaluated and tested in full (eval.c:11-12, 35) */
    double *raw_ab;          /* device, or NULL.  Set: the evaluation applies no test -- it writes f w (untested)
                                as the value and Sum|terms| w here, at the point's slot (0 where skipped or
                                outside the box): the outputs of the reference's GPU kernel (gpu_kernel.h:22-36),
                                whose caller applies eval.c's tests on the host */
} sfx_eval_params_t;

/* Per-child flag byte written by the evaluation kernels. */
#define SFX_FL_VALID 1u      /* a lattice node (inside the box) */
#define SFX_FL_NEAR  2u      /* "near" the fallback (eval.c:37) */
#define SFX_FL_DD    4u      /* needs the double-double recompute */
#define SFX_FL_SKIP  8u      /* evaluated without the |.| sums */
#define SFX_FL_UNRES 16u     /* unresolved even in double-double (set by the fallback kernel) */
#define SFX_FL_SHELL1 32u    /* node lists at step st > 0: max_d |k_d| == K       (the e1 shell, sum.c:178) */
#define SFX_FL_SHELL2 64u    /* node lists at step st > 0: max_d |k_d| == K - st  (the e2 shell) */
#define SFX_FL_SKIPCLS 128u  /* node lists: the node is of the skip class (evaluated without the |.| sums unless ABCHECK) */

#ifdef __cplusplus
}
#endif
