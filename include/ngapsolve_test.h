/*
 * ngapsolve_test.h — per-kernel hooks of libngapsolve.so.
 *
 * No counterpart in the reference (HYPRE exposes no such entry points). They
 * exist so parity tests can drive each GPU kernel through the C ABI on host
 * buffers and compare it with the CPU oracle, and so the benchmark can time
 * one kernel on the handle's own stream with HIP events. All arrays are host
 * arrays of the named level's size (x fastest), FP64; for prec == 1 they are
 * converted to FP32 on the device, the FP32 kernel runs, and results are
 * converted back. A handle must have been set up first.
 */
#ifndef NGAPSOLVE_TEST_H
#define NGAPSOLVE_TEST_H

#include "ngapsolve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids for ngapt_time_kernel */
#define NGAPT_K_SPMV 0          /* fine-level 7-point SpMV (+ fused dot in FP64) */
#define NGAPT_K_RBGS 1          /* fine-level fused red+black Gauss-Seidel sweep (one launch) */
#define NGAPT_K_RESTRICT 2      /* fine-level residual, then restriction to level 1 (two launches) */
#define NGAPT_K_PROLONG 3       /* level 1 -> fine prolongation + correction */
#define NGAPT_K_UPDATE 4        /* PCG x += alpha p, r -= alpha q, ||r||^2 */
#define NGAPT_K_EXACT 5         /* residual with the exact 7-array operator */
#define NGAPT_K_SWEEP2 6        /* fine-level fused pair of sweeps (one launch) */
#define NGAPT_K_SWEEP2_ZERO 7   /* the same pair from a zero iterate (no v in) */

int ngapt_nlevels(const ngap_t *h);
/* Makes the n-th next device allocation of the handle fail with
 * NGAP_ERR_ALLOC (n = 0: none), so a test can check that a setup which ran
 * out of memory part-way leaves the handle usable by the next setup. */
int ngapt_fail_alloc(ngap_t *h, int n);
int ngapt_level_info(const ngap_t *h, int l, int *dims3, int *cdir);
int ngapt_level_coefs(const ngap_t *h, int l, double *c, double *ax, double *ay, double *az, double *pm, double *pp);
int ngapt_apply(ngap_t *h, int l, int prec, const double *x, double *y);
/* one symmetric Gauss-Seidel sweep; order bit 0: colour updated first
 * (0 red, 1 black); bit 1: start from a zero iterate (v ignored on input);
 * bit 2: two sweeps of that order in one launch (the post-smoothing pair);
 * bits 1 and 2: the first two sweeps from a zero iterate in one launch;
 * without bits 1 and 2, bit 3: the 64-wide tiled sweep kernel also on levels
 * whose rows fit one workgroup row, and bit 4: the sweep kernel itself from a
 * zero iterate */
int ngapt_rbgs(ngap_t *h, int l, int prec, const double *f, double *v, int order);
/* residual + restriction, prolongation + correction of level l; prec bit 0:
 * FP32, bit 1: the one-cell-per-thread transfer kernels on every level (the
 * solver takes the 16-byte vector form where a level allows it) */
int ngapt_resid_restrict(ngap_t *h, int l, int prec, const double *f, const double *v, double *fc);
int ngapt_prolong_add(ngap_t *h, int l, int prec, const double *vc, double *v);
int ngapt_coarse_solve(ngap_t *h, int prec, const double *f, double *v);
int ngapt_vcycle(ngap_t *h, int prec, const double *r, double *z);
int ngapt_apply_exact(ngap_t *h, const double *x, double *y);
int ngapt_dot(ngap_t *h, long long n, const double *x, const double *y, double *out);
/* Times `reps` back-to-back launches of one fine-level kernel with HIP events
 * on the handle's stream. avg_ms = mean launch duration; bytes = algorithmic
 * bytes one launch moves (each array touched once). */
int ngapt_time_kernel(ngap_t *h, int kernel_id, int prec, int reps, double *avg_ms, double *bytes);

#ifdef __cplusplus
}
#endif

#endif
