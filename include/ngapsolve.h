/*
 * ngapsolve.h — C ABI of the MI355X pressure Poisson solver.
 *
 * Drop-in for NGA2's `linsol` plugin as served by `hypre_str` with method
 * `pcg_pfmg` (PCG preconditioned by one PFMG-style semicoarsening V(1,1)
 * cycle). Plain C, no torch types; Fortran binds it with iso_c_binding
 * (nga2_amd/fortran/ngap_c.f90), Python with ctypes (nga2_amd/_lib.py).
 *
 * Reference interface each entry point replaces (paths in the NGA2 tree):
 *   ngap_create   hypre_str_from_args      src/solver/hypre_str_class.f90:68-103
 *   ngap_init     hypre_str_init           src/solver/hypre_str_class.f90:109-145
 *   ngap_setup    hypre_str_setup          src/solver/hypre_str_class.f90:149-318
 *                 (pcg_pfmg branch :264-284)
 *   ngap_solve    hypre_str_solve          src/solver/hypre_str_class.f90:322-386
 *   ngap_destroy  hypre_str_destroy        src/solver/hypre_str_class.f90:390-422
 *   ngap_last_stats  fields it/rerr/aerr of linsol (src/solver/linsol_class.f90:45-48)
 *                 and hypre_str_log/print (:426-463)
 *
 * Data contract (linsol_class.f90:29-35): the caller owns
 *   opr(nst, imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_)   FP64, stencil fastest
 *   rhs, sol(imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_)   FP64, i fastest
 * (Fortran column-major, ghost layers included). Only interior cells
 * [imin_..imax_]x[jmin_..jmax_]x[kmin_..kmax_] are read or written.
 * Rows with opr(1)==0 (the stencil centre) do not participate: they become
 * identity rows and their rhs and sol are taken as 0 (hypre_str_class.f90:
 * 174-175, 339-345). Periodic directions wrap with the global period
 * (:131-135). Ghost cells of sol are NOT written: the caller syncs them,
 * as hypre_str_solve does with cfg%sync (:380).
 *
 * One handle = one HIP stream on the current device. Calls are synchronous
 * (they return after results are on the host, or in device memory for the
 * _device variants). A handle must not be used from two threads at once.
 *
 * Return codes: 0 = OK; < 0 = misuse or invalid argument; > 0 = runtime
 * error. Non-convergence is not an error (the reference accepts it
 * silently): ngap_solve returns 0 with it == maxit and rerr > rcvg, and
 * ngap_stats.status == NGAP_STATUS_MAXIT.
 */
#ifndef NGAPSOLVE_H
#define NGAPSOLVE_H

#ifdef __cplusplus
extern "C" {
#endif

/* methods: hypre_str_class.f90:17-22 */
#define NGAP_PFMG 1
#define NGAP_PCG_PFMG 2
#define NGAP_GMRES_PFMG 3
#define NGAP_SMG 4
#define NGAP_PCG_SMG 5
#define NGAP_GMRES_SMG 6

/* precision modes */
#define NGAP_FP64 0  /* FP64 PCG + FP64 V-cycle (the plain method) */
#define NGAP_MIXED 1 /* FP64 flexible PCG + FP32 V-cycle, FP64 coarsest solve */

/* return codes */
#define NGAP_OK 0
#define NGAP_ERR_ARG (-1)         /* null pointer, bad size or bound */
#define NGAP_ERR_STENCIL (-2)     /* stencil is not the 7-point set */
#define NGAP_ERR_NOT_INIT (-3)    /* setup before init */
#define NGAP_ERR_NOT_SETUP (-4)   /* solve before setup (hypre_str_class.f90:330) */
#define NGAP_ERR_METHOD (-5)      /* method not provided (hypre_str_class.f90:311) */
#define NGAP_ERR_UNSUPPORTED (-6) /* PCG: periodic direction with an odd cell count >= 3;
                                     fourier3d: a direction of more than one cell not periodic */
#define NGAP_ERR_NOT_CIRCULANT (-7) /* fourier3d: operator not uniform in space */
#define NGAP_ERR_HIP 1            /* a HIP runtime call failed */
#define NGAP_ERR_NONFINITE 2      /* NaN/Inf in opr, rhs or the initial guess */
#define NGAP_ERR_BREAKDOWN 3      /* <p,Ap> <= 0 or non-finite in the FP64 iteration */
#define NGAP_ERR_ALLOC 4          /* device or host allocation failed */
#define NGAP_ERR_COARSE 5         /* coarsest-level factorisation failed */
#define NGAP_ERR_STAGE 6          /* the correction stage returned an error */
#define NGAP_ERR_SINGULAR 7       /* fourier3d: transformed operator near zero or NaN */

/* ngap_stats.status */
#define NGAP_STATUS_CONVERGED 0
#define NGAP_STATUS_MAXIT 1
/* PCG stopped before maxit: the recursive residual found no new value 10%
 * below its best for NGAP_STAG_WINDOW iterations in a row, as when the
 * right-hand side of a singular operator has a component in its null space
 * (a periodic or all-Neumann box whose rhs does not sum to zero). Like
 * MAXIT it is not an error: it and rerr/true_rerr hold the last iterate. */
#define NGAP_STATUS_STAGNATED 2
#define NGAP_STAG_WINDOW 25

typedef struct ngap_t ngap_t;

/* Grid of one rank (single GPU: local == global). Mirrors cfg%nx.., the local
 * interior bounds imin_..kmax_ and the array bounds imino_..kmaxo_. */
typedef struct ngap_grid {
    int nx, ny, nz;
    int imin_, imax_, jmin_, jmax_, kmin_, kmax_;
    int imino_, imaxo_, jmino_, jmaxo_, kmino_, kmaxo_;
    int xper, yper, zper;
} ngap_grid;

/* Build-specific knobs (no counterpart in hypre_str, whose PFMG settings are
 * fixed at :267-272). Fill with ngap_default_opts. */
typedef struct ngap_opts {
    int precision;    /* NGAP_FP64 (default) or NGAP_MIXED */
    int nu_pre;       /* pre-smoothing RB sweeps per level; 0 (default): 1 in FP64, 2 in NGAP_MIXED */
    int nu_post;      /* post-smoothing BR sweeps per level; 0 (default): 1 in FP64, 2 in NGAP_MIXED */
    int coarsest_max; /* stop coarsening at <= this many cells, default 128 */
    int check_finite; /* scan opr/rhs/guess for NaN/Inf, default 1 */
    int verbose;      /* 0 quiet, 1 one line per solve on stderr */
    int graphs;       /* replay the V-cycle as a HIP graph, default 1 */
    int kdim;         /* GMRES restart length (HYPRE_StructGMRESSetKDim; HYPRE's default 5) */
    int skip_relax;   /* 1 (default): on an isotropic fine operator (at most 10% of its cells with
                         one direction's couplings more than twice another's), smooth only the coarse
                         levels whose couplings are isotropic (max over the directions of sum |a_d|
                         <= 2 x min); the others restrict and interpolate without smoothing, after
                         HYPRE PFMG's SkipRelax (on by default there). 0: smooth every level */
} ngap_opts;

typedef struct ngap_stats {
    int status;          /* NGAP_STATUS_* */
    int it;              /* PCG iterations */
    int nlevels;         /* multigrid levels including the fine level */
    int coarse_singular; /* coarsest operator had vanishing row sums */
    int precision;       /* mode actually used by the last solve */
    int fallback;        /* bit 0: a mixed solve fell back to the FP64 V-cycle;
                            bit 1: the correction stage was dropped (guess rejected or breakdown) */
    int restarts;        /* residual replacements */
    int stage_calls;     /* calls of the correction stage (ngap_set_stage) in the last solve */
    int skip_relax;      /* 1: the last setup skipped relaxation on some levels (ngap_opts.skip_relax) */
    int relaxed_levels;  /* levels the V-cycle smooths (the coarsest, solved, included) */
    double rerr;         /* relative 2-norm of the recursive residual at exit */
    double true_rerr;    /* ||b - A x||_2/||b||_2 with the exact 7-array opr */
    double aerr;         /* ||b - A x||_2 */
    double cnorm_ratio;  /* sqrt(<r,Mr>/<r0,Mr0>) of the last iterate (HYPRE-style) */
    double asym;         /* max |opr(+d,p) - opr(-d,p+e_d)| / |opr(1,p)| seen at setup */
    double setup_ms;     /* wall time of the last ngap_setup* */
    double solve_ms;     /* wall time of the last ngap_solve* */
    double copy_ms;      /* host<->device copies inside the last ngap_solve */
    long long cells;     /* fine-level cells */
    long long device_bytes; /* device memory held by the handle */
} ngap_stats;

/* Fills *o with the defaults. Always returns 0. */
int ngap_default_opts(ngap_opts *o);

/* Replaces hypre_str_from_args (hypre_str_class.f90:68-103). stc is the
 * Fortran array stc(nst,3) in column-major order. Methods NGAP_PCG_PFMG and
 * NGAP_GMRES_PFMG are provided (others -> NGAP_ERR_METHOD at ngap_setup).
 * NGAP_GMRES_PFMG (:286-303; the implicit velocity solve of
 * incomp_class.f90:1756-1803) is flexible GMRES(kdim), right-preconditioned
 * by one V-cycle built on the symmetric part (A + A^T)/2 of the operator,
 * stopping on the true relative residual ||b - A x||/||b|| <= rcvg. */
int ngap_create(ngap_t **h, const ngap_grid *g, int nst, const int *stc, int method);

/* Replaces hypre_str_init (:109-145): checks the stencil map and the grid. */
int ngap_init(ngap_t *h);

/* Replaces hypre_str_setup (:149-318): deep-copies opr to the device,
 * builds the multigrid hierarchy. maxlevel > 0 caps the number of levels
 * (HYPRE_StructPFMGSetMaxLevels, :272). o may be NULL for defaults.
 * Calling it again rebuilds (the reference destroys first, :158). */
int ngap_setup(ngap_t *h, const double *opr, int maxit, double rcvg, int maxlevel, const ngap_opts *o);

/* Same as ngap_setup with opr already in device memory (same layout).
 * No counterpart in the reference: for callers that assemble the operator
 * on the GPU. The call synchronizes the device before reading d_opr, so the
 * producer may run on any stream. It reads the interior cells' entries once
 * (ghost entries are not read) and keeps its own copy of them, so d_opr may
 * change after the call. */
int ngap_setup_device(ngap_t *h, const double *d_opr, int maxit, double rcvg, int maxlevel, const ngap_opts *o);

/* Replaces hypre_str_solve (:322-386). sol holds the initial guess on entry
 * and the solution on exit (interior only). it/rerr/aerr may be NULL. */
int ngap_solve(ngap_t *h, const double *rhs, double *sol, int *it, double *rerr, double *aerr);

/* Same as ngap_solve with rhs and sol in device memory (same layout).
 * No counterpart in the reference. The device is synchronized before d_rhs
 * and d_sol are read, and the handle's stream before the call returns. */
int ngap_solve_device(ngap_t *h, const double *d_rhs, double *d_sol, int *it, double *rerr, double *aerr);

/* res = rhs - A sol with the exact operator of the last setup, for rhs and
 * sol in device memory (ghosted layout, as ngap_solve_device) and res a
 * device vector of the nx*ny*nz interior cells, x fastest. *rnorm (host,
 * may be NULL) = ||res||_2. No counterpart in the reference: it produces the
 * inputs of a correction stage (training pairs of a learned guess). */
int ngap_residual_device(ngap_t *h, const double *d_rhs, const double *d_sol, double *d_res, double *rnorm);

/* y = A x for device vectors of the nx*ny*nz interior cells (x fastest),
 * with the fine-level FP64 operator the PCG iterates with. No counterpart in
 * the reference: it supplies the residual-norm loss of a learned stage. */
int ngap_apply_device(ngap_t *h, const double *d_x, double *d_y);

/* Changes maxit/rcvg without rebuilding the hierarchy. */
int ngap_set_tolerance(ngap_t *h, int maxit, double rcvg);

/* Replaces hypre_str_destroy (:390-422) and frees the handle. */
int ngap_destroy(ngap_t *h);

/* Statistics of the last setup/solve (replaces reading it/rerr/aerr and the
 * log/print routines, :426-463). */
int ngap_last_stats(const ngap_t *h, ngap_stats *s);

/* Correction stage: fn(ctx, r, e, n, stream) writes e ~ A^{-1} r for r and e
 * device FP64 vectors of the n interior cells (x fastest), launching its work
 * on `stream` (the handle's hipStream_t) without synchronizing it.
 * No counterpart in the reference (hypre_str has no learned stage). flags:
 *   NGAP_STAGE_GUESS    x0 += fn(b - A x0) before the first iteration; the
 *                       guess is kept only if it lowers ||b - A x0||
 *   NGAP_STAGE_PRECOND  in the first precond_iters iterations the
 *                       preconditioner is z = V(r) + fn(r - A V(r)), with
 *                       flexible (Polak-Ribiere) PCG in either precision.
 * A breakdown with the stage active drops the stage and restarts. The
 * stopping rule is unchanged, so every solve still meets rcvg with the exact
 * operator. fn == NULL removes the stage. */
typedef int (*ngap_stage_fn)(void *ctx, const double *r, double *e, long long n, void *stream);
#define NGAP_STAGE_GUESS 1
#define NGAP_STAGE_PRECOND 2
int ngap_set_stage(ngap_t *h, ngap_stage_fn fn, void *ctx, int flags, int precond_iters);

/* Static message for a return code. */
const char *ngap_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif
