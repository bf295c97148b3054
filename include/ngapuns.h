/* The unstructured (IJ) linear solver of NGA2, hypre_uns
 * (src/solver/hypre_uns_class.f90), on the GPU (libngapuns.so).
 *
 * Cells with |opr(1)| <= 10 eps at init do not take part (prep_umap,
 * hypre_uns_class.f90:562-607); the others form a CSR matrix, one row per cell
 * in i-fastest order, columns from the stencil with the entries to
 * non-participating cells dropped (:186-230). The methods and their numbers
 * are hypre_uns's (:17-24); the multigrid is BoomerAMG as hypre_uns configures
 * it (:27-33, :256-266): classical strength 0.25, PMIS coarsening, extended+i
 * interpolation with at most 4 weights per row, Galerkin coarse operators,
 * one smoothing sweep before and after the coarse correction. The hierarchy is
 * built on the host at setup; every cycle, product and Krylov step runs on
 * the GPU. The smoother is l1-Jacobi (HYPRE's relax type 18) in place of the
 * hybrid Gauss-Seidel hypre_uns asks for, which is sequential within a
 * process. pcg_parasail and gmres_pilut are not provided.
 *
 * Layout: opr(nst, imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_) and the
 * ghosted rhs/sol, as for ngapsolve.h; return codes are those of ngapsolve.h. */
#ifndef NGAPUNS_H
#define NGAPUNS_H

#include "ngapsolve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hypre_uns_class.f90:17-24 */
#define NGAP_UNS_AMG 1
#define NGAP_UNS_PCG_AMG 2
#define NGAP_UNS_GMRES_AMG 3
#define NGAP_UNS_BICGSTAB_AMG 4
#define NGAP_UNS_PCG 5
#define NGAP_UNS_PCG_PARASAIL 6
#define NGAP_UNS_GMRES 7
#define NGAP_UNS_GMRES_PILUT 8

typedef struct ngap_uns ngap_uns;
typedef struct ngap_amg ngap_amg;

/* Build-specific knobs; hypre_uns's values are the defaults (ngap_uns_default_opts). */
typedef struct ngap_uns_opts {
    double strong_threshold; /* 0.25 (amg_strong_threshold, :29) */
    double max_row_sum;      /* 0.9 (BoomerAMG default) */
    int pmax;                /* interpolation weights kept per row, 4 (BoomerAMG default) */
    int max_coarse;          /* coarsening stops at <= this many rows; dense inverse there, 512
                                (at most 4096; NGAP_ERR_COARSE when coarsening stalls above 4096 rows) */
    int max_levels;          /* 25 (BoomerAMG default) */
    int nu;                  /* smoothing sweeps before and after the coarse correction, 1 (:264) */
    int kdim;                /* GMRES restart length, 5 (gmres_kdim, :28) */
    int graphs;              /* replay the V-cycle as a HIP graph, 1 */
    int verbose;             /* 1: one line per setup and solve on stderr */
    int precision;           /* NGAP_FP64 (default), or NGAP_MIXED: the multigrid cycle in FP32 under the
                                FP64 Krylov iteration (PCG then takes the flexible beta); no counterpart
                                in hypre_uns */
} ngap_uns_opts;

typedef struct ngap_uns_stats {
    int status;           /* NGAP_STATUS_* */
    int it;               /* iterations of the last solve */
    int nlevels;          /* multigrid levels */
    int coarse_singular;  /* the coarsest operator's rows sum to zero */
    long long rows;       /* participating cells */
    long long nnz;        /* fine-level nonzeros */
    double op_complexity; /* sum of nonzeros over levels / fine nonzeros */
    double rerr;          /* ||b - A x|| / ||b|| at exit (true residual) */
    double aerr;          /* ||b - A x|| */
    double setup_ms;      /* wall time of the last setup (host hierarchy + upload) */
    double hier_ms;       /* host time of the hierarchy inside it */
    double solve_ms;      /* wall time of the last solve */
    double vcycle_ms;     /* one V-cycle, from HIP events on the last solve */
} ngap_uns_stats;

int ngap_uns_default_opts(ngap_uns_opts *o);

/* hypre_uns_from_args (:84-119): stc(nst,3) column-major; method 1..8. */
int ngap_uns_create(ngap_uns **h, const ngap_grid *g, int nst, const int *stc, int method);

/* hypre_uns_init (:124-149): the cell numbering from opr(1,:,:,:) as it is
 * now (prep_umap, :562-607). */
int ngap_uns_init(ngap_uns *h, const double *opr);

/* hypre_uns_setup (:153-426): the matrix from opr and the solver's hierarchy.
 * NGAP_ERR_METHOD for pcg_parasail and gmres_pilut; NGAP_ERR_ARG for knobs out
 * of range or a level with more than 2^31-1 nonzeros; NGAP_ERR_ALLOC when a
 * host or device allocation fails. */
int ngap_uns_setup(ngap_uns *h, const double *opr, int maxit, double rcvg, const ngap_uns_opts *o);

/* hypre_uns_solve (:430-500): rhs and the guess from the participating cells
 * of the ghosted arrays; the solution is written to those cells only (the
 * caller syncs the ghosts, as :497 does). */
int ngap_uns_solve(ngap_uns *h, const double *rhs, double *sol, int *it, double *rerr, double *aerr);
/* The same with device pointers (ghosted layout). */
int ngap_uns_solve_device(ngap_uns *h, const double *d_rhs, double *d_sol, int *it, double *rerr, double *aerr);

/* One V-cycle of the hierarchy from a zero guess, x = M b, on compact vectors
 * of ngap_uns_rows() entries (device pointers). */
int ngap_uns_vcycle_device(ngap_uns *h, const double *d_b, double *d_x);

/* hypre_uns_destroy (:504-544) and the handle. */
int ngap_uns_destroy(ngap_uns *h);

long long ngap_uns_rows(const ngap_uns *h);
/* The cell numbering over the ghosted box (ind, :47-49), -1 where a cell
 * does not take part. */
int ngap_uns_index(const ngap_uns *h, int *ind);
int ngap_uns_last_stats(const ngap_uns *h, ngap_uns_stats *s);
/* The handle's hierarchy (valid until the next setup or destroy). */
const ngap_amg *ngap_uns_hierarchy(const ngap_uns *h);

/* Host-side pieces, usable without a GPU. */
/* prep_umap: ind over the ghosted box, returns the participating count (or < 0). */
long long ngap_uns_umap_host(const ngap_grid *g, const double *opr, int nst, int *ind);
/* The CSR rows of :186-230 (with repeated columns merged). Pass col/val NULL
 * to get the nonzero count only. */
long long ngap_uns_assemble_host(const ngap_grid *g, int nst, const int *stc, const int *ind, const double *opr,
                                 int *rowptr, int *col, double *val);
/* The multigrid hierarchy of a CSR matrix (n rows). */
int ngap_amg_build(ngap_amg **H, int n, const int *rowptr, const int *col, const double *val,
                   const ngap_uns_opts *o);
int ngap_amg_nlevels(const ngap_amg *H);
/* Level l: rows, nonzeros, and (l < nlevels-1) the interpolation's nonzeros. */
int ngap_amg_level_sizes(const ngap_amg *H, int l, int *n, int *nnz, int *pnnz);
/* Copies of level l: A (rowptr n+1, col/val nnz), the C/F split (1 C, -1 F)
 * and P (rowptr n+1, col/val pnnz); any pointer may be NULL. */
int ngap_amg_level(const ngap_amg *H, int l, int *rowptr, int *col, double *val, int *cf, int *prowptr, int *pcol,
                   double *pval);
void ngap_amg_free(ngap_amg *H);

#ifdef __cplusplus
}
#endif

#endif
