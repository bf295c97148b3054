/*
 * FP64 CPU oracle for the pressure Poisson solve: PCG preconditioned by one
 * PFMG-style semicoarsening V(1,1) cycle.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product path (nga2_amd, libngapsolve.so) never links or calls it.
 *
 * What it restates (reference paths relative to the NGA2 tree):
 *  - hypre_str%setup, method pcg_pfmg (src/solver/hypre_str_class.f90:264-284):
 *    PCG(maxit, rcvg) preconditioned by one PFMG cycle from a zero guess,
 *    RelaxType 2 (red/black Gauss-Seidel, RB before and BR after the coarse
 *    correction, hence a symmetric preconditioner), RAPType 1 (7-point
 *    coarse operators), MaxLevels = maxlevel when maxlevel > 0.
 *  - hypre_str%setup :171-179 — rows with opr(1)==0 become identity rows.
 *  - hypre_str%solve :336-348 — rhs and guess are zeroed on those rows.
 *  - HYPRE itself is not in the reference tree (tools/GNUMake/packages/
 *    Make.hypre:1-8 only links -lHYPRE, no version pinned), so the multigrid
 *    below is this project's own PFMG-style design; its coarsening rule,
 *    interpolation weights and coarse operators are stated here and in the
 *    GPU code identically so both build the same hierarchy.
 *
 * Algorithm (shared with the GPU library):
 *  Level-0 operator is stored symmetrically: diagonal c and, per direction
 *  d in {x,y,z}, the coupling a_d(p) between cell p and its -d neighbour.
 *  The +d coupling of p is a_d(p+e_d). a_d(p) = 0.5*(opr(-d,p)+opr(+d,p-e_d))
 *  when both cells participate and the neighbour exists, else 0.
 *  Directions with n_d == 1 are folded: periodic self couplings go into the
 *  diagonal, non-periodic ones are dropped (HYPRE ignores off-grid entries).
 *
 *  Coarsening: one direction per level (semicoarsening). Candidates are the
 *  directions with n_d a multiple of 4 (periodic: the coarse length stays
 *  even, as red/black colouring needs across the wrap) or n_d >= 2
 *  (non-periodic);
 *  the candidate with the largest sum |a_d| wins (ties within 1e-6 relative
 *  go to the lower index). Coarse point J along d is fine point 2J; fine odd
 *  points o = 2J+1 interpolate from J and J+1 with operator-dependent weights
 *      den = c(o) + sum_t (a_t^-(o) + a_t^+(o)),  Pm = -a_d^-(o)/den,
 *      Pp = -a_d^+(o)/den   (0 when den <= 0),
 *  restriction R = P^T. The coarse operator is 7-point:
 *      a_d^c(J)  = -a_d^-(o) a_d^+(o) / den(o),  o = 2J-1   (1-D Galerkin)
 *      a_t^c(J)  = a_t(e) + 0.5 (Pm(J)+Pm(J-e_t)) a_t(o_r)
 *                        + 0.5 (Pp(J-e_d)+Pp(J-e_d-e_t)) a_t(o_l)
 *      c^c(J)    = s^c(J) - sum of the 6 coarse couplings of J,
 *      s^c(J)    = s(e) + Pm(J) s(o_r) + Pp(J-e_d) s(o_l),  s = row sum
 *  which keeps symmetry, the constant null space of periodic/Neumann
 *  problems and the M-matrix sign pattern.
 *  Coarsening stops at maxlevel levels, at <= coarsest_max cells or when no
 *  direction can be coarsened. The coarsest level is solved exactly by dense
 *  Cholesky (rank-one regularised when its row sums vanish, i.e. singular)
 *  when it has <= 1024 cells, otherwise by 8 symmetric RB/BR sweeps.
 *
 *  Red cells have (i+j+k) even. Periodic directions of odd length >= 3 are
 *  rejected (red/black colouring is inconsistent across the wrap).
 *
 *  PCG (plain): beta = <r+,z+>/<r,z>; flexible variant (used to check the
 *  mixed-precision GPU path): beta = -alpha <z+,q>/<r,z> (Polak-Ribiere).
 *  Stop when ||r||_2/||b||_2 <= rcvg on the recursive residual; then the
 *  true residual with the exact 7-array operator is computed; if it misses
 *  rcvg the residual is replaced and iteration continues.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXLEV 64

typedef struct {
    int n[3];
    int per[3];
    long N;
    int cdir; /* direction coarsened to build level l+1, -1 on coarsest */
    double *c, *a[3];
    double *pm, *pp; /* interpolation weights (for the transfer from level l-1), coarse-indexed */
    double *f, *v, *t;
} lvl_t;

typedef struct {
    int nlev;
    lvl_t L[ORC_MAXLEV];
    double *opr7; /* exact operator, [N][7], identity rows applied */
    unsigned char *part;
    int nu_pre, nu_post;
    int coarse_dense; /* 1: dense Cholesky on coarsest */
    int coarse_singular;
    double *chol; /* n x n lower factor, row-major */
    int nc;
    double *bufs[8];
    unsigned char relax[ORC_MAXLEV]; /* 1: level l is smoothed in the V-cycle (see orc_set_skip_relax) */
} orc_t;

static inline long idx3(const int *n, int i, int j, int k) { return (long)i + (long)n[0] * ((long)j + (long)n[1] * k); }

/* neighbour coordinate along a direction: wraps when periodic, -1 when absent */
static inline int nbm(int i, int n, int per) { return i > 0 ? i - 1 : (per ? n - 1 : -1); }
static inline int nbp(int i, int n, int per) { return i < n - 1 ? i + 1 : (per ? 0 : -1); }

static void *xcalloc(long n, long sz) { return calloc((size_t)(n > 0 ? n : 1), (size_t)sz); }

/* ---------------------------------------------------------------- operator */

/* y = A_l x (symmetric 4-array storage) */
static void lvl_apply(const lvl_t *L, const double *x, double *y) {
    const int *n = L->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++) {
        int km = nbm(k, n[2], L->per[2]), kp = nbp(k, n[2], L->per[2]);
        for (int j = 0; j < n[1]; j++) {
            int jm = nbm(j, n[1], L->per[1]), jp = nbp(j, n[1], L->per[1]);
            for (int i = 0; i < n[0]; i++) {
                int im = nbm(i, n[0], L->per[0]), ip = nbp(i, n[0], L->per[0]);
                long p = idx3(n, i, j, k);
                double s = L->c[p] * x[p];
                if (im >= 0) s += L->a[0][p] * x[idx3(n, im, j, k)];
                if (ip >= 0) { long q = idx3(n, ip, j, k); s += L->a[0][q] * x[q]; }
                if (jm >= 0) s += L->a[1][p] * x[idx3(n, i, jm, k)];
                if (jp >= 0) { long q = idx3(n, i, jp, k); s += L->a[1][q] * x[q]; }
                if (km >= 0) s += L->a[2][p] * x[idx3(n, i, j, km)];
                if (kp >= 0) { long q = idx3(n, i, j, kp); s += L->a[2][q] * x[q]; }
                y[p] = s;
            }
        }
    }
}

/* y = A x with the exact 7-coefficient operator as handed over
 * (stencil order: centre, +x, -x, +y, -y, +z, -z). Couplings pointing out of a
 * non-periodic box are ignored, as HYPRE's struct matrix does. */
static void exact_apply(const orc_t *o, const double *x, double *y) {
    const lvl_t *L = &o->L[0];
    const int *n = L->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++) {
        int km = nbm(k, n[2], L->per[2]), kp = nbp(k, n[2], L->per[2]);
        for (int j = 0; j < n[1]; j++) {
            int jm = nbm(j, n[1], L->per[1]), jp = nbp(j, n[1], L->per[1]);
            for (int i = 0; i < n[0]; i++) {
                int im = nbm(i, n[0], L->per[0]), ip = nbp(i, n[0], L->per[0]);
                long p = idx3(n, i, j, k);
                const double *A = o->opr7 + 7 * p;
                double s = A[0] * x[p];
                if (ip >= 0) s += A[1] * x[idx3(n, ip, j, k)];
                if (im >= 0) s += A[2] * x[idx3(n, im, j, k)];
                if (jp >= 0) s += A[3] * x[idx3(n, i, jp, k)];
                if (jm >= 0) s += A[4] * x[idx3(n, i, jm, k)];
                if (kp >= 0) s += A[5] * x[idx3(n, i, j, kp)];
                if (km >= 0) s += A[6] * x[idx3(n, i, j, km)];
                y[p] = s;
            }
        }
    }
}

/* deterministic dot: per-plane partials summed in plane order */
static double dotp(long nplane, long plane, const double *x, const double *y, double *part) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nplane; k++) {
        double s = 0.0;
        const double *a = x + k * plane, *b = y + k * plane;
        for (long p = 0; p < plane; p++) s += a[p] * b[p];
        part[k] = s;
    }
    double s = 0.0;
    for (long k = 0; k < nplane; k++) s += part[k];
    return s;
}

/* ---------------------------------------------------------------- smoother */

/* one Gauss-Seidel half sweep over cells of colour `col` ((i+j+k)&1 == col) */
static void gs_colour(const lvl_t *L, const double *f, double *v, int col) {
    const int *n = L->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++) {
        int km = nbm(k, n[2], L->per[2]), kp = nbp(k, n[2], L->per[2]);
        for (int j = 0; j < n[1]; j++) {
            int jm = nbm(j, n[1], L->per[1]), jp = nbp(j, n[1], L->per[1]);
            int i0 = (col + j + k) & 1;
            for (int i = i0; i < n[0]; i += 2) {
                int im = nbm(i, n[0], L->per[0]), ip = nbp(i, n[0], L->per[0]);
                long p = idx3(n, i, j, k);
                double s = f[p];
                if (im >= 0) s -= L->a[0][p] * v[idx3(n, im, j, k)];
                if (ip >= 0) { long q = idx3(n, ip, j, k); s -= L->a[0][q] * v[q]; }
                if (jm >= 0) s -= L->a[1][p] * v[idx3(n, i, jm, k)];
                if (jp >= 0) { long q = idx3(n, i, jp, k); s -= L->a[1][q] * v[q]; }
                if (km >= 0) s -= L->a[2][p] * v[idx3(n, i, j, km)];
                if (kp >= 0) { long q = idx3(n, i, j, kp); s -= L->a[2][q] * v[q]; }
                v[p] = s / L->c[p];
            }
        }
    }
}

/* ---------------------------------------------------------------- transfers */

/* coordinates of a coarse cell's fine even point and its odd neighbours */
typedef struct { int fe[3]; int have_l, have_r; int ol[3], orr[3]; } tmap;

static inline void coarse_map(const lvl_t *F, int d, const int *J, tmap *m) {
    for (int q = 0; q < 3; q++) m->fe[q] = J[q];
    m->fe[d] = 2 * J[d];
    for (int q = 0; q < 3; q++) { m->ol[q] = m->fe[q]; m->orr[q] = m->fe[q]; }
    int l = nbm(m->fe[d], F->n[d], F->per[d]);
    int r = (m->fe[d] + 1 < F->n[d]) ? m->fe[d] + 1 : -1; /* periodic n even: never wraps */
    m->have_l = l >= 0; m->ol[d] = l;
    m->have_r = r >= 0; m->orr[d] = r;
}

/* fc = R (f - A v) on level l -> l+1 */
static void resid_restrict(const orc_t *o, int l, const double *f, const double *v, double *fc) {
    const lvl_t *F = &o->L[l], *C = &o->L[l + 1];
    int d = F->cdir;
    double *r = F->t;
    lvl_apply(F, v, r);
    for (long p = 0; p < F->N; p++) r[p] = f[p] - r[p];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                int J[3] = {i, j, k};
                tmap m; coarse_map(F, d, J, &m);
                long pc = idx3(C->n, i, j, k);
                double s = r[idx3(F->n, m.fe[0], m.fe[1], m.fe[2])];
                if (m.have_r) s += C->pm[pc] * r[idx3(F->n, m.orr[0], m.orr[1], m.orr[2])];
                if (m.have_l) {
                    int Jl[3] = {i, j, k};
                    Jl[d] = nbm(J[d], C->n[d], C->per[d]);
                    s += C->pp[idx3(C->n, Jl[0], Jl[1], Jl[2])] * r[idx3(F->n, m.ol[0], m.ol[1], m.ol[2])];
                }
                fc[pc] = s;
            }
}

/* v += P vc on level l from l+1 */
static void prolong_add(const orc_t *o, int l, const double *vc, double *v) {
    const lvl_t *F = &o->L[l], *C = &o->L[l + 1];
    int d = F->cdir;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < F->n[2]; k++)
        for (int j = 0; j < F->n[1]; j++)
            for (int i = 0; i < F->n[0]; i++) {
                int I[3] = {i, j, k};
                long p = idx3(F->n, i, j, k);
                int fd = I[d];
                int J[3] = {i, j, k};
                if ((fd & 1) == 0) {
                    J[d] = fd / 2;
                    v[p] += vc[idx3(C->n, J[0], J[1], J[2])];
                } else {
                    J[d] = (fd - 1) / 2;
                    long pl = idx3(C->n, J[0], J[1], J[2]);
                    double s = C->pm[pl] * vc[pl];
                    int jr = nbp(J[d], C->n[d], C->per[d]);
                    if (jr >= 0 && !(C->per[d] == 0 && J[d] + 1 >= C->n[d])) {
                        int Jr[3] = {J[0], J[1], J[2]}; Jr[d] = jr;
                        s += C->pp[pl] * vc[idx3(C->n, Jr[0], Jr[1], Jr[2])];
                    }
                    v[p] += s;
                }
            }
}

/* ---------------------------------------------------------------- hierarchy */

static double coupling_plus(const lvl_t *L, int d, int i, int j, int k) {
    int c[3] = {i, j, k};
    int q = nbp(c[d], L->n[d], L->per[d]);
    if (q < 0) return 0.0;
    c[d] = q;
    return L->a[d][idx3(L->n, c[0], c[1], c[2])];
}

static double row_sum(const lvl_t *L, int i, int j, int k) {
    long p = idx3(L->n, i, j, k);
    double s = L->c[p];
    for (int d = 0; d < 3; d++) s += L->a[d][p] + coupling_plus(L, d, i, j, k);
    return s;
}

static int can_coarsen(const lvl_t *L, int d) {
    if (L->per[d]) return (L->n[d] % 4) == 0; /* coarse periodic length stays even */
    return L->n[d] >= 2;
}

static int alloc_level(lvl_t *L) {
    L->c = xcalloc(L->N, sizeof(double));
    for (int d = 0; d < 3; d++) L->a[d] = xcalloc(L->N, sizeof(double));
    L->pm = xcalloc(L->N, sizeof(double));
    L->pp = xcalloc(L->N, sizeof(double));
    L->f = xcalloc(L->N, sizeof(double));
    L->v = xcalloc(L->N, sizeof(double));
    L->t = xcalloc(L->N, sizeof(double));
    return (L->c && L->a[0] && L->a[1] && L->a[2] && L->pm && L->pp && L->f && L->v && L->t) ? 0 : -1;
}

static void free_level(lvl_t *L) {
    free(L->c); for (int d = 0; d < 3; d++) free(L->a[d]);
    free(L->pm); free(L->pp); free(L->f); free(L->v); free(L->t);
    memset(L, 0, sizeof(*L));
}

/* pick the coarsening direction, -1 when none */
static int choose_dir(const lvl_t *L) {
    double str[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) {
        if (!can_coarsen(L, d)) { str[d] = -1.0; continue; }
        double s = 0.0;
        for (long p = 0; p < L->N; p++) s += fabs(L->a[d][p]);
        str[d] = s;
    }
    int best = -1;
    for (int d = 0; d < 3; d++) {
        if (str[d] < 0) continue;
        if (best < 0 || str[d] > str[best] * (1.0 + 1e-6)) best = d;
    }
    return best;
}

/* build level l+1 from level l along F->cdir */
static int build_coarse(orc_t *o, int l) {
    lvl_t *F = &o->L[l], *C = &o->L[l + 1];
    int d = F->cdir;
    for (int q = 0; q < 3; q++) { C->n[q] = F->n[q]; C->per[q] = F->per[q]; }
    C->n[d] = F->per[d] ? F->n[d] / 2 : (F->n[d] + 1) / 2;
    C->N = (long)C->n[0] * C->n[1] * C->n[2];
    C->cdir = -1;
    if (alloc_level(C)) return -1;
    double *sc = C->f; /* scratch: coarse row sums */

    /* interpolation weights of the odd point to the right of each coarse cell */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                int J[3] = {i, j, k};
                tmap m; coarse_map(F, d, J, &m);
                long pc = idx3(C->n, i, j, k);
                C->pm[pc] = 0.0; C->pp[pc] = 0.0;
                if (!m.have_r) continue;
                int *o3 = m.orr;
                long po = idx3(F->n, o3[0], o3[1], o3[2]);
                double am = F->a[d][po], ap = coupling_plus(F, d, o3[0], o3[1], o3[2]);
                double den = F->c[po];
                for (int t = 0; t < 3; t++) if (t != d) den += F->a[t][po] + coupling_plus(F, t, o3[0], o3[1], o3[2]);
                if (den > 0.0) { C->pm[pc] = -am / den; C->pp[pc] = -ap / den; }
            }

    /* minus-direction couplings and row sums */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                int J[3] = {i, j, k};
                tmap m; coarse_map(F, d, J, &m);
                long pc = idx3(C->n, i, j, k);
                long pe = idx3(F->n, m.fe[0], m.fe[1], m.fe[2]);
                /* index of the coarse cell left of J along d (owner of the left odd point's weights) */
                long pcl = -1;
                if (m.have_l) {
                    int Jl[3] = {i, j, k}; Jl[d] = nbm(J[d], C->n[d], C->per[d]);
                    pcl = idx3(C->n, Jl[0], Jl[1], Jl[2]);
                }
                /* along d */
                if (m.have_l) {
                    long po = idx3(F->n, m.ol[0], m.ol[1], m.ol[2]);
                    double am = F->a[d][po], ap = coupling_plus(F, d, m.ol[0], m.ol[1], m.ol[2]);
                    double den = F->c[po];
                    for (int t = 0; t < 3; t++) if (t != d) den += F->a[t][po] + coupling_plus(F, t, m.ol[0], m.ol[1], m.ol[2]);
                    C->a[d][pc] = den > 0.0 ? -am * ap / den : 0.0;
                } else {
                    C->a[d][pc] = 0.0;
                }
                /* transverse */
                for (int t = 0; t < 3; t++) {
                    if (t == d) continue;
                    int Jt[3] = {i, j, k};
                    Jt[t] = nbm(J[t], C->n[t], C->per[t]);
                    if (Jt[t] < 0) { C->a[t][pc] = 0.0; continue; }
                    long pct = idx3(C->n, Jt[0], Jt[1], Jt[2]);
                    double s = F->a[t][pe];
                    if (m.have_r) {
                        long po = idx3(F->n, m.orr[0], m.orr[1], m.orr[2]);
                        s += 0.5 * (C->pm[pc] + C->pm[pct]) * F->a[t][po];
                    }
                    if (m.have_l) {
                        long po = idx3(F->n, m.ol[0], m.ol[1], m.ol[2]);
                        int Jlt[3] = {Jt[0], Jt[1], Jt[2]}; Jlt[d] = nbm(J[d], C->n[d], C->per[d]);
                        long pclt = idx3(C->n, Jlt[0], Jlt[1], Jlt[2]);
                        s += 0.5 * (C->pp[pcl] + C->pp[pclt]) * F->a[t][po];
                    }
                    C->a[t][pc] = s;
                }
                /* row sum */
                double s = row_sum(F, m.fe[0], m.fe[1], m.fe[2]);
                if (m.have_r) s += C->pm[pc] * row_sum(F, m.orr[0], m.orr[1], m.orr[2]);
                if (m.have_l) s += C->pp[pcl] * row_sum(F, m.ol[0], m.ol[1], m.ol[2]);
                sc[pc] = s;
            }
    /* periodic direction of length 2 on the coarse level: both couplings of a
     * pair exist (through the wrap and directly) and the kernel sums both. */

    /* diagonal from row sum */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                long pc = idx3(C->n, i, j, k);
                double s = sc[pc];
                for (int t = 0; t < 3; t++) s -= C->a[t][pc] + coupling_plus(C, t, i, j, k);
                C->c[pc] = s;
            }
    memset(C->f, 0, sizeof(double) * C->N);
    return 0;
}

/* dense coarsest operator, Cholesky factor in o->chol */
static int factor_coarsest(orc_t *o) {
    lvl_t *C = &o->L[o->nlev - 1];
    int n = (int)C->N;
    o->nc = n;
    o->chol = xcalloc((long)n * n, sizeof(double));
    if (!o->chol) return -1;
    double *M = o->chol;
    double ssum = 0.0, csum = 0.0;
    for (int k = 0; k < C->n[2]; k++)
        for (int j = 0; j < C->n[1]; j++)
            for (int i = 0; i < C->n[0]; i++) {
                long p = idx3(C->n, i, j, k);
                M[p * n + p] += C->c[p];
                int c3[3] = {i, j, k};
                for (int d = 0; d < 3; d++) {
                    int q = nbm(c3[d], C->n[d], C->per[d]);
                    if (q < 0) continue;
                    int e3[3] = {i, j, k}; e3[d] = q;
                    long pq = idx3(C->n, e3[0], e3[1], e3[2]);
                    M[p * n + pq] += C->a[d][p];
                    M[pq * n + p] += C->a[d][p];
                }
                ssum += fabs(row_sum(C, i, j, k));
                csum += fabs(C->c[p]);
            }
    o->coarse_singular = ssum <= 1e-10 * csum;
    if (o->coarse_singular) {
        double g = csum / ((double)n * n);
        for (long q = 0; q < (long)n * n; q++) M[q] += g;
    }
    /* in-place Cholesky, lower triangle */
    for (int jj = 0; jj < n; jj++) {
        double s = M[(long)jj * n + jj];
        for (int kk = 0; kk < jj; kk++) s -= M[(long)jj * n + kk] * M[(long)jj * n + kk];
        if (!(s > 0.0)) return -2;
        double ljj = sqrt(s);
        M[(long)jj * n + jj] = ljj;
        for (int ii = jj + 1; ii < n; ii++) {
            double t = M[(long)ii * n + jj];
            for (int kk = 0; kk < jj; kk++) t -= M[(long)ii * n + kk] * M[(long)jj * n + kk];
            M[(long)ii * n + jj] = t / ljj;
        }
    }
    for (int ii = 0; ii < n; ii++)
        for (int jj = ii + 1; jj < n; jj++) M[(long)ii * n + jj] = 0.0;
    return 0;
}

static void coarse_solve(const orc_t *o, const double *f, double *v) {
    const lvl_t *C = &o->L[o->nlev - 1];
    if (o->coarse_dense) {
        int n = o->nc;
        const double *M = o->chol;
        for (int i = 0; i < n; i++) {
            double s = f[i];
            for (int k = 0; k < i; k++) s -= M[(long)i * n + k] * v[k];
            v[i] = s / M[(long)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = v[i];
            for (int k = i + 1; k < n; k++) s -= M[(long)k * n + i] * v[k];
            v[i] = s / M[(long)i * n + i];
        }
    } else {
        memset(v, 0, sizeof(double) * C->N);
        for (int s = 0; s < 8; s++) {
            gs_colour(C, f, v, 0); gs_colour(C, f, v, 1);
            gs_colour(C, f, v, 1); gs_colour(C, f, v, 0);
        }
    }
}

/* ---------------------------------------------------------------- V-cycle */

static void vcycle(orc_t *o, const double *r, double *z) {
    int nl = o->nlev;
    /* level 0 uses r/z directly */
    const double *f0 = r;
    for (int l = 0; l < nl - 1; l++) {
        lvl_t *L = &o->L[l];
        const double *f = l == 0 ? f0 : L->f;
        double *v = l == 0 ? z : L->v;
        memset(v, 0, sizeof(double) * L->N);
        if (o->relax[l])
            for (int s = 0; s < o->nu_pre; s++) { gs_colour(L, f, v, 0); gs_colour(L, f, v, 1); }
        resid_restrict(o, l, f, v, o->L[l + 1].f);
    }
    {
        lvl_t *C = &o->L[nl - 1];
        const double *f = nl == 1 ? f0 : C->f;
        double *v = nl == 1 ? z : C->v;
        coarse_solve(o, f, v);
    }
    for (int l = nl - 2; l >= 0; l--) {
        lvl_t *L = &o->L[l];
        const double *f = l == 0 ? f0 : L->f;
        double *v = l == 0 ? z : L->v;
        prolong_add(o, l, o->L[l + 1].v, v);
        if (o->relax[l])
            for (int s = 0; s < o->nu_post; s++) { gs_colour(L, f, v, 1); gs_colour(L, f, v, 0); }
    }
}

/* ---------------------------------------------------------------- PCG core */

typedef void (*matvec_fn)(void *ctx, const double *x, double *y);
typedef double (*dot_fn)(void *ctx, const double *x, const double *y);

typedef struct {
    long n;
    matvec_fn A, M, Aexact;
    dot_fn dot;
    void *ctx;
} pcg_ops;

/* returns iterations; rerr = recursive relative residual at exit,
 * trerr = true relative residual with Aexact */
static int pcg_core(const pcg_ops *op, const double *b, double *x, int maxit, double rcvg, int flexible,
                    double *rerr, double *trerr, double *w /* 4n scratch */) {
    long n = op->n;
    double *r = w, *z = w + n, *p = w + 2 * n, *q = w + 3 * n;
    double bn = sqrt(op->dot(op->ctx, b, b));
    if (bn == 0.0) {
        memset(x, 0, sizeof(double) * n);
        *rerr = 0.0; *trerr = 0.0;
        return 0;
    }
    op->Aexact(op->ctx, x, r);
    for (long i = 0; i < n; i++) r[i] = b[i] - r[i];
    double rel = sqrt(op->dot(op->ctx, r, r)) / bn;
    int it = 0;
    int restart = 1;
    double rho = 0.0;
    while (1) {
        if (rel <= rcvg) {
            /* certify with the exact operator */
            op->Aexact(op->ctx, x, q);
            for (long i = 0; i < n; i++) q[i] = b[i] - q[i];
            double tr = sqrt(op->dot(op->ctx, q, q)) / bn;
            if (tr <= rcvg || it >= maxit) { *rerr = rel; *trerr = tr; return it; }
            memcpy(r, q, sizeof(double) * n);
            rel = tr;
            restart = 1;
        }
        if (it >= maxit) break;
        if (restart) {
            op->M(op->ctx, r, z);
            rho = op->dot(op->ctx, r, z);
            memcpy(p, z, sizeof(double) * n);
            restart = 0;
        }
        op->A(op->ctx, p, q);
        double pq = op->dot(op->ctx, p, q);
        double alpha = rho / pq;
        for (long i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        it++;
        rel = sqrt(op->dot(op->ctx, r, r)) / bn;
        if (rel <= rcvg || it >= maxit) continue;
        op->M(op->ctx, r, z);
        double rho_new = op->dot(op->ctx, r, z);
        double beta;
        if (flexible) beta = -alpha * op->dot(op->ctx, z, q) / rho;
        else beta = rho_new / rho;
        for (long i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rho = rho_new;
    }
    op->Aexact(op->ctx, x, q);
    for (long i = 0; i < n; i++) q[i] = b[i] - q[i];
    *rerr = rel;
    *trerr = sqrt(op->dot(op->ctx, q, q)) / bn;
    return it;
}

/* stencil context */
static void st_A(void *c, const double *x, double *y) { orc_t *o = c; lvl_apply(&o->L[0], x, y); }
static void st_Ax(void *c, const double *x, double *y) { orc_t *o = c; exact_apply(o, x, y); }
static void st_M(void *c, const double *r, double *z) { vcycle((orc_t *)c, r, z); }
static double st_dot(void *c, const double *x, const double *y) {
    orc_t *o = c;
    long plane = (long)o->L[0].n[0] * o->L[0].n[1];
    return dotp(o->L[0].n[2], plane, x, y, o->bufs[0]);
}

/* ================================================================ exports */

void orc_destroy(orc_t *o) {
    if (!o) return;
    for (int l = 0; l < o->nlev; l++) free_level(&o->L[l]);
    free(o->opr7); free(o->part); free(o->chol);
    for (int q = 0; q < 8; q++) free(o->bufs[q]);
    free(o);
}

/* error codes: 0 ok, -1 alloc, -2 bad arguments, -3 odd periodic, -4 coarse factor failed */
int orc_create(orc_t **out, int nx, int ny, int nz, int xper, int yper, int zper, const double *opr7, int maxlevel,
               int coarsest_max, int nu_pre, int nu_post) {
    *out = NULL;
    if (nx < 1 || ny < 1 || nz < 1 || !opr7) return -2;
    int n3[3] = {nx, ny, nz}, p3[3] = {xper != 0, yper != 0, zper != 0};
    for (int d = 0; d < 3; d++)
        if (p3[d] && n3[d] >= 3 && (n3[d] & 1)) return -3;
    orc_t *o = calloc(1, sizeof(orc_t));
    if (!o) return -1;
    o->nu_pre = nu_pre > 0 ? nu_pre : 1;
    o->nu_post = nu_post > 0 ? nu_post : 1;
    lvl_t *L = &o->L[0];
    for (int d = 0; d < 3; d++) { L->n[d] = n3[d]; L->per[d] = p3[d]; }
    L->N = (long)nx * ny * nz;
    if (alloc_level(L)) { orc_destroy(o); return -1; }
    o->nlev = 1;
    o->opr7 = xcalloc(7 * L->N, sizeof(double));
    o->part = xcalloc(L->N, 1);
    o->bufs[0] = xcalloc(nz, sizeof(double));
    if (!o->opr7 || !o->part || !o->bufs[0]) { orc_destroy(o); return -1; }
    /* identity rows (hypre_str_class.f90:174-175) */
    for (long p = 0; p < L->N; p++) {
        o->part[p] = fabs(opr7[7 * p]) > 0.0;
        if (o->part[p]) memcpy(o->opr7 + 7 * p, opr7 + 7 * p, 7 * sizeof(double));
        else { o->opr7[7 * p] = 1.0; }
    }
    /* symmetric storage */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long p = idx3(L->n, i, j, k);
                const double *A = o->opr7 + 7 * p;
                double c = A[0];
                int c3[3] = {i, j, k};
                for (int d = 0; d < 3; d++) {
                    double am = A[2 + 2 * d], ap = A[1 + 2 * d];
                    if (n3[d] == 1) {
                        if (p3[d] && o->part[p]) c += am + ap; /* self coupling through the wrap */
                        L->a[d][p] = 0.0;
                        continue;
                    }
                    int q = nbm(c3[d], n3[d], p3[d]);
                    if (q < 0 || !o->part[p]) { L->a[d][p] = 0.0; continue; }
                    int e3[3] = {i, j, k}; e3[d] = q;
                    long pq = idx3(L->n, e3[0], e3[1], e3[2]);
                    if (!o->part[pq]) { L->a[d][p] = 0.0; continue; }
                    L->a[d][p] = 0.5 * (am + o->opr7[7 * pq + 1 + 2 * d]);
                }
                L->c[p] = c;
            }
    if (coarsest_max < 1) coarsest_max = 512;
    while (1) {
        lvl_t *F = &o->L[o->nlev - 1];
        F->cdir = -1;
        if (maxlevel > 0 && o->nlev >= maxlevel) break;
        if (F->N <= coarsest_max) break;
        if (o->nlev >= ORC_MAXLEV) break;
        int d = choose_dir(F);
        if (d < 0) break;
        F->cdir = d;
        if (build_coarse(o, o->nlev - 1)) { orc_destroy(o); return -1; }
        o->nlev++;
    }
    for (int l = 0; l < ORC_MAXLEV; l++) o->relax[l] = 1;
    lvl_t *C = &o->L[o->nlev - 1];
    o->coarse_dense = C->N <= 1024;
    if (o->coarse_dense) {
        int rc = factor_coarsest(o);
        if (rc) { orc_destroy(o); return rc == -1 ? -1 : -4; }
    }
    *out = o;
    return 0;
}

/* Skipped relaxation (PFMG's SkipRelax idea, rule of this project): with
 * skip != 0 a level l with 0 < l < nlev-1 is smoothed only when its couplings
 * are isotropic: over the directions with n_d > 1 and s_d = sum_p |a_d(p)| > 0,
 * max s_d <= 2 min s_d (a level with fewer than two such directions counts as
 * isotropic); the factor 2 carries a 1e-6 relative margin, so a ratio of 2 up
 * to the rounding of the sums counts as isotropic whatever the summation order. On a skipped level the V-cycle restricts the right-hand side
 * unchanged on the way down (v = 0) and sets v = P v_coarse on the way up.
 * After semicoarsening x, y, z of an isotropic grid the strengths go
 * (1,1,1) -> (1/2,2,2) -> (1,1,4) -> (2,2,2): every third level is smoothed,
 * as in full coarsening. The finest level is always smoothed, the coarsest
 * always solved, and a locally anisotropic fine operator (skip_gate) smooths
 * every level. skip == 0 smooths every level. */
/* The skip applies only to a locally isotropic fine operator: per cell with
 * two or more directions of n_d > 1, s_d = |a_d(p)| + |a_d(p + e_d)|; a cell
 * with some s_d > 0 counts, and is anisotropic when max_d s_d > 2 (1 + 1e-6)
 * min_d s_d. The gate is open when at most 10% of the counted cells are
 * anisotropic (a stretched wall grid fails it and smooths every level). */
static int skip_gate(const lvl_t *L) {
    double cnt = 0.0, an = 0.0;
    for (int k = 0; k < L->n[2]; k++)
        for (int j = 0; j < L->n[1]; j++)
            for (int i = 0; i < L->n[0]; i++) {
                long p = idx3(L->n, i, j, k);
                int c3[3] = {i, j, k};
                double mn = INFINITY, mx = 0.0;
                int nd = 0;
                for (int d = 0; d < 3; d++) {
                    if (L->n[d] < 2) continue;
                    double sd = fabs(L->a[d][p]);
                    int q = nbp(c3[d], L->n[d], L->per[d]);
                    if (q >= 0) {
                        int e3[3] = {i, j, k};
                        e3[d] = q;
                        sd += fabs(L->a[d][idx3(L->n, e3[0], e3[1], e3[2])]);
                    }
                    nd++;
                    if (sd < mn) mn = sd;
                    if (sd > mx) mx = sd;
                }
                if (nd < 2 || !(mx > 0.0)) continue;
                cnt += 1.0;
                if (mx > 2.0 * (1.0 + 1e-6) * mn) an += 1.0;
            }
    return cnt > 0.0 && an * 10.0 <= cnt;
}

void orc_set_skip_relax(orc_t *o, int skip) {
    /* a locally anisotropic fine operator (gate closed) smooths every level */
    int keep = 1;
    if (skip && o->nlev > 2 && !skip_gate(&o->L[0])) keep = o->nlev;
    for (int l = 0; l < o->nlev; l++) {
        o->relax[l] = 1;
        if (!skip || l < keep || l == o->nlev - 1) continue;
        const lvl_t *L = &o->L[l];
        double smin = INFINITY, smax = 0.0;
        int cnt = 0;
        for (int d = 0; d < 3; d++) {
            if (L->n[d] < 2) continue;
            double sd = 0.0;
            for (long p = 0; p < L->N; p++) sd += fabs(L->a[d][p]);
            if (!(sd > 0.0)) continue;
            cnt++;
            if (sd < smin) smin = sd;
            if (sd > smax) smax = sd;
        }
        o->relax[l] = cnt < 2 || smax <= 2.0 * (1.0 + 1e-6) * smin;
    }
}

int orc_relax(const orc_t *o, int l) { return (l >= 0 && l < o->nlev) ? o->relax[l] : -1; }

/* set one level's flag directly (test hook: cycles with other skip patterns) */
int orc_set_relax_level(orc_t *o, int l, int relax) {
    if (l < 0 || l >= o->nlev) return -2;
    o->relax[l] = relax != 0;
    return 0;
}

int orc_nlevels(const orc_t *o) { return o->nlev; }
int orc_coarse_singular(const orc_t *o) { return o->coarse_singular; }
int orc_coarse_dense(const orc_t *o) { return o->coarse_dense; }

int orc_level_info(const orc_t *o, int l, int *dims, int *cdir) {
    if (l < 0 || l >= o->nlev) return -2;
    for (int d = 0; d < 3; d++) dims[d] = o->L[l].n[d];
    *cdir = o->L[l].cdir;
    return 0;
}

/* copy out c, a_x, a_y, a_z and (for l >= 1) the interpolation weights */
int orc_level_coefs(const orc_t *o, int l, double *c, double *ax, double *ay, double *az, double *pm, double *pp) {
    if (l < 0 || l >= o->nlev) return -2;
    const lvl_t *L = &o->L[l];
    size_t b = sizeof(double) * L->N;
    if (c) memcpy(c, L->c, b);
    if (ax) memcpy(ax, L->a[0], b);
    if (ay) memcpy(ay, L->a[1], b);
    if (az) memcpy(az, L->a[2], b);
    if (pm) memcpy(pm, L->pm, b);
    if (pp) memcpy(pp, L->pp, b);
    return 0;
}

int orc_apply(const orc_t *o, int l, const double *x, double *y) {
    if (l < 0 || l >= o->nlev) return -2;
    lvl_apply(&o->L[l], x, y);
    return 0;
}

void orc_apply_exact(const orc_t *o, const double *x, double *y) { exact_apply(o, x, y); }

/* order 0: red then black; 1: black then red */
int orc_rbgs(const orc_t *o, int l, const double *f, double *v, int order) {
    if (l < 0 || l >= o->nlev) return -2;
    const lvl_t *L = &o->L[l];
    gs_colour(L, f, v, order ? 1 : 0);
    gs_colour(L, f, v, order ? 0 : 1);
    return 0;
}

int orc_resid_restrict(const orc_t *o, int l, const double *f, const double *v, double *fc) {
    if (l < 0 || l + 1 >= o->nlev) return -2;
    resid_restrict(o, l, f, v, fc);
    return 0;
}

int orc_prolong_add(const orc_t *o, int l, const double *vc, double *v) {
    if (l < 0 || l + 1 >= o->nlev) return -2;
    prolong_add(o, l, vc, v);
    return 0;
}

int orc_coarse_solve(const orc_t *o, const double *f, double *v) {
    coarse_solve(o, f, v);
    return 0;
}

void orc_vcycle(orc_t *o, const double *r, double *z) { vcycle(o, r, z); }

/* Solve; rhs and sol are interior arrays (x fastest). sol holds the guess on
 * entry. Non-participating rows get rhs = sol = 0 (hypre_str_class.f90:339-345). */
int orc_solve(orc_t *o, const double *rhs, double *sol, int maxit, double rcvg, int flexible, int *it, double *rerr,
              double *trerr) {
    long N = o->L[0].N;
    double *b = xcalloc(N, sizeof(double)), *w = xcalloc(4 * N, sizeof(double));
    if (!b || !w) { free(b); free(w); return -1; }
    for (long p = 0; p < N; p++) {
        b[p] = o->part[p] ? rhs[p] : 0.0;
        if (!o->part[p]) sol[p] = 0.0;
    }
    pcg_ops op = {N, st_A, st_M, st_Ax, st_dot, o};
    *it = pcg_core(&op, b, sol, maxit, rcvg, flexible, rerr, trerr, w);
    free(b); free(w);
    return 0;
}

/* ---- dense PCG (for checking the CG recurrences on small SPD matrices) ---- */
typedef struct { int n; const double *A, *Minv; } dense_ctx;
static void dn_A(void *c, const double *x, double *y) {
    const dense_ctx *d = c;
    for (int i = 0; i < d->n; i++) { double s = 0; for (int j = 0; j < d->n; j++) s += d->A[(long)i * d->n + j] * x[j]; y[i] = s; }
}
static void dn_M(void *c, const double *x, double *y) {
    const dense_ctx *d = c;
    if (!d->Minv) { memcpy(y, x, sizeof(double) * d->n); return; }
    for (int i = 0; i < d->n; i++) { double s = 0; for (int j = 0; j < d->n; j++) s += d->Minv[(long)i * d->n + j] * x[j]; y[i] = s; }
}
static double dn_dot(void *c, const double *x, const double *y) {
    const dense_ctx *d = c; double s = 0; for (int i = 0; i < d->n; i++) s += x[i] * y[i]; return s;
}
int orc_pcg_dense(int n, const double *A, const double *Minv, const double *b, double *x, int maxit, double rcvg,
                  int flexible, int *it, double *rerr, double *trerr) {
    double *w = xcalloc(4L * n, sizeof(double));
    if (!w) return -1;
    dense_ctx c = {n, A, Minv};
    pcg_ops op = {n, dn_A, dn_M, dn_A, dn_dot, &c};
    *it = pcg_core(&op, b, x, maxit, rcvg, flexible, rerr, trerr, w);
    free(w);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following parallel regions (the CPU baseline's 1-core run) */
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
