/* Driver of the CPU oracle (oracle/pcg_pfmg.c) for the AddressSanitizer /
 * UndefinedBehaviorSanitizer build of tests/test_sanitizers.py: on grids with
 * periodic and wall directions, odd sizes, a one-cell direction and identity
 * rows it builds the hierarchy, runs every level kernel, the V-cycle, the
 * coarsest solve and PCG, and the dense PCG. Exit status 0 when every call
 * returns 0 and the solves converge. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct orc_t orc_t;
int orc_create(orc_t **out, int nx, int ny, int nz, int xper, int yper, int zper, const double *opr7, int maxlevel,
               int coarsest_max, int nu_pre, int nu_post);
void orc_destroy(orc_t *o);
int orc_nlevels(const orc_t *o);
int orc_level_info(const orc_t *o, int l, int *dims, int *cdir);
int orc_apply(const orc_t *o, int l, const double *x, double *y);
int orc_rbgs(const orc_t *o, int l, const double *f, double *v, int order);
int orc_resid_restrict(const orc_t *o, int l, const double *f, const double *v, double *fc);
int orc_prolong_add(const orc_t *o, int l, const double *vc, double *v);
int orc_coarse_solve(const orc_t *o, const double *f, double *v);
void orc_vcycle(orc_t *o, const double *r, double *z);
int orc_solve(orc_t *o, const double *rhs, double *sol, int maxit, double rcvg, int flexible, int *it, double *rerr,
              double *true_rerr);
int orc_pcg_dense(int n, const double *A, const double *Minv, const double *b, double *x, int maxit, double rcvg,
                  int flexible, int *it, double *rerr, double *true_rerr);

static double urand(unsigned long long *s) {
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

/* symmetric 7-point operator [N][7] (centre, +x, -x, +y, -y, +z, -z) with
 * random positive face weights; cells inside the block are identity rows */
static double *make_opr(int nx, int ny, int nz, const int per[3], int block, unsigned long long seed) {
    const int n[3] = {nx, ny, nz};
    const long long N = (long long)nx * ny * nz;
    double *w = malloc(sizeof(double) * 3 * N), *o = calloc(7 * N, sizeof(double));
    for (long long q = 0; q < 3 * N; q++) w[q] = 0.5 + 1.5 * urand(&seed);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long long p = i + (long long)nx * (j + (long long)ny * k);
                const int c[3] = {i, j, k};
                const int in = block && i >= nx / 3 && i < nx / 2 && j >= ny / 3 && j < ny / 2;
                if (in) continue; /* identity row: opr(1) = 0 */
                double diag = 0.0;
                for (int d = 0; d < 3; d++)
                    for (int s = 1; s >= -1; s -= 2) {
                        int q[3] = {c[0], c[1], c[2]};
                        q[d] += s;
                        if (q[d] < 0 || q[d] >= n[d]) {
                            if (!per[d] || n[d] == 1) continue;
                            q[d] = (q[d] + n[d]) % n[d];
                        }
                        if (block && q[0] >= nx / 3 && q[0] < nx / 2 && q[1] >= ny / 3 && q[1] < ny / 2) continue;
                        const long long pq = q[0] + (long long)nx * (q[1] + (long long)ny * q[2]);
                        const double a = w[3 * (s > 0 ? pq : p) + d];
                        o[7 * p + 1 + 2 * d + (s > 0 ? 0 : 1)] = -a;
                        diag += a;
                    }
                o[7 * p] = diag;
            }
    free(w);
    return o;
}

int main(void) {
    const int cases[][7] = {
        /* nx ny nz xper yper zper block */
        {12, 10, 8, 0, 1, 1, 0}, {16, 16, 1, 1, 1, 1, 0}, {9, 7, 5, 0, 0, 0, 1}, {16, 12, 8, 1, 0, 1, 1},
        {40, 6, 18, 1, 0, 0, 0},
    };
    int fails = 0;
    unsigned long long seed = 7;
    for (size_t c = 0; c < sizeof(cases) / sizeof(cases[0]); c++) {
        const int *cs = cases[c];
        const int per[3] = {cs[3], cs[4], cs[5]};
        const long long N = (long long)cs[0] * cs[1] * cs[2];
        double *opr = make_opr(cs[0], cs[1], cs[2], per, cs[6], 11 + c);
        orc_t *o = NULL;
        int rc = orc_create(&o, cs[0], cs[1], cs[2], per[0], per[1], per[2], opr, 0, 16, 1, 1);
        if (rc) {
            printf("case %zu: create rc %d\n", c, rc);
            fails++;
            free(opr);
            continue;
        }
        const int L = orc_nlevels(o);
        for (int l = 0; l < L; l++) {
            int d[3], cd;
            orc_level_info(o, l, d, &cd);
            const long long n = (long long)d[0] * d[1] * d[2];
            double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n);
            for (long long q = 0; q < n; q++) x[q] = urand(&seed) - 0.5, f[q] = urand(&seed) - 0.5;
            fails += orc_apply(o, l, x, y) != 0;
            fails += orc_rbgs(o, l, f, x, 0) != 0;
            fails += orc_rbgs(o, l, f, x, 1) != 0;
            if (l + 1 < L) {
                int dc[3], cdc;
                orc_level_info(o, l + 1, dc, &cdc);
                const long long nc = (long long)dc[0] * dc[1] * dc[2];
                double *fc = calloc(nc, sizeof(double));
                fails += orc_resid_restrict(o, l, f, x, fc) != 0;
                fails += orc_prolong_add(o, l, fc, x) != 0;
                free(fc);
            } else {
                fails += orc_coarse_solve(o, f, x) != 0;
            }
            free(x);
            free(y);
            free(f);
        }
        double *b = malloc(sizeof(double) * N), *x = calloc(N, sizeof(double)), *z = malloc(sizeof(double) * N);
        double mean = 0.0;
        long long act = 0;
        for (long long q = 0; q < N; q++) b[q] = opr[7 * q] != 0.0 ? urand(&seed) - 0.5 : 0.0;
        for (long long q = 0; q < N; q++) mean += b[q], act += opr[7 * q] != 0.0;
        for (long long q = 0; q < N; q++)
            if (opr[7 * q] != 0.0) b[q] -= mean / act;
        orc_vcycle(o, b, z);
        int it = 0;
        double rerr = 0, tr = 0;
        rc = orc_solve(o, b, x, 300, 1e-10, 0, &it, &rerr, &tr);
        printf("case %zu: %d levels, %d iterations, true residual %.2e\n", c, L, it, tr);
        if (rc || !(tr <= 1e-9)) fails++;
        orc_destroy(o);
        free(opr);
        free(b);
        free(x);
        free(z);
    }
    /* dense PCG on a small SPD matrix */
    {
        const int n = 9;
        double A[81], b[9], x[9] = {0};
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) A[i * n + j] = (i == j) ? 4.0 : (abs(i - j) == 1 ? -1.0 : 0.0);
        for (int i = 0; i < n; i++) b[i] = 1.0 + i;
        int it = 0;
        double rerr = 0, tr = 0;
        if (orc_pcg_dense(n, A, NULL, b, x, 100, 1e-12, 0, &it, &rerr, &tr) || !(tr <= 1e-11)) fails++;
    }
    return fails ? 1 : 0;
}
