// Hierarchy build of the PFMG-style multigrid (gfx950): coarsening strength,
// interpolation weights and 7-point coarse operators, FP64 -> FP32 copies.
//
// The algorithm (coarsening rule, interpolation weights, 7-point coarse
// operator, red/black order) is the one stated in oracle/pcg_pfmg.c; the
// floating-point operations are issued in the same order so the FP64 GPU
// hierarchy matches the CPU oracle to rounding.
#include "ngap_internal.h"

namespace ngap {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ void decomp(const Geo &g, long long p, int &i, int &j, int &k) {
    i = (int)(p % g.n0);
    long long r = p / g.n0;
    j = (int)(r % g.n1);
    k = (int)(r / g.n1);
}

__device__ __forceinline__ long long lin(const Geo &g, int i, int j, int k) {
    return (long long)i + (long long)g.n0 * ((long long)j + (long long)g.n1 * k);
}

__device__ __forceinline__ int gdim(const Geo &g, int d) { return d == 0 ? g.n0 : (d == 1 ? g.n1 : g.n2); }
__device__ __forceinline__ int gper(const Geo &g, int d) { return d == 0 ? g.p0 : (d == 1 ? g.p1 : g.p2); }

// ---------------------------------------------------------------- hierarchy build (FP64)

__device__ __forceinline__ double coupling_plus(const Geo &g, int d, const double *const *a, int i, int j, int k) {
    int c3[3] = {i, j, k};
    int q = nbp(c3[d], gdim(g, d), gper(g, d));
    if (q < 0) return 0.0;
    c3[d] = q;
    return a[d][lin(g, c3[0], c3[1], c3[2])];
}

struct LevPtr {
    const double *c;
    const double *a[3];
};

__device__ __forceinline__ double row_sum(const Geo &g, const LevPtr &L, int i, int j, int k) {
    long long p = lin(g, i, j, k);
    double s = L.c[p];
    for (int d = 0; d < 3; d++) s += L.a[d][p] + coupling_plus(g, d, L.a, i, j, k);
    return s;
}

// interpolation weights of the odd fine point right of each coarse cell
__global__ void __launch_bounds__(NT) k_coarse_weights(Geo F, Geo C, int d, LevPtr L, double *__restrict__ pm,
                                                       double *__restrict__ pp) {
    for (long long pc = blockIdx.x * (long long)NT + threadIdx.x; pc < C.N; pc += (long long)gridDim.x * NT) {
        int J[3];
        decomp(C, pc, J[0], J[1], J[2]);
        int o[3] = {J[0], J[1], J[2]};
        o[d] = 2 * J[d] + 1;
        double wm = 0.0, wp = 0.0;
        if (o[d] < gdim(F, d)) {
            long long po = lin(F, o[0], o[1], o[2]);
            double am = L.a[d][po], ap = coupling_plus(F, d, L.a, o[0], o[1], o[2]);
            double den = L.c[po];
            for (int t = 0; t < 3; t++)
                if (t != d) den += L.a[t][po] + coupling_plus(F, t, L.a, o[0], o[1], o[2]);
            if (den > 0.0) {
                wm = -am / den;
                wp = -ap / den;
            }
        }
        pm[pc] = wm;
        pp[pc] = wp;
    }
}

// coarse minus-couplings along every direction and coarse row sums
__global__ void __launch_bounds__(NT) k_coarse_couplings(Geo F, Geo C, int d, LevPtr L, const double *__restrict__ pm,
                                                         const double *__restrict__ pp, double *__restrict__ ca0,
                                                         double *__restrict__ ca1, double *__restrict__ ca2,
                                                         double *__restrict__ sc) {
    double *ca[3] = {ca0, ca1, ca2};
    for (long long pc = blockIdx.x * (long long)NT + threadIdx.x; pc < C.N; pc += (long long)gridDim.x * NT) {
        int J[3];
        decomp(C, pc, J[0], J[1], J[2]);
        int fe[3] = {J[0], J[1], J[2]};
        fe[d] = 2 * J[d];
        const int nd = gdim(F, d), per = gper(F, d);
        int lo = nbm(fe[d], nd, per);
        int ro = (fe[d] + 1 < nd) ? fe[d] + 1 : -1;
        int ol[3] = {fe[0], fe[1], fe[2]}, orr[3] = {fe[0], fe[1], fe[2]};
        ol[d] = lo;
        orr[d] = ro;
        const bool have_l = lo >= 0, have_r = ro >= 0;
        long long pe = lin(F, fe[0], fe[1], fe[2]);
        long long pcl = -1;
        if (have_l) {
            int Jl[3] = {J[0], J[1], J[2]};
            Jl[d] = nbm(J[d], gdim(C, d), gper(C, d));
            pcl = lin(C, Jl[0], Jl[1], Jl[2]);
        }
        // along d: 1-D Galerkin through the left odd point
        if (have_l) {
            long long po = lin(F, ol[0], ol[1], ol[2]);
            double am = L.a[d][po], ap = coupling_plus(F, d, L.a, ol[0], ol[1], ol[2]);
            double den = L.c[po];
            for (int t = 0; t < 3; t++)
                if (t != d) den += L.a[t][po] + coupling_plus(F, t, L.a, ol[0], ol[1], ol[2]);
            ca[d][pc] = den > 0.0 ? -am * ap / den : 0.0;
        } else {
            ca[d][pc] = 0.0;
        }
        // transverse: collapse with symmetrised weights
        for (int t = 0; t < 3; t++) {
            if (t == d) continue;
            int Jt[3] = {J[0], J[1], J[2]};
            Jt[t] = nbm(J[t], gdim(C, t), gper(C, t));
            if (Jt[t] < 0) {
                ca[t][pc] = 0.0;
                continue;
            }
            long long pct = lin(C, Jt[0], Jt[1], Jt[2]);
            double s = L.a[t][pe];
            if (have_r) {
                long long po = lin(F, orr[0], orr[1], orr[2]);
                s += 0.5 * (pm[pc] + pm[pct]) * L.a[t][po];
            }
            if (have_l) {
                long long po = lin(F, ol[0], ol[1], ol[2]);
                int Jlt[3] = {Jt[0], Jt[1], Jt[2]};
                Jlt[d] = nbm(J[d], gdim(C, d), gper(C, d));
                long long pclt = lin(C, Jlt[0], Jlt[1], Jlt[2]);
                s += 0.5 * (pp[pcl] + pp[pclt]) * L.a[t][po];
            }
            ca[t][pc] = s;
        }
        double s = row_sum(F, L, fe[0], fe[1], fe[2]);
        if (have_r) s += pm[pc] * row_sum(F, L, orr[0], orr[1], orr[2]);
        if (have_l) s += pp[pcl] * row_sum(F, L, ol[0], ol[1], ol[2]);
        sc[pc] = s;
    }
}

__global__ void __launch_bounds__(NT) k_coarse_diag(Geo C, const double *__restrict__ sc, double *__restrict__ cc,
                                                    const double *__restrict__ ca0, const double *__restrict__ ca1,
                                                    const double *__restrict__ ca2) {
    const double *a[3] = {ca0, ca1, ca2};
    for (long long pc = blockIdx.x * (long long)NT + threadIdx.x; pc < C.N; pc += (long long)gridDim.x * NT) {
        int i, j, k;
        decomp(C, pc, i, j, k);
        double s = sc[pc];
        for (int t = 0; t < 3; t++) s -= a[t][pc] + coupling_plus(C, t, a, i, j, k);
        cc[pc] = s;
    }
}

// sums of |a_d| per direction (coarsening strength); part has 3*MAX_PARTS slots
__global__ void __launch_bounds__(NT) k_strength(Geo g, const double *__restrict__ a0, const double *__restrict__ a1,
                                                 const double *__restrict__ a2, double *__restrict__ part) {
    __shared__ double sm[NT / 64];
    double s0 = 0, s1 = 0, s2 = 0;
    for (long long p = blockIdx.x * (long long)NT + threadIdx.x; p < g.N; p += (long long)gridDim.x * NT) {
        s0 += fabs(a0[p]);
        s1 += fabs(a1[p]);
        s2 += fabs(a2[p]);
    }
    double b0 = block_sum(s0, sm);
    double b1 = block_sum(s1, sm);
    double b2 = block_sum(s2, sm);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = b0;
        part[MAX_PARTS + blockIdx.x] = b1;
        part[2 * MAX_PARTS + blockIdx.x] = b2;
    }
}

// Local anisotropy of a level: per cell with two or more directions of
// n_d > 1, s_d = |a_d(p)| + |a_d(p + e_d)| (its two couplings along d); the
// cell counts (part) when some s_d > 0 and anisotropic (part2) when
// max_d s_d > 2 (1 + 1e-6) min_d s_d. Block partials as doubles (exact counts).
__global__ void __launch_bounds__(NT) k_aniso_count(Geo g, const double *__restrict__ a0,
                                                    const double *__restrict__ a1, const double *__restrict__ a2,
                                                    double *__restrict__ part) {
    __shared__ double sm[NT / 64];
    const double *a[3] = {a0, a1, a2};
    const int n[3] = {g.n0, g.n1, g.n2}, per[3] = {g.p0, g.p1, g.p2};
    double cnt = 0.0, an = 0.0;
    for (long long p = blockIdx.x * (long long)NT + threadIdx.x; p < g.N; p += (long long)gridDim.x * NT) {
        const int i = (int)(p % g.n0), j = (int)((p / g.n0) % g.n1), k = (int)(p / ((long long)g.n0 * g.n1));
        const int c3[3] = {i, j, k};
        const long long stride[3] = {1, g.n0, (long long)g.n0 * g.n1};
        double mn = INFINITY, mx = 0.0;
        int nd = 0;
        for (int d = 0; d < 3; d++) {
            if (n[d] < 2) continue;
            double sd = fabs(a[d][p]);
            const int q = nbp(c3[d], n[d], per[d]);
            if (q >= 0) sd += fabs(a[d][p + (long long)(q - c3[d]) * stride[d]]);
            nd++;
            mn = fmin(mn, sd);
            mx = fmax(mx, sd);
        }
        if (nd < 2 || !(mx > 0.0)) continue;
        cnt += 1.0;
        if (mx > 2.0 * (1.0 + 1e-6) * mn) an += 1.0;
    }
    const double bc = block_sum(cnt, sm);
    const double ba = block_sum(an, sm);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = bc;
        part[MAX_PARTS + blockIdx.x] = ba;
    }
}

__global__ void __launch_bounds__(NT) k_to_f32(long long n, const double *__restrict__ s, float *__restrict__ d) {
    for (long long p = blockIdx.x * (long long)NT + threadIdx.x; p < n; p += (long long)gridDim.x * NT) d[p] = (float)s[p];
}

__global__ void __launch_bounds__(NT) k_to_f64(long long n, const float *__restrict__ s, double *__restrict__ d) {
    for (long long p = blockIdx.x * (long long)NT + threadIdx.x; p < n; p += (long long)gridDim.x * NT) d[p] = (double)s[p];
}

// ---------------------------------------------------------------- host launchers

// grid of the grid-stride streaming kernels: at most 1024 blocks, so the
// single-block finalize of their partial sums stays short
int grid_for(long long n) {
    long long b = (n + NT - 1) / NT;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

void launch_coarse_build(hipStream_t s, const Geo &F, const Geo &C, int d, const double *c, const double *a0,
                         const double *a1, const double *a2, double *pm, double *pp, double *cc, double *ca0,
                         double *ca1, double *ca2, double *sc) {
    LevPtr L{c, {a0, a1, a2}};
    long long b = (C.N + NT - 1) / NT;
    if (b > 65535) b = 65535;
    hipLaunchKernelGGL(k_coarse_weights, dim3((unsigned)b), dim3(NT), 0, s, F, C, d, L, pm, pp);
    hipLaunchKernelGGL(k_coarse_couplings, dim3((unsigned)b), dim3(NT), 0, s, F, C, d, L, (const double *)pm,
                       (const double *)pp, ca0, ca1, ca2, sc);
    hipLaunchKernelGGL(k_coarse_diag, dim3((unsigned)b), dim3(NT), 0, s, C, (const double *)sc, cc,
                       (const double *)ca0, (const double *)ca1, (const double *)ca2);
}

void launch_strength(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2, double *part,
                     int *nparts) {
    int b = grid_for(g.N);
    *nparts = b;
    hipLaunchKernelGGL(k_strength, dim3(b), dim3(NT), 0, s, g, a0, a1, a2, part);
}

void launch_aniso_count(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2,
                        double *part, int *nparts) {
    int b = grid_for(g.N);
    *nparts = b;
    hipLaunchKernelGGL(k_aniso_count, dim3(b), dim3(NT), 0, s, g, a0, a1, a2, part);
}

void launch_to_f32(hipStream_t s, long long n, const double *src, float *dst) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(NT), 0, s, n, src, dst);
}

void launch_to_f64(hipStream_t s, long long n, const float *src, double *dst) {
    hipLaunchKernelGGL(k_to_f64, dim3(grid_for(n)), dim3(NT), 0, s, n, src, dst);
}

}  // namespace ngap
