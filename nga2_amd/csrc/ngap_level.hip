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

// Cell indices of a level fit in int (setup rejects levels of 2^31 cells or
// more), so the build kernels index in 32 bits: 64-bit divisions and
// multiplies cost the coarse build several times its memory time.

__device__ __forceinline__ int lin(const Geo &g, int i, int j, int k) { return i + g.n0 * (j + g.n1 * k); }

__device__ __forceinline__ int gdim(const Geo &g, int d) { return d == 0 ? g.n0 : (d == 1 ? g.n1 : g.n2); }
__device__ __forceinline__ int gper(const Geo &g, int d) { return d == 0 ? g.p0 : (d == 1 ? g.p1 : g.p2); }

// One thread per cell: t = i + n0 j indexes the (i, j) plane, k runs over
// blockIdx.y (strided past 65535 planes). Returns false for threads past
// the plane.
__device__ __forceinline__ bool plane_cell(const Geo &g, int &t, int &i, int &j) {
    t = (int)(blockIdx.x * NT + threadIdx.x);
    const unsigned plane = (unsigned)g.n0 * (unsigned)g.n1;
    if ((unsigned)t >= plane) return false;
    j = (int)((unsigned)t / (unsigned)g.n0);
    i = t - j * g.n0;
    return true;
}

dim3 plane_grid(const Geo &g) {
    const long long plane = (long long)g.n0 * g.n1;
    return dim3((unsigned)((plane + NT - 1) / NT), (unsigned)(g.n2 < 65535 ? g.n2 : 65535));
}

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
    int p = lin(g, i, j, k);
    double s = L.c[p];
#pragma unroll
    for (int d = 0; d < 3; d++) s += L.a[d][p] + coupling_plus(g, d, L.a, i, j, k);
    return s;
}

// interpolation weights of the odd fine point right of each coarse cell;
// FP32 copies too when pm32 is given (the mixed-precision cycle's). The
// coarsening direction d is a template argument so that the per-direction
// index arrays stay in registers.
template <int d>
__global__ void __launch_bounds__(NT) k_coarse_weights(Geo F, Geo C, LevPtr L, double *__restrict__ pm,
                                                       double *__restrict__ pp, float *__restrict__ pm32,
                                                       float *__restrict__ pp32) {
    int t, i0, j0;
    if (!plane_cell(C, t, i0, j0)) return;
    const int plane = C.n0 * C.n1;
    for (int k0 = blockIdx.y; k0 < C.n2; k0 += gridDim.y) {
        const int pc = t + plane * k0;
        int o[3] = {i0, j0, k0};
        o[d] = 2 * o[d] + 1;
        double wm = 0.0, wp = 0.0;
        if (o[d] < gdim(F, d)) {
            int po = lin(F, o[0], o[1], o[2]);
            double am = L.a[d][po], ap = coupling_plus(F, d, L.a, o[0], o[1], o[2]);
            double den = L.c[po];
#pragma unroll
            for (int t2 = 0; t2 < 3; t2++)
                if (t2 != d) den += L.a[t2][po] + coupling_plus(F, t2, L.a, o[0], o[1], o[2]);
            if (den > 0.0) {
                wm = -am / den;
                wp = -ap / den;
            }
        }
        pm[pc] = wm;
        pp[pc] = wp;
        if (pm32) {
            pm32[pc] = (float)wm;
            pp32[pc] = (float)wp;
        }
    }
}

// coarse minus-couplings along every direction and coarse row sums
template <int d>
__global__ void __launch_bounds__(NT) k_coarse_couplings(Geo F, Geo C, LevPtr L, const double *__restrict__ pm,
                                                         const double *__restrict__ pp, double *__restrict__ ca0,
                                                         double *__restrict__ ca1, double *__restrict__ ca2,
                                                         double *__restrict__ sc, float *__restrict__ cs0,
                                                         float *__restrict__ cs1, float *__restrict__ cs2) {
    double *ca[3] = {ca0, ca1, ca2};
    float *cs[3] = {cs0, cs1, cs2};
    int t0, i0, j0;
    if (!plane_cell(C, t0, i0, j0)) return;
    const int plane = C.n0 * C.n1;
    for (int k0 = blockIdx.y; k0 < C.n2; k0 += gridDim.y) {
        const int pc = t0 + plane * k0;
        const int J[3] = {i0, j0, k0};
        int fe[3] = {J[0], J[1], J[2]};
        fe[d] = 2 * J[d];
        const int nd = gdim(F, d), per = gper(F, d);
        int lo = nbm(fe[d], nd, per);
        int ro = (fe[d] + 1 < nd) ? fe[d] + 1 : -1;
        int ol[3] = {fe[0], fe[1], fe[2]}, orr[3] = {fe[0], fe[1], fe[2]};
        ol[d] = lo;
        orr[d] = ro;
        const bool have_l = lo >= 0, have_r = ro >= 0;
        int pe = lin(F, fe[0], fe[1], fe[2]);
        int pcl = -1;
        if (have_l) {
            int Jl[3] = {J[0], J[1], J[2]};
            Jl[d] = nbm(J[d], gdim(C, d), gper(C, d));
            pcl = lin(C, Jl[0], Jl[1], Jl[2]);
        }
        double out[3];
        // along d: 1-D Galerkin through the left odd point
        if (have_l) {
            int po = lin(F, ol[0], ol[1], ol[2]);
            double am = L.a[d][po], ap = coupling_plus(F, d, L.a, ol[0], ol[1], ol[2]);
            double den = L.c[po];
#pragma unroll
            for (int t = 0; t < 3; t++)
                if (t != d) den += L.a[t][po] + coupling_plus(F, t, L.a, ol[0], ol[1], ol[2]);
            out[d] = den > 0.0 ? -am * ap / den : 0.0;
        } else {
            out[d] = 0.0;
        }
        // transverse: collapse with symmetrised weights
#pragma unroll
        for (int t = 0; t < 3; t++) {
            if (t == d) continue;
            int Jt[3] = {J[0], J[1], J[2]};
            Jt[t] = nbm(J[t], gdim(C, t), gper(C, t));
            if (Jt[t] < 0) {
                out[t] = 0.0;
                continue;
            }
            int pct = lin(C, Jt[0], Jt[1], Jt[2]);
            double s = L.a[t][pe];
            if (have_r) {
                int po = lin(F, orr[0], orr[1], orr[2]);
                s += 0.5 * (pm[pc] + pm[pct]) * L.a[t][po];
            }
            if (have_l) {
                int po = lin(F, ol[0], ol[1], ol[2]);
                int Jlt[3] = {Jt[0], Jt[1], Jt[2]};
                Jlt[d] = nbm(J[d], gdim(C, d), gper(C, d));
                int pclt = lin(C, Jlt[0], Jlt[1], Jlt[2]);
                s += 0.5 * (pp[pcl] + pp[pclt]) * L.a[t][po];
            }
            out[t] = s;
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
            ca[t][pc] = out[t];
            if (cs0) cs[t][pc] = (float)out[t];
        }
        double s = row_sum(F, L, fe[0], fe[1], fe[2]);
        if (have_r) s += pm[pc] * row_sum(F, L, orr[0], orr[1], orr[2]);
        if (have_l) s += pp[pcl] * row_sum(F, L, ol[0], ol[1], ol[2]);
        sc[pc] = s;
    }
}

__global__ void __launch_bounds__(NT) k_coarse_diag(Geo C, const double *__restrict__ sc, double *__restrict__ cc,
                                                    const double *__restrict__ ca0, const double *__restrict__ ca1,
                                                    const double *__restrict__ ca2, float *__restrict__ cc32) {
    const double *a[3] = {ca0, ca1, ca2};
    int t, i, j;
    if (!plane_cell(C, t, i, j)) return;
    const int plane = C.n0 * C.n1;
    for (int k = blockIdx.y; k < C.n2; k += gridDim.y) {
        const int pc = t + plane * k;
        double s = sc[pc];
#pragma unroll
        for (int d = 0; d < 3; d++) s -= a[d][pc] + coupling_plus(C, d, a, i, j, k);
        cc[pc] = s;
        if (cc32) cc32[pc] = (float)s;
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
    const int stride[3] = {1, g.n0, g.n0 * g.n1};
    // 64-bit loop counter (its last step past N must not overflow); p fits in int
    for (long long pl = (long long)blockIdx.x * NT + threadIdx.x; pl < g.N; pl += (long long)gridDim.x * NT) {
        const int p = (int)pl;
        const unsigned r = (unsigned)p / (unsigned)g.n0;
        const int i = p - (int)r * g.n0, j = (int)(r % (unsigned)g.n1), k = (int)(r / (unsigned)g.n1);
        const int c3[3] = {i, j, k};
        double mn = INFINITY, mx = 0.0;
        int nd = 0;
        for (int d = 0; d < 3; d++) {
            if (n[d] < 2) continue;
            double sd = fabs(a[d][p]);
            const int q = nbp(c3[d], n[d], per[d]);
            if (q >= 0) sd += fabs(a[d][p + (q - c3[d]) * stride[d]]);
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

// Dense (pseudo-)inverse of a coarsest level of n <= COARSE_LDS_MAX cells in
// one workgroup, the matrix held in LDS (129 KB at n = 128):
// - assembly of the symmetric 7-point matrix, row p by thread p, and its row
//   sums; the level counts as singular when sum |row sums| <= 1e-10 sum |c|,
//   and then every entry gets csum / n^2 (the rank-one shift that makes the
//   inverse a pseudo-inverse on mean-free vectors);
// - Cholesky A = L L^T (right-looking, in place);
// - W = L^-1 in place, column by column from the last, 8 threads per row;
// - A^-1 = W^T W into Ainv (n x n, symmetric).
// Rows are LDM = 129 doubles apart, so the column accesses of the
// factorisation fall on different LDS banks.
// out[0] = 1 when singular, out[1] = 1 when a pivot is not positive.
constexpr int NTC = 1024;
constexpr int LDM = COARSE_LDS_MAX + 1;
__global__ void __launch_bounds__(NTC) k_coarse_inverse(Geo g, const double *__restrict__ c,
                                                        const double *__restrict__ a0, const double *__restrict__ a1,
                                                        const double *__restrict__ a2, double *__restrict__ Ainv,
                                                        double *__restrict__ out) {
    __shared__ double M[COARSE_LDS_MAX * LDM];
    __shared__ double col[COARSE_LDS_MAX];
    __shared__ double red[2 * (NTC / 64)];
    __shared__ double gam;
    __shared__ int sing, bad;
    const int n = (int)g.N, t = threadIdx.x;
    for (int e = t; e < n * LDM; e += NTC) M[e] = 0.0;
    __syncthreads();
    double rsa = 0.0, ca = 0.0;
    if (t < n) {
        const double *a[3] = {a0, a1, a2};
        const int r = t / g.n0;
        const int c3[3] = {t - r * g.n0, r % g.n1, r / g.n1};
        const int n3[3] = {g.n0, g.n1, g.n2}, p3[3] = {g.p0, g.p1, g.p2};
        const int st3[3] = {1, g.n0, g.n0 * g.n1};
        double *row = M + t * LDM;
        row[t] += c[t];
        double rs = c[t];
        for (int d = 0; d < 3; d++) {
            const int qm = nbm(c3[d], n3[d], p3[d]);
            if (qm >= 0) row[t + (qm - c3[d]) * st3[d]] += a[d][t];
            rs += a[d][t];
            const int qp = nbp(c3[d], n3[d], p3[d]);
            if (qp >= 0) {
                const int pq = t + (qp - c3[d]) * st3[d];
                row[pq] += a[d][pq];
                rs += a[d][pq];
            }
        }
        rsa = fabs(rs);
        ca = fabs(c[t]);
    }
    rsa = wave_sum(rsa);
    ca = wave_sum(ca);
    if ((t & 63) == 0) {
        red[t >> 6] = rsa;
        red[NTC / 64 + (t >> 6)] = ca;
    }
    __syncthreads();
    if (t == 0) {
        double ss = 0.0, cs = 0.0;
        for (int w = 0; w < NTC / 64; w++) {
            ss += red[w];
            cs += red[NTC / 64 + w];
        }
        sing = ss <= 1e-10 * cs;
        gam = cs / ((double)n * n);
        bad = 0;
    }
    __syncthreads();
    if (sing)
        for (int e = t; e < n * n; e += NTC) {
            const int i = e / n;
            M[i * LDM + (e - i * n)] += gam;
        }
    __syncthreads();
    // Cholesky: column j scaled by the pivot, then the trailing lower triangle
    // updated by 32 x 32 thread tiles
    for (int j = 0; j < n; j++) {
        if (t == 0) {
            const double d = M[j * LDM + j];
            if (!(d > 0.0)) bad = 1;
            M[j * LDM + j] = sqrt(d);
        }
        __syncthreads();
        const double ljj = M[j * LDM + j];
        for (int i = j + 1 + t; i < n; i += NTC) M[i * LDM + j] /= ljj;
        __syncthreads();
        for (int i = j + 1 + (t >> 5); i < n; i += NTC / 32) {
            const double lij = M[i * LDM + j];
            for (int k = j + 1 + (t & 31); k <= i; k += 32) M[i * LDM + k] -= lij * M[k * LDM + j];
        }
        __syncthreads();
    }
    // W = L^-1: W[j][j] = 1 / L[j][j], W[i][j] = -W[j][j] sum_{k=j+1..i} W[i][k] L[k][j];
    // each row's sum by 8 consecutive lanes
    for (int j = n - 1; j >= 0; j--) {
        const double wjj = 1.0 / M[j * LDM + j];
        for (int i = j + 1 + t; i < n; i += NTC) col[i] = M[i * LDM + j];
        __syncthreads();
        for (int i = j + 1 + (t >> 3); i < n; i += NTC / 8) {
            double sacc = 0.0;
            for (int k = j + 1 + (t & 7); k <= i; k += 8) sacc += M[i * LDM + k] * col[k];
            sacc += __shfl_xor(sacc, 1, 64);
            sacc += __shfl_xor(sacc, 2, 64);
            sacc += __shfl_xor(sacc, 4, 64);
            if ((t & 7) == 0) M[i * LDM + j] = -wjj * sacc;
        }
        if (t == 0) M[j * LDM + j] = wjj;
        __syncthreads();
    }
    // A^-1[a][b] = sum_{i >= max(a, b)} W[i][a] W[i][b]
    for (int e = t; e < n * n; e += NTC) {
        const int ra = e / n, cb = e - ra * n;
        double sacc = 0.0;
        for (int i = ra > cb ? ra : cb; i < n; i++) sacc += M[i * LDM + ra] * M[i * LDM + cb];
        Ainv[e] = sacc;
    }
    if (t == 0) {
        out[0] = sing ? 1.0 : 0.0;
        out[1] = bad ? 1.0 : 0.0;
    }
}

void launch_coarse_inverse(hipStream_t s, const Geo &g, const double *c, const double *a0, const double *a1,
                           const double *a2, double *Ainv, double *out) {
    hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(NTC), 0, s, g, c, a0, a1, a2, Ainv, out);
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
                         double *ca1, double *ca2, double *sc, const LevelData<float> *c32) {
    LevPtr L{c, {a0, a1, a2}};
    const dim3 gr = plane_grid(C);
    float *n32 = nullptr;
    float *w0 = c32 ? c32->pm : n32, *w1 = c32 ? c32->pp : n32;
    float *s0 = c32 ? c32->a0 : n32, *s1 = c32 ? c32->a1 : n32, *s2 = c32 ? c32->a2 : n32;
    const double *cpm = pm, *cpp = pp;
    if (d == 0) {
        hipLaunchKernelGGL(k_coarse_weights<0>, gr, dim3(NT), 0, s, F, C, L, pm, pp, w0, w1);
        hipLaunchKernelGGL(k_coarse_couplings<0>, gr, dim3(NT), 0, s, F, C, L, cpm, cpp, ca0, ca1, ca2, sc, s0, s1, s2);
    } else if (d == 1) {
        hipLaunchKernelGGL(k_coarse_weights<1>, gr, dim3(NT), 0, s, F, C, L, pm, pp, w0, w1);
        hipLaunchKernelGGL(k_coarse_couplings<1>, gr, dim3(NT), 0, s, F, C, L, cpm, cpp, ca0, ca1, ca2, sc, s0, s1, s2);
    } else {
        hipLaunchKernelGGL(k_coarse_weights<2>, gr, dim3(NT), 0, s, F, C, L, pm, pp, w0, w1);
        hipLaunchKernelGGL(k_coarse_couplings<2>, gr, dim3(NT), 0, s, F, C, L, cpm, cpp, ca0, ca1, ca2, sc, s0, s1, s2);
    }
    hipLaunchKernelGGL(k_coarse_diag, gr, dim3(NT), 0, s, C, (const double *)sc, cc, (const double *)ca0,
                       (const double *)ca1, (const double *)ca2, c32 ? c32->c : n32);
}

// grid of the setup reductions: one partial per block, up to MAX_PARTS blocks
// (few long grid-stride loops leave the loads' latency exposed)
int grid_parts(long long n) {
    long long b = (n + NT - 1) / NT;
    if (b < 1) b = 1;
    if (b > MAX_PARTS) b = MAX_PARTS;
    return (int)b;
}

void launch_strength(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2, double *part,
                     int *nparts) {
    // three streams without neighbours: 1024 blocks reach the bandwidth, and
    // the single-block finalize of three partial arrays stays short
    int b = grid_for(g.N);
    *nparts = b;
    hipLaunchKernelGGL(k_strength, dim3(b), dim3(NT), 0, s, g, a0, a1, a2, part);
}

void launch_aniso_count(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2,
                        double *part, int *nparts) {
    int b = grid_parts(g.N);
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
