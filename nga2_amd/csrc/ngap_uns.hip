// GPU side of the unstructured solver (libngapuns.so): NGA2's hypre_uns
// (src/solver/hypre_uns_class.f90) with the multigrid hierarchy built on the
// host (ngap_amg.cpp) and every cycle, sparse product and Krylov step on the
// device.
//
// Data in HBM: per level the CSR matrix (int32 row pointers and columns, FP64
// values), the interpolation P and its transpose R in CSR, 1/d for the
// l1-Jacobi smoother, and three vectors (right-hand side, iterate, second
// iterate / residual); the coarsest level's dense inverse. One kernel covers
// the four sparse products: y = A x, y = f - A x, y = x + (f - A x)/d and
// y += A x, with TPR lanes per row (TPR from the level's mean row length) and
// the row sum reduced across those lanes by shuffles.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "ngap_uns_internal.h"

namespace {

#define UCHK(x)                                                                                          \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                          \
            fprintf(stderr, "[ngap_uns] %s failed: %s\n", #x, hipGetErrorString(e_));                    \
            return NGAP_ERR_HIP;                                                                         \
        }                                                                                                \
    } while (0)

#define RCHK(x)                   \
    do {                          \
        const int r_ = (x);       \
        if (r_) return r_;        \
    } while (0)

constexpr int NPART = 1024;     // partials of a dot product
constexpr int MAX_VBLOCKS = 8192; // grid cap of the vector kernels (partials buffer size)

// y = A x; y = f - A x; y = x + (f - A x) / d; y += A x; and, for the first
// sweep from a zero guess (x = f / d, never stored): y = f - A (f / d) and
// y = f / d + A x
enum { CSR_MUL = 0, CSR_RESID = 1, CSR_JACOBI = 2, CSR_ADD = 3, CSR_RESID0 = 4, CSR_ADD0 = 5 };

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

// T: the cycle's value type (FP64, or FP32 for the mixed-precision cycle)
template <typename T, int TPR, int MODE>
__global__ void __launch_bounds__(256) k_csr(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                             const T *__restrict__ val, const T *__restrict__ x,
                                             const T *__restrict__ f, const T *__restrict__ dinv,
                                             T *__restrict__ y) {
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    const int row = (int)(t / TPR), g = (int)(t % TPR);
    if (row >= n) return;  // whole TPR groups leave together
    T s = T(0);
    const int q1 = rp[row + 1];
    if (MODE == CSR_RESID0) {
        for (int q = rp[row] + g; q < q1; q += TPR) {
            const int c = col[q];
            s += val[q] * mul_rn(f[c], dinv[c]);
        }
    } else {
        for (int q = rp[row] + g; q < q1; q += TPR) s += val[q] * x[col[q]];
    }
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, TPR);
    if (g == 0) {
        if (MODE == CSR_MUL) y[row] = s;
        if (MODE == CSR_RESID) y[row] = f[row] - s;
        if (MODE == CSR_JACOBI) y[row] = x[row] + (f[row] - s) * dinv[row];
        if (MODE == CSR_ADD) y[row] += s;
        if (MODE == CSR_RESID0) y[row] = f[row] - s;
        if (MODE == CSR_ADD0) y[row] = mul_rn(f[row], dinv[row]) + s;
    }
}

template <typename T>
__global__ void k_scale(int n, const T *__restrict__ f, const T *__restrict__ dinv, T *__restrict__ x) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) x[i] = f[i] * dinv[i];
}

// coarsest level: x = Ainv f, one wavefront per row, FP64 sums
template <typename T>
__global__ void __launch_bounds__(256) k_dense(int n, const double *__restrict__ Ainv, const T *__restrict__ f,
                                               T *__restrict__ x) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += Ainv[(long long)row * n + j] * (double)f[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) x[row] = (T)s;
}

// precision conversions around the FP32 cycle
__global__ void k_to_f32(int n, const double *__restrict__ a, float *__restrict__ b) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = (float)a[i];
}
__global__ void k_to_f64(int n, const float *__restrict__ a, double *__restrict__ b) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = (double)a[i];
}

// one FP64 partial of <a, b> per workgroup (grid-stride, NPART workgroups)
__global__ void __launch_bounds__(256) k_dot(int n, const double *__restrict__ a, const double *__restrict__ b,
                                             double *__restrict__ part) {
    __shared__ double sm[4];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += a[i] * b[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}

// out[slot] = sum of the partials in fixed order
__global__ void __launch_bounds__(256) k_fin(const double *__restrict__ part, int np, double *__restrict__ out,
                                             int slot) {
    __shared__ double sm[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[slot] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}

// PCG scalars on the device: scal[0] rho = <r, z>, scal[1] <p, A p>,
// scal[2] <r, r>, scal[3] alpha, scal[4] beta
__device__ double block_sum(double s) {
    __shared__ double sm[4];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    return (sm[0] + sm[1]) + (sm[2] + sm[3]);
}

// scal[1] = sum of the partials, alpha = rho / scal[1]
__global__ void __launch_bounds__(256) k_fin_alpha(const double *__restrict__ part, int np, double *scal) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        scal[1] = s;
        scal[3] = scal[0] / s;
    }
}

// rho1 = sum of the partials, beta = rho1 / rho, rho = rho1
__global__ void __launch_bounds__(256) k_fin_beta(const double *__restrict__ part, int np, double *scal) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        scal[4] = s / scal[0];
        scal[0] = s;
    }
}

// x += alpha p, r -= alpha q, one partial of <r, r> per workgroup
__global__ void __launch_bounds__(256) k_pcg_xr(int n, const double *__restrict__ scal, const double *__restrict__ p,
                                                const double *__restrict__ q, double *__restrict__ x,
                                                double *__restrict__ r, double *__restrict__ part) {
    const double a = scal[3];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] += a * p[i];
        const double ri = r[i] - a * q[i];
        r[i] = ri;
        s += ri * ri;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// partials of <r, z> and <z, q> (flexible PCG), two per workgroup
__global__ void __launch_bounds__(256) k_dot2(int n, const double *__restrict__ r, const double *__restrict__ z,
                                              const double *__restrict__ q, double *__restrict__ part) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        s1 += r[i] * z[i];
        s2 += z[i] * q[i];
    }
    s1 = block_sum(s1);
    __syncthreads();
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

// flexible beta = -alpha <z1, q> / rho (r1 - r = -alpha q), rho = <r1, z1>
__global__ void __launch_bounds__(256) k_fin_beta_flex(const double *__restrict__ part, int np, double *scal) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) {
        s1 += part[2 * i];
        s2 += part[2 * i + 1];
    }
    s1 = block_sum(s1);
    __syncthreads();
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        scal[4] = -scal[3] * s2 / scal[0];
        scal[0] = s1;
    }
}

// p = z + beta p
__global__ void k_pcg_p(int n, const double *__restrict__ scal, const double *__restrict__ z, double *__restrict__ p) {
    const double b = scal[4];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = z[i] + b * p[i];
}

// y = a x + b y (x may be y)
__global__ void k_axpby(int n, double a, const double *x, double b, double *y) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) y[i] = a * x[i] + b * y[i];
}

// z = x + a y + b w (y, w may be null; z may be x or y: each element is read before it is written)
__global__ void k_lin3(int n, const double *x, double a, const double *y, double b, const double *w, double *z) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double v = x[i];
        if (y) v += a * y[i];
        if (w) v += b * w[i];
        z[i] = v;
    }
}

__global__ void k_mul(int n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ y) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) y[i] = a[i] * b[i];
}

__global__ void k_gather(int n, const int *__restrict__ map, const double *__restrict__ src, double *__restrict__ dst) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[i] = src[map[i]];
}

__global__ void k_scatter(int n, const int *__restrict__ map, const double *__restrict__ src, double *__restrict__ dst) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[map[i]] = src[i];
}

int vec_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > MAX_VBLOCKS) b = MAX_VBLOCKS;
    return (int)(b < 1 ? 1 : b);
}

int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// ------------------------------------------------------------------ device structures
struct DCsr {
    int n = 0, ncol = 0, tpr = 1;
    long long nnz = 0;
    int *rp = nullptr, *col = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;  // FP32 copy for the mixed-precision cycle
    const double *v(double) const { return val; }
    const float *v(float) const { return val32; }
};

template <typename T>
struct LVec {
    T *dinv = nullptr, *f = nullptr, *x = nullptr, *w = nullptr, *r = nullptr;
};

struct DLevel {
    DCsr A, P, R;  // P: this level's rows x next level's rows; R = P^T
    LVec<double> d;
    LVec<float> s;
    LVec<double> &vec(double) { return d; }
    LVec<float> &vec(float) { return s; }
};

template <typename T, int MODE>
void launch_csr_mode(hipStream_t s, const DCsr &M, const T *x, const T *f, const T *dinv, T *y) {
    const long long threads = (long long)M.n * M.tpr;
    const dim3 grid((unsigned)((threads + 255) / 256));
    const T *val = M.v(T());
    switch (M.tpr) {
        case 1: hipLaunchKernelGGL((k_csr<T, 1, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        case 2: hipLaunchKernelGGL((k_csr<T, 2, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        case 4: hipLaunchKernelGGL((k_csr<T, 4, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        case 8: hipLaunchKernelGGL((k_csr<T, 8, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        case 16: hipLaunchKernelGGL((k_csr<T, 16, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        case 32: hipLaunchKernelGGL((k_csr<T, 32, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
        default: hipLaunchKernelGGL((k_csr<T, 64, MODE>), grid, dim3(256), 0, s, M.n, M.rp, M.col, val, x, f, dinv, y); break;
    }
}

template <typename T>
void csr_mul(hipStream_t s, const DCsr &M, const T *x, T *y) {
    launch_csr_mode<T, CSR_MUL>(s, M, x, nullptr, nullptr, y);
}
template <typename T>
void csr_resid(hipStream_t s, const DCsr &M, const T *f, const T *x, T *y) {
    launch_csr_mode<T, CSR_RESID>(s, M, x, f, nullptr, y);
}
template <typename T>
void csr_jacobi(hipStream_t s, const DCsr &M, const T *f, const T *dinv, const T *x, T *y) {
    launch_csr_mode<T, CSR_JACOBI>(s, M, x, f, dinv, y);
}
template <typename T>
void csr_add(hipStream_t s, const DCsr &M, const T *x, T *y) {
    launch_csr_mode<T, CSR_ADD>(s, M, x, nullptr, nullptr, y);
}
template <typename T>
void csr_resid0(hipStream_t s, const DCsr &M, const T *f, const T *dinv, T *y) {
    launch_csr_mode<T, CSR_RESID0>(s, M, nullptr, f, dinv, y);
}
template <typename T>
void csr_add0(hipStream_t s, const DCsr &P, const T *f, const T *dinv, const T *x, T *y) {
    launch_csr_mode<T, CSR_ADD0>(s, P, x, f, dinv, y);
}

}  // namespace

// ------------------------------------------------------------------ handle
struct ngap_uns {
    ngap_grid g{};
    int nst = 0, method = 0;
    std::vector<int> stc;
    int inited = 0, setup_done = 0;
    std::vector<int> ind;     // ghosted numbering
    std::vector<int> map;     // row -> ghosted linear index
    long long ntot = 0;       // ghosted cells
    int rows = 0;
    int maxit = 0;
    double rcvg = 0.0;
    ngap_uns_opts o{};
    ngap_amg *hier = nullptr;
    std::vector<DLevel> L;
    double *Ainv = nullptr;
    int nc = 0;
    int *d_map = nullptr;
    double *dgi = nullptr;  // 1/a_ii of the fine matrix (Jacobi-scaled methods)
    // Krylov vectors (rows entries each)
    double *b = nullptr, *x = nullptr, *r = nullptr, *vin = nullptr, *vout = nullptr, *p = nullptr, *q = nullptr,
           *s = nullptr, *t = nullptr, *v = nullptr, *rh = nullptr, *ph = nullptr, *sh = nullptr;
    float *vin32 = nullptr, *vout32 = nullptr;  // the FP32 cycle's input and output
    double *V = nullptr, *Z = nullptr;  // GMRES bases
    double *grhs = nullptr, *gsol = nullptr;  // ghosted staging for the host entry point
    double *part = nullptr, *scal = nullptr, *h_scal = nullptr;
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    hipGraphExec_t vgraph = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool time_next = false, timed = false;
    ngap_uns_stats st{};
};

namespace {

int dalloc(ngap_uns *h, void **p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return NGAP_ERR_ALLOC;
    }
    h->allocs.push_back(*p);
    return 0;
}

template <typename T>
int upload(ngap_uns *h, T **dst, const std::vector<T> &src) {
    RCHK(dalloc(h, (void **)dst, sizeof(T) * src.size()));
    if (!src.empty()) UCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return 0;
}

int upload_csr(ngap_uns *h, DCsr &D, const ngap::amg::Csr &M, bool f32 = false) {
    D.n = M.n;
    D.ncol = M.ncol;
    D.nnz = (long long)M.col.size();
    // lanes per row: a quarter of the mean row length rounded up to a power of
    // two (measured at 128^3 against a half, the whole and an eighth: 0.50 ms
    // per V-cycle against 0.52, 0.63 and 0.70)
    const int mean = M.n > 0 ? (int)((D.nnz + M.n - 1) / M.n) : 1;
    D.tpr = std::min(64, std::max(1, pow2ceil(mean) / 4));
    RCHK(upload(h, &D.rp, M.rp));
    RCHK(upload(h, &D.col, M.col));
    RCHK(upload(h, &D.val, M.val));
    if (f32) RCHK(upload(h, &D.val32, std::vector<float>(M.val.begin(), M.val.end())));
    return 0;
}

// the vectors and 1/d of one level in precision T
template <typename T>
int level_vectors(ngap_uns *h, LVec<T> &V, const std::vector<double> &d, int l) {
    const int n = (int)d.size();
    std::vector<T> di(n);
    for (int i = 0; i < n; i++) di[i] = d[i] > 0.0 ? (T)(1.0 / d[i]) : T(0);
    RCHK(upload(h, &V.dinv, di));
    if (l > 0) {
        RCHK(dalloc(h, (void **)&V.f, sizeof(T) * n));
        RCHK(dalloc(h, (void **)&V.x, sizeof(T) * n));
    }
    RCHK(dalloc(h, (void **)&V.w, sizeof(T) * n));
    RCHK(dalloc(h, (void **)&V.r, sizeof(T) * n));
    return 0;
}

void free_device(ngap_uns *h) {
    if (h->vgraph) (void)hipGraphExecDestroy(h->vgraph);
    h->vgraph = nullptr;
    for (void *p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->h_scal) (void)hipHostFree(h->h_scal);
    h->h_scal = nullptr;
    h->L.clear();
    h->Ainv = nullptr;
    h->vin32 = h->vout32 = nullptr;
    h->setup_done = 0;
}

bool is_amg(int m) {
    return m == NGAP_UNS_AMG || m == NGAP_UNS_PCG_AMG || m == NGAP_UNS_GMRES_AMG || m == NGAP_UNS_BICGSTAB_AMG;
}

// ------------------------------------------------------------------ cycle
// One V(nu, nu) cycle from a zero guess, vout = M vin. On each level above the
// coarsest the first sweep writes w and every later launch switches buffers,
// so after 2 nu - 1 switches the result lies in x (vout on level 0). With one
// sweep on a sparse level the first one (x = f / d) is never stored: the
// residual reads f / d, and the prolongation writes f / d + P e into w.
template <typename T>
void vcycle(ngap_uns *h, const T *vin, T *vout) {
    hipStream_t s = h->stream;
    const int nl = (int)h->L.size(), nu = h->o.nu;
    if (nl == 1) {
        const int n = h->L[0].A.n;
        hipLaunchKernelGGL(k_dense<T>, dim3((n + 3) / 4), dim3(256), 0, s, n, h->Ainv, vin, vout);
        return;
    }
    std::vector<T *> cur(nl);
    for (int l = 0; l < nl - 1; l++) {
        DLevel &D = h->L[l];
        LVec<T> &V = D.vec(T());
        const T *f = l == 0 ? vin : V.f;
        T *xf = l == 0 ? vout : V.x;
        T *bufs[2] = {V.w, xf};
        // fused only on sparse levels: with long rows the two gathers (f and
        // 1/d) cost more than the stored x (measured at 128^3: 110 us fused
        // against 86 us unfused at 30 nonzeros per row, 55 against 59 at 7)
        if (nu == 1 && D.A.nnz <= 12LL * D.A.n) {
            // the sweep from zero, x = f / d, stays implicit: the residual and
            // the prolongation read f / d
            cur[l] = nullptr;
            csr_resid0<T>(s, D.A, f, V.dinv, V.r);
        } else {
            int nb = 0;
            hipLaunchKernelGGL(k_scale<T>, dim3(vec_blocks(D.A.n)), dim3(256), 0, s, D.A.n, f, (const T *)V.dinv,
                               bufs[nb]);
            for (int sw = 1; sw < nu; sw++) {
                csr_jacobi<T>(s, D.A, f, V.dinv, bufs[nb], bufs[nb ^ 1]);
                nb ^= 1;
            }
            cur[l] = bufs[nb];
            csr_resid<T>(s, D.A, f, cur[l], V.r);
        }
        csr_mul<T>(s, D.R, V.r, h->L[l + 1].vec(T()).f);
    }
    {
        DLevel &C = h->L[nl - 1];
        LVec<T> &V = C.vec(T());
        hipLaunchKernelGGL(k_dense<T>, dim3((C.A.n + 3) / 4), dim3(256), 0, s, C.A.n, h->Ainv, (const T *)V.f, V.x);
    }
    for (int l = nl - 2; l >= 0; l--) {
        DLevel &D = h->L[l];
        LVec<T> &V = D.vec(T());
        const T *f = l == 0 ? vin : V.f;
        T *xf = l == 0 ? vout : V.x;
        T *c = cur[l];
        const T *ec = h->L[l + 1].vec(T()).x;
        if (!c) {
            c = V.w;
            csr_add0<T>(s, D.P, f, V.dinv, ec, c);
        } else {
            csr_add<T>(s, D.P, ec, c);
        }
        for (int sw = 0; sw < nu; sw++) {
            T *nxt = c == xf ? V.w : xf;
            csr_jacobi<T>(s, D.A, f, V.dinv, c, nxt);
            c = nxt;
        }
    }
}

// the cycle on the handle's buffers: FP64 vin -> vout, or the FP32 cycle on
// vin32 -> vout32
void vcycle_any(ngap_uns *h) {
    if (h->o.precision == NGAP_MIXED)
        vcycle<float>(h, h->vin32, h->vout32);
    else
        vcycle<double>(h, h->vin, h->vout);
}

int run_vcycle(ngap_uns *h) {
    const bool timed = h->time_next;
    if (timed) UCHK(hipEventRecord(h->e0, h->stream));
    const bool mixed = h->o.precision == NGAP_MIXED;
    if (mixed)
        hipLaunchKernelGGL(k_to_f32, dim3(vec_blocks(h->rows)), dim3(256), 0, h->stream, h->rows, h->vin, h->vin32);
    if (h->vgraph) {
        UCHK(hipGraphLaunch(h->vgraph, h->stream));
    } else {
        vcycle_any(h);
        UCHK(hipGetLastError());
    }
    if (mixed)
        hipLaunchKernelGGL(k_to_f64, dim3(vec_blocks(h->rows)), dim3(256), 0, h->stream, h->rows, h->vout32, h->vout);
    if (timed) {
        UCHK(hipEventRecord(h->e1, h->stream));
        h->time_next = false;
        h->timed = true;
    }
    return 0;
}

// vout = M vin for the method's preconditioner
int precond(ngap_uns *h) {
    if (is_amg(h->method)) return run_vcycle(h);
    hipLaunchKernelGGL(k_mul, dim3(vec_blocks(h->rows)), dim3(256), 0, h->stream, h->rows, h->vin, h->dgi, h->vout);
    UCHK(hipGetLastError());
    return 0;
}

// <a, b> on the stream, result to the host
int dot(ngap_uns *h, const double *a, const double *b, double *out) {
    hipLaunchKernelGGL(k_dot, dim3(NPART), dim3(256), 0, h->stream, h->rows, a, b, h->part);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(256), 0, h->stream, h->part, NPART, h->scal, 0);
    UCHK(hipMemcpyAsync(h->h_scal, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    UCHK(hipStreamSynchronize(h->stream));
    *out = h->h_scal[0];
    return 0;
}

void axpby(ngap_uns *h, double a, const double *x, double b, double *y) {
    hipLaunchKernelGGL(k_axpby, dim3(vec_blocks(h->rows)), dim3(256), 0, h->stream, h->rows, a, x, b, y);
}
void lin3(ngap_uns *h, const double *x, double a, const double *y, double b, const double *w, double *z) {
    hipLaunchKernelGGL(k_lin3, dim3(vec_blocks(h->rows)), dim3(256), 0, h->stream, h->rows, x, a, y, b, w, z);
}
void copy(ngap_uns *h, const double *x, double *z) { lin3(h, x, 0.0, nullptr, 0.0, nullptr, z); }

// r = b - A x, ||r||
int residual(ngap_uns *h, double *nrm) {
    csr_resid(h->stream, h->L[0].A, h->b, h->x, h->r);
    double rr;
    RCHK(dot(h, h->r, h->r, &rr));
    *nrm = std::sqrt(rr);
    return 0;
}

// ------------------------------------------------------------------ Krylov drivers
// r is kept in vin, so PCG's z = M r needs no copy. alpha, beta and rho stay
// on the device; the host reads <p, Ap> and <r, r> once per iteration for the
// breakdown and stopping tests.
int run_pcg(ngap_uns *h, double bn, int &it) {
    hipStream_t s = h->stream;
    const DCsr &A = h->L[0].A;
    const int n = h->rows, nb = vec_blocks(n);
    double *r = h->vin, *z = h->vout;
    csr_resid(s, A, h->b, h->x, r);
    double rr;
    RCHK(dot(h, r, r, &rr));
    if (std::sqrt(rr) / bn < h->rcvg) return 0;
    RCHK(precond(h));
    copy(h, z, h->p);
    hipLaunchKernelGGL(k_dot, dim3(NPART), dim3(256), 0, s, n, r, z, h->part);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(256), 0, s, h->part, NPART, h->scal, 0);
    while (it < h->maxit) {
        csr_mul(s, A, h->p, h->q);
        hipLaunchKernelGGL(k_dot, dim3(NPART), dim3(256), 0, s, n, h->p, h->q, h->part);
        hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(256), 0, s, h->part, NPART, h->scal);
        hipLaunchKernelGGL(k_pcg_xr, dim3(nb), dim3(256), 0, s, n, h->scal, h->p, h->q, h->x, r, h->part);
        hipLaunchKernelGGL(k_fin, dim3(1), dim3(256), 0, s, h->part, nb, h->scal, 2);
        UCHK(hipMemcpyAsync(h->h_scal, h->scal, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
        UCHK(hipStreamSynchronize(s));
        const double pq = h->h_scal[1];
        rr = h->h_scal[2];
        if (!(pq > 0.0) || !std::isfinite(pq)) return NGAP_ERR_BREAKDOWN;
        it++;
        if (std::sqrt(rr) / bn < h->rcvg) break;
        RCHK(precond(h));
        if (h->o.precision == NGAP_MIXED && is_amg(h->method)) {
            // the FP32 cycle is not exactly a fixed symmetric operator: flexible beta
            hipLaunchKernelGGL(k_dot2, dim3(NPART), dim3(256), 0, s, n, r, z, h->q, h->part);
            hipLaunchKernelGGL(k_fin_beta_flex, dim3(1), dim3(256), 0, s, h->part, NPART, h->scal);
        } else {
            hipLaunchKernelGGL(k_dot, dim3(NPART), dim3(256), 0, s, n, r, z, h->part);
            hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(256), 0, s, h->part, NPART, h->scal);
        }
        hipLaunchKernelGGL(k_pcg_p, dim3(nb), dim3(256), 0, s, n, h->scal, z, h->p);
    }
    return 0;
}

int run_amg(ngap_uns *h, double bn, int &it) {
    double nr;
    csr_resid(h->stream, h->L[0].A, h->b, h->x, h->vin);
    RCHK(dot(h, h->vin, h->vin, &nr));
    nr = std::sqrt(nr);
    while (nr / bn >= h->rcvg && it < h->maxit) {
        RCHK(run_vcycle(h));
        axpby(h, 1.0, h->vout, 1.0, h->x);
        csr_resid(h->stream, h->L[0].A, h->b, h->x, h->vin);
        RCHK(dot(h, h->vin, h->vin, &nr));
        nr = std::sqrt(nr);
        it++;
    }
    return 0;
}

// restarted GMRES(kdim), right preconditioning, modified Gram-Schmidt,
// Givens rotations; the true residual is taken at every restart
int run_gmres(ngap_uns *h, double bn, int &it) {
    const int m = h->o.kdim, n = h->rows;
    const DCsr &A = h->L[0].A;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    double nr;
    RCHK(residual(h, &nr));
    while (nr / bn >= h->rcvg && it < h->maxit) {
        double *V0 = h->V;
        axpby(h, 1.0 / nr, h->r, 0.0, V0);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = nr;
        int k = 0;
        for (; k < m && it < h->maxit; k++) {
            double *Vk = h->V + (size_t)k * n, *Vk1 = h->V + (size_t)(k + 1) * n, *Zk = h->Z + (size_t)k * n;
            copy(h, Vk, h->vin);
            RCHK(precond(h));
            copy(h, h->vout, Zk);
            csr_mul(h->stream, A, Zk, Vk1);
            for (int j = 0; j <= k; j++) {
                double hj;
                RCHK(dot(h, h->V + (size_t)j * n, Vk1, &hj));
                H[(size_t)j * m + k] = hj;
                axpby(h, -hj, h->V + (size_t)j * n, 1.0, Vk1);
            }
            double hk;
            RCHK(dot(h, Vk1, Vk1, &hk));
            hk = std::sqrt(hk);
            H[(size_t)(k + 1) * m + k] = hk;
            for (int j = 0; j < k; j++) {
                const double a = H[(size_t)j * m + k], b = H[(size_t)(j + 1) * m + k];
                H[(size_t)j * m + k] = cs[j] * a + sn[j] * b;
                H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
            }
            const double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
            const double rn = std::hypot(a, b);
            cs[k] = rn > 0.0 ? a / rn : 1.0;
            sn[k] = rn > 0.0 ? b / rn : 0.0;
            H[(size_t)k * m + k] = rn;
            H[(size_t)(k + 1) * m + k] = 0.0;
            gv[k + 1] = -sn[k] * gv[k];
            gv[k] = cs[k] * gv[k];
            it++;
            if (!(rn > 0.0)) return NGAP_ERR_BREAKDOWN;
            if (std::fabs(gv[k + 1]) / bn < h->rcvg || hk == 0.0) {
                k++;
                break;
            }
            axpby(h, 0.0, Vk1, 1.0 / hk, Vk1);
        }
        for (int j = k - 1; j >= 0; j--) {
            double sacc = gv[j];
            for (int l = j + 1; l < k; l++) sacc -= H[(size_t)j * m + l] * y[l];
            y[j] = sacc / H[(size_t)j * m + j];
        }
        for (int j = 0; j < k; j++) axpby(h, y[j], h->Z + (size_t)j * n, 1.0, h->x);
        RCHK(residual(h, &nr));
    }
    return 0;
}

// right-preconditioned BiCGStab
int run_bicgstab(ngap_uns *h, double bn, int &it) {
    const DCsr &A = h->L[0].A;
    hipStream_t s = h->stream;
    double nr;
    RCHK(residual(h, &nr));
    copy(h, h->r, h->rh);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    hipMemsetAsync(h->v, 0, sizeof(double) * h->rows, s);
    hipMemsetAsync(h->p, 0, sizeof(double) * h->rows, s);
    while (nr / bn >= h->rcvg && it < h->maxit) {
        double rho1;
        RCHK(dot(h, h->rh, h->r, &rho1));
        if (rho1 == 0.0 || !std::isfinite(rho1)) return NGAP_ERR_BREAKDOWN;
        const double beta = (rho1 / rho) * (alpha / omega);
        // p = r + beta (p - omega v)
        lin3(h, h->r, beta, h->p, -beta * omega, h->v, h->p);
        copy(h, h->p, h->vin);
        RCHK(precond(h));
        copy(h, h->vout, h->ph);
        csr_mul(s, A, h->ph, h->v);
        double rv;
        RCHK(dot(h, h->rh, h->v, &rv));
        if (rv == 0.0 || !std::isfinite(rv)) return NGAP_ERR_BREAKDOWN;
        alpha = rho1 / rv;
        lin3(h, h->r, -alpha, h->v, 0.0, nullptr, h->s);
        it++;
        double ss;
        RCHK(dot(h, h->s, h->s, &ss));
        if (std::sqrt(ss) / bn < h->rcvg) {
            axpby(h, alpha, h->ph, 1.0, h->x);
            copy(h, h->s, h->r);
            nr = std::sqrt(ss);
            break;
        }
        copy(h, h->s, h->vin);
        RCHK(precond(h));
        copy(h, h->vout, h->sh);
        csr_mul(s, A, h->sh, h->t);
        double ts, tt;
        RCHK(dot(h, h->t, h->s, &ts));
        RCHK(dot(h, h->t, h->t, &tt));
        if (tt == 0.0) return NGAP_ERR_BREAKDOWN;
        omega = ts / tt;
        lin3(h, h->x, alpha, h->ph, omega, h->sh, h->x);
        lin3(h, h->s, -omega, h->t, 0.0, nullptr, h->r);
        RCHK(dot(h, h->r, h->r, &nr));
        nr = std::sqrt(nr);
        rho = rho1;
        if (omega == 0.0) return NGAP_ERR_BREAKDOWN;
    }
    return 0;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int do_solve(ngap_uns *h, int *it, double *rerr, double *aerr) {
    const double t0 = now_ms();
    int iters = 0;
    double bb;
    RCHK(dot(h, h->b, h->b, &bb));
    const double bn = std::sqrt(bb);
    if (bn == 0.0) {
        hipMemsetAsync(h->x, 0, sizeof(double) * h->rows, h->stream);
        h->st.rerr = 0.0;
        h->st.aerr = 0.0;
    } else {
        // the first V-cycle of the solve is timed with events on the stream
        h->time_next = is_amg(h->method);
        h->timed = false;
        int rc = 0;
        switch (h->method) {
            case NGAP_UNS_AMG: rc = run_amg(h, bn, iters); break;
            case NGAP_UNS_PCG_AMG:
            case NGAP_UNS_PCG: rc = run_pcg(h, bn, iters); break;
            case NGAP_UNS_GMRES_AMG:
            case NGAP_UNS_GMRES: rc = run_gmres(h, bn, iters); break;
            case NGAP_UNS_BICGSTAB_AMG: rc = run_bicgstab(h, bn, iters); break;
            default: rc = NGAP_ERR_METHOD;
        }
        if (rc) return rc;
        double nr;
        RCHK(residual(h, &nr));
        h->st.rerr = nr / bn;
        h->st.aerr = nr;
        if (h->timed) {
            float ms = 0.f;
            UCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
            h->st.vcycle_ms = ms;
        }
    }
    h->st.it = iters;
    h->st.status = h->st.rerr < h->rcvg ? NGAP_STATUS_CONVERGED : NGAP_STATUS_MAXIT;
    h->st.solve_ms = now_ms() - t0;
    if (it) *it = iters;
    if (rerr) *rerr = h->st.rerr;
    if (aerr) *aerr = h->st.aerr;
    if (h->o.verbose)
        fprintf(stderr, "[ngap_uns] method %d: it %d, rerr %.3e, %.3f ms\n", h->method, iters, h->st.rerr,
                h->st.solve_ms);
    return 0;
}

int setup_impl(ngap_uns *h, const double *opr, int maxit, double rcvg, const ngap_uns_opts *o) {
    if (!h || !opr) return NGAP_ERR_ARG;
    if (!h->inited) return NGAP_ERR_NOT_INIT;
    if (h->method == NGAP_UNS_PCG_PARASAIL || h->method == NGAP_UNS_GMRES_PILUT) return NGAP_ERR_METHOD;
    if (maxit < 0 || !(rcvg >= 0.0)) return NGAP_ERR_ARG;
    ngap_uns_opts op;
    ngap_uns_default_opts(&op);
    if (o) op = *o;
    if (!ngap::amg::opts_valid(op)) return NGAP_ERR_ARG;
    if (h->rows < 1) return NGAP_ERR_ARG;
    const double t0 = now_ms();
    free_device(h);
    if (h->hier) {
        ngap_amg_free(h->hier);
        h->hier = nullptr;
    }
    h->o = op;
    h->maxit = maxit;
    h->rcvg = rcvg;
    // the matrix (hypre_uns_class.f90:186-230)
    ngap::amg::Csr A;
    A.n = A.ncol = h->rows;
    A.rp.resize(h->rows + 1);
    const long long nnz = ngap::assemble(h->g, h->nst, h->stc.data(), h->ind.data(), opr, nullptr, nullptr, nullptr);
    if (nnz >= (1LL << 31)) return NGAP_ERR_ARG;
    A.col.resize(nnz);
    A.val.resize(nnz);
    ngap::assemble(h->g, h->nst, h->stc.data(), h->ind.data(), opr, A.rp.data(), A.col.data(), A.val.data());
    for (double v : A.val)
        if (!std::isfinite(v)) return NGAP_ERR_NONFINITE;
    std::vector<double> dgi(h->rows);
    for (int i = 0; i < h->rows; i++) {
        double d = 0.0;
        for (int q = A.rp[i]; q < A.rp[i + 1]; q++)
            if (A.col[q] == i) d += A.val[q];
        dgi[i] = d != 0.0 ? 1.0 / d : 0.0;
    }
    // the hierarchy (host)
    auto *H = new (std::nothrow) ngap_amg;
    if (!H) return NGAP_ERR_ALLOC;
    const double th = now_ms();
    int rc = 0;
    if (is_amg(h->method)) {
        rc = ngap::amg::build_checked(H->H, std::move(A), op);
    } else {
        H->H.A.push_back(std::move(A));
        H->H.d.emplace_back();
    }
    h->st.hier_ms = now_ms() - th;
    if (rc) {
        delete H;
        return rc;
    }
    h->hier = H;
    const ngap::amg::Hier &Hh = H->H;
    // device copies
    const int nl = (int)Hh.A.size();
    h->L.resize(nl);
    long long nnz_all = 0;
    const bool mixed = is_amg(h->method) && op.precision == NGAP_MIXED;
    for (int l = 0; l < nl; l++) {
        DLevel &D = h->L[l];
        RCHK(upload_csr(h, D.A, Hh.A[l], mixed));
        nnz_all += D.A.nnz;
        if (is_amg(h->method)) {
            if (l < nl - 1) {
                RCHK(upload_csr(h, D.P, Hh.P[l], mixed));
                RCHK(upload_csr(h, D.R, ngap::amg::transpose(Hh.P[l]), mixed));
            }
            if (mixed)
                RCHK(level_vectors<float>(h, D.s, Hh.d[l], l));
            else
                RCHK(level_vectors<double>(h, D.d, Hh.d[l], l));
        }
    }
    if (is_amg(h->method)) {
        RCHK(upload(h, &h->Ainv, Hh.Ainv));
        h->nc = Hh.A.back().n;
    }
    RCHK(upload(h, &h->dgi, dgi));
    RCHK(upload(h, &h->d_map, h->map));
    const size_t vb = sizeof(double) * (size_t)h->rows;
    double **vecs[] = {&h->b, &h->x, &h->r, &h->vin, &h->vout, &h->p, &h->q, &h->s, &h->t, &h->v, &h->rh, &h->ph, &h->sh};
    for (double **pv : vecs) RCHK(dalloc(h, (void **)pv, vb));
    if (mixed) {
        RCHK(dalloc(h, (void **)&h->vin32, sizeof(float) * (size_t)h->rows));
        RCHK(dalloc(h, (void **)&h->vout32, sizeof(float) * (size_t)h->rows));
    }
    if (h->method == NGAP_UNS_GMRES || h->method == NGAP_UNS_GMRES_AMG) {
        RCHK(dalloc(h, (void **)&h->V, vb * (op.kdim + 1)));
        RCHK(dalloc(h, (void **)&h->Z, vb * op.kdim));
    }
    RCHK(dalloc(h, (void **)&h->grhs, sizeof(double) * h->ntot));
    RCHK(dalloc(h, (void **)&h->gsol, sizeof(double) * h->ntot));
    RCHK(dalloc(h, (void **)&h->part, sizeof(double) * MAX_VBLOCKS));
    RCHK(dalloc(h, (void **)&h->scal, sizeof(double) * 8));
    UCHK(hipHostMalloc((void **)&h->h_scal, sizeof(double) * 8));
    // the V-cycle as a graph on the handle's stream
    if (is_amg(h->method) && op.graphs) {
        hipGraph_t gr = nullptr;
        UCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        vcycle_any(h);
        UCHK(hipStreamEndCapture(h->stream, &gr));
        const hipError_t e = hipGraphInstantiate(&h->vgraph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            h->vgraph = nullptr;
            UCHK(e);
        }
    }
    UCHK(hipStreamSynchronize(h->stream));
    h->st.nlevels = nl;
    h->st.coarse_singular = Hh.singular;
    h->st.rows = h->rows;
    h->st.nnz = h->L[0].A.nnz;
    h->st.op_complexity = (double)nnz_all / (double)h->L[0].A.nnz;
    h->setup_done = 1;
    h->st.setup_ms = now_ms() - t0;
    if (op.verbose)
        fprintf(stderr, "[ngap_uns] setup: %d rows, %d levels, operator complexity %.2f, %.1f ms (hierarchy %.1f)\n",
                h->rows, nl, h->st.op_complexity, h->st.setup_ms, h->st.hier_ms);
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int ngap_uns_create(ngap_uns **hp, const ngap_grid *g, int nst, const int *stc, int method) {
    if (!hp || !g || !stc || nst < 1 || nst > 64) return NGAP_ERR_ARG;
    if (method < NGAP_UNS_AMG || method > NGAP_UNS_GMRES_PILUT) return NGAP_ERR_METHOD;
    const long long nxo = g->imaxo_ - g->imino_ + 1, nyo = g->jmaxo_ - g->jmino_ + 1, nzo = g->kmaxo_ - g->kmino_ + 1;
    if (g->imax_ - g->imin_ + 1 != g->nx || g->jmax_ - g->jmin_ + 1 != g->ny || g->kmax_ - g->kmin_ + 1 != g->nz)
        return NGAP_ERR_ARG;  // one rank owns the whole grid
    if (nxo < 1 || nyo < 1 || nzo < 1 || nxo * nyo * nzo >= (1LL << 31)) return NGAP_ERR_ARG;
    for (int st = 0; st < nst; st++) {
        const int di = stc[st], dj = stc[nst + st], dk = stc[2 * nst + st];
        if (g->imin_ + di < g->imino_ || g->imax_ + di > g->imaxo_ || g->jmin_ + dj < g->jmino_ ||
            g->jmax_ + dj > g->jmaxo_ || g->kmin_ + dk < g->kmino_ || g->kmax_ + dk > g->kmaxo_)
            return NGAP_ERR_STENCIL;  // an offset reaches past the ghost layer
    }
    auto *h = new (std::nothrow) ngap_uns;
    if (!h) return NGAP_ERR_ALLOC;
    h->g = *g;
    h->nst = nst;
    h->stc.assign(stc, stc + 3 * nst);
    h->method = method;
    h->ntot = nxo * nyo * nzo;
    ngap_uns_default_opts(&h->o);
    if (hipStreamCreate(&h->stream) != hipSuccess || hipEventCreate(&h->e0) != hipSuccess ||
        hipEventCreate(&h->e1) != hipSuccess) {
        delete h;
        return NGAP_ERR_HIP;
    }
    *hp = h;
    return 0;
}

int ngap_uns_init(ngap_uns *h, const double *opr) {
    if (!h || !opr) return NGAP_ERR_ARG;
    try {
        h->ind.assign(h->ntot, -1);
        h->map.reserve(h->ntot);
    } catch (const std::bad_alloc &) {
        return NGAP_ERR_ALLOC;
    }
    const long long n = ngap::umap(h->g, opr, h->nst, h->ind.data());
    if (n < 0 || n >= (1LL << 31)) return NGAP_ERR_ARG;
    h->rows = (int)n;
    h->map.clear();
    const long long nxo = h->g.imaxo_ - h->g.imino_ + 1, nyo = h->g.jmaxo_ - h->g.jmino_ + 1;
    for (int k = h->g.kmin_; k <= h->g.kmax_; k++)
        for (int j = h->g.jmin_; j <= h->g.jmax_; j++)
            for (int i = h->g.imin_; i <= h->g.imax_; i++) {
                const long long p = (i - h->g.imino_) + nxo * ((j - h->g.jmino_) + nyo * (long long)(k - h->g.kmino_));
                if (h->ind[p] >= 0) h->map.push_back((int)p);
            }
    h->inited = 1;
    h->setup_done = 0;
    return 0;
}

int ngap_uns_setup(ngap_uns *h, const double *opr, int maxit, double rcvg, const ngap_uns_opts *o) {
    // host allocations (the matrix, the hierarchy) report failure as a code;
    // the handle is then left without a setup
    try {
        return setup_impl(h, opr, maxit, rcvg, o);
    } catch (const std::bad_alloc &) {
        return NGAP_ERR_ALLOC;
    }
}

int ngap_uns_solve_device(ngap_uns *h, const double *d_rhs, double *d_sol, int *it, double *rerr, double *aerr) {
    if (!h || !d_rhs || !d_sol) return NGAP_ERR_ARG;
    if (!h->setup_done) return NGAP_ERR_NOT_SETUP;
    if (it) *it = -1;
    UCHK(hipDeviceSynchronize());
    const int nb = vec_blocks(h->rows);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, h->stream, h->rows, h->d_map, d_rhs, h->b);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, h->stream, h->rows, h->d_map, (const double *)d_sol, h->x);
    RCHK(do_solve(h, it, rerr, aerr));
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, h->stream, h->rows, h->d_map, h->x, d_sol);
    UCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int ngap_uns_solve(ngap_uns *h, const double *rhs, double *sol, int *it, double *rerr, double *aerr) {
    if (!h || !rhs || !sol) return NGAP_ERR_ARG;
    if (!h->setup_done) return NGAP_ERR_NOT_SETUP;
    const size_t bytes = sizeof(double) * h->ntot;
    UCHK(hipMemcpyAsync(h->grhs, rhs, bytes, hipMemcpyHostToDevice, h->stream));
    UCHK(hipMemcpyAsync(h->gsol, sol, bytes, hipMemcpyHostToDevice, h->stream));
    UCHK(hipStreamSynchronize(h->stream));
    RCHK(ngap_uns_solve_device(h, h->grhs, h->gsol, it, rerr, aerr));
    UCHK(hipMemcpy(sol, h->gsol, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int ngap_uns_vcycle_device(ngap_uns *h, const double *d_b, double *d_x) {
    if (!h || !d_b || !d_x) return NGAP_ERR_ARG;
    if (!h->setup_done) return NGAP_ERR_NOT_SETUP;
    if (!is_amg(h->method)) return NGAP_ERR_METHOD;
    UCHK(hipDeviceSynchronize());
    copy(h, d_b, h->vin);
    RCHK(run_vcycle(h));
    copy(h, h->vout, d_x);
    UCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int ngap_uns_destroy(ngap_uns *h) {
    if (!h) return NGAP_ERR_ARG;
    (void)hipDeviceSynchronize();
    free_device(h);
    if (h->hier) ngap_amg_free(h->hier);
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

long long ngap_uns_rows(const ngap_uns *h) { return h ? (long long)h->rows : NGAP_ERR_ARG; }

int ngap_uns_index(const ngap_uns *h, int *ind) {
    if (!h || !ind) return NGAP_ERR_ARG;
    if (!h->inited) return NGAP_ERR_NOT_INIT;
    std::memcpy(ind, h->ind.data(), sizeof(int) * h->ntot);
    return 0;
}

int ngap_uns_last_stats(const ngap_uns *h, ngap_uns_stats *s) {
    if (!h || !s) return NGAP_ERR_ARG;
    *s = h->st;
    return 0;
}

const ngap_amg *ngap_uns_hierarchy(const ngap_uns *h) { return h ? h->hier : nullptr; }

}  // extern "C"
