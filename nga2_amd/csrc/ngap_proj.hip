// Device-side projection: operator assembly, divergence, pressure gradient,
// mean removal and ghost fill over the caller's ghosted device arrays (the
// formulas of incomp_class.f90 / tpns_class.f90 cited in ngapproj.h).
//
// Every floating-point expression is evaluated in the order the reference
// writes it, with contraction into FMA turned off, so the results are the
// IEEE values of that order (bit-identical to a plain FP64 evaluation on the
// host). The only exception is the fixed-order block reduction of the mean
// in shift_p.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ngapproj.h"

#pragma clang fp contract(off)

namespace {

constexpr int PT = 256;

struct Box {
    int nxo, nyo, nzo;      // ghosted extents
    int no_x, no_y, no_z;   // ghost widths (imin_ - imino_ ...)
    int nx, ny, nz;         // interior extents
    int px, py, pz;         // periodic
    long long No;           // nxo * nyo * nzo
};

__device__ __forceinline__ void split(const Box &b, long long q, int &i, int &j, int &k) {
    i = (int)(q % b.nxo);
    const long long r = q / b.nxo;
    j = (int)(r % b.nyo);
    k = (int)(r / b.nyo);
}

__device__ __forceinline__ bool interior(const Box &b, int i, int j, int k) {
    return i >= b.no_x && i < b.no_x + b.nx && j >= b.no_y && j < b.no_y + b.ny && k >= b.no_z && k < b.no_z + b.nz;
}

int blocks_for(long long n) {
    long long v = (n + PT - 1) / PT;
    if (v > 65535LL * 16) v = 65535LL * 16;
    return (int)(v < 1 ? 1 : v);
}

// opr over all ghosted cells (interior: the formula; ghosts: 0)
template <bool RHO>
__global__ void __launch_bounds__(PT) k_laplacian(Box b, const double *__restrict__ dpx, const double *__restrict__ dpy,
                                                  const double *__restrict__ dpz, const double *__restrict__ dux,
                                                  const double *__restrict__ duy, const double *__restrict__ duz,
                                                  const double *__restrict__ vol, const double *__restrict__ ru,
                                                  const double *__restrict__ rv, const double *__restrict__ rw,
                                                  double *__restrict__ opr) {
    const long long sx = 1, sy = b.nxo, sz = (long long)b.nxo * b.nyo;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        double *o = opr + 7 * q;
        if (!interior(b, i, j, k)) {
            for (int s = 0; s < 7; s++) o[s] = 0.0;
            continue;
        }
        const double *dp[3] = {dpx, dpy, dpz};
        const double *du[3] = {dux, duy, duz};
        const double *rf[3] = {ru, rv, rw};
        const long long st[3] = {sx, sy, sz};
        double cu[3], cl[3], ou[3], ol[3];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const long long e = q + st[ax];
            const double p0 = dp[ax][2 * q], p1 = dp[ax][2 * q + 1];
            const double u0 = du[ax][2 * q], u1 = du[ax][2 * q + 1];
            const double u0up = du[ax][2 * e], u1up = du[ax][2 * e + 1];
            if (RHO) {
                const double rup = rf[ax][e], rlo = rf[ax][q];
                cu[ax] = p1 * u0up / rup;
                cl[ax] = p0 * u1 / rlo;
                ou[ax] = p1 * u1up / rup;
                ol[ax] = p0 * u0 / rlo;
            } else {
                cu[ax] = p1 * u0up;
                cl[ax] = p0 * u1;
                ou[ax] = p1 * u1up;
                ol[ax] = p0 * u0;
            }
        }
        const double diag = cu[0] + cl[0] + cu[1] + cl[1] + cu[2] + cl[2];
        const double vq = vol[q];
        o[0] = -diag * vq;
        o[1] = -ou[0] * vq;
        o[2] = -ol[0] * vq;
        o[3] = -ou[1] * vq;
        o[4] = -ol[1] * vq;
        o[5] = -ou[2] * vq;
        o[6] = -ol[2] * vq;
    }
}

__global__ void __launch_bounds__(PT) k_div(Box b, const double *__restrict__ dpx, const double *__restrict__ dpy,
                                            const double *__restrict__ dpz, const double *__restrict__ u,
                                            const double *__restrict__ v, const double *__restrict__ w,
                                            double *__restrict__ div) {
    const long long sy = b.nxo, sz = (long long)b.nxo * b.nyo;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (!interior(b, i, j, k)) {
            div[q] = 0.0;
            continue;
        }
        const double tx = dpx[2 * q] * u[q] + dpx[2 * q + 1] * u[q + 1];
        const double ty = dpy[2 * q] * v[q] + dpy[2 * q + 1] * v[q + sy];
        const double tz = dpz[2 * q] * w[q] + dpz[2 * q + 1] * w[q + sz];
        div[q] = tx + ty + tz;
    }
}

// g = divu(-1) P(c - e) + divu(0) P(c) along one axis
__global__ void __launch_bounds__(PT) k_pgrad(Box b, const double *__restrict__ du, long long stride,
                                              const double *__restrict__ pres, double *__restrict__ g) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (!interior(b, i, j, k)) {
            g[q] = 0.0;
            continue;
        }
        g[q] = du[2 * q] * pres[q - stride] + du[2 * q + 1] * pres[q];
    }
}

// per-block partials of sum(vol * VF * A) (or of vol * VF with A == NULL) over interior cells
__global__ void __launch_bounds__(PT) k_integrate(Box b, const double *__restrict__ vol, const double *__restrict__ vf,
                                                  const double *__restrict__ a, double *__restrict__ part) {
    __shared__ double sm[PT / 64];
    double acc = 0.0;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (!interior(b, i, j, k)) continue;
        const double t = vol[q] * vf[q];
        acc += a ? t * a[q] : t;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < PT / 64; w++) s += sm[w];
        part[blockIdx.x] = s;
    }
}

// out[0] = sum(part[0..n)) in fixed order
__global__ void __launch_bounds__(PT) k_sum(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double sm[PT];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += PT) a += part[i];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int w = PT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

// P -= tot on interior cells with VF > 0, tot = num[0] / fluid volume
__global__ void __launch_bounds__(PT) k_shift(Box b, const double *__restrict__ vf, const double *__restrict__ num,
                                              double fvol, double *__restrict__ pres) {
    const double tot = num[0] / fvol;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (interior(b, i, j, k) && vf[q] > 0.0) pres[q] = pres[q] - tot;
    }
}

// periodic ghost fill along one axis: ghost layer g (0..no-1) below and above
// takes the interior cell at the periodic image, every other index unchanged
__global__ void __launch_bounds__(PT) k_sync(Box b, int ax, double *__restrict__ a) {
    const int n = ax == 0 ? b.nx : (ax == 1 ? b.ny : b.nz);
    const int no = ax == 0 ? b.no_x : (ax == 1 ? b.no_y : b.no_z);
    const long long st = ax == 0 ? 1 : (ax == 1 ? (long long)b.nxo : (long long)b.nxo * b.nyo);
    // cells of one ghost slab: the extents of the other two axes
    const long long e0 = ax == 0 ? b.nyo : b.nxo;
    const long long e1 = ax == 2 ? b.nyo : b.nzo;
    const long long slab = e0 * e1;
    const long long total = slab * 2 * no;
    for (long long t = blockIdx.x * (long long)PT + threadIdx.x; t < total; t += (long long)gridDim.x * PT) {
        const long long s = t % slab;
        const int gl = (int)(t / slab);  // 0..2no-1: low ghosts then high ghosts
        const long long c0 = s % e0, c1 = s / e0;
        long long base;
        if (ax == 0) base = (long long)b.nxo * (c0 + (long long)b.nyo * c1);
        else if (ax == 1) base = c0 + (long long)b.nxo * b.nyo * c1;
        else base = c0 + (long long)b.nxo * c1;
        int dst, src;
        if (gl < no) {
            const int g = gl;
            dst = g;
            src = no + (((g - no) % n) + n) % n;
        } else {
            const int g = gl - no;
            dst = n + no + g;
            src = no + (g % n);
        }
        a[base + dst * st] = a[base + src * st];
    }
}

}  // namespace

struct ngap_proj {
    Box b{};
    int device = 0;
    hipStream_t s = nullptr;
    double *dp[3] = {nullptr, nullptr, nullptr};
    double *du[3] = {nullptr, nullptr, nullptr};
    double *vol = nullptr, *vf = nullptr;
    double *part = nullptr, *num = nullptr;
    int nb = 0;
    double fvol = 0.0;
    std::vector<void *> allocs;
    ~ngap_proj() {
        for (void *q : allocs) (void)hipFree(q);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

int up(ngap_proj *P, double **d, const double *h, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(double) * (n ? n : 1)) != hipSuccess) return NGAP_ERR_ALLOC;
    P->allocs.push_back(q);
    *d = (double *)q;
    if (h && hipMemcpy(q, h, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) return NGAP_ERR_HIP;
    return 0;
}

int fin(ngap_proj *P) {
    if (hipGetLastError() != hipSuccess) return NGAP_ERR_HIP;
    return hipStreamSynchronize(P->s) == hipSuccess ? 0 : NGAP_ERR_HIP;
}

int sync_all(ngap_proj *P, double *a) {
    const Box &b = P->b;
    const int per[3] = {b.px, b.py, b.pz};
    const int no[3] = {b.no_x, b.no_y, b.no_z};
    for (int ax = 0; ax < 3; ax++) {
        if (!per[ax] || no[ax] == 0) continue;
        const long long e0 = ax == 0 ? b.nyo : b.nxo;
        const long long e1 = ax == 2 ? b.nyo : b.nzo;
        hipLaunchKernelGGL(k_sync, dim3(blocks_for(e0 * e1 * 2 * no[ax])), dim3(PT), 0, P->s, b, ax, a);
    }
    return 0;
}

}  // namespace

extern "C" {

int ngap_proj_create(ngap_proj **pp, const ngap_grid *g, const double *divp_x, const double *divp_y,
                     const double *divp_z, const double *divu_x, const double *divu_y, const double *divu_z,
                     const double *vol, const double *vf) {
    if (!pp) return NGAP_ERR_ARG;
    *pp = nullptr;
    if (!g || !divp_x || !divp_y || !divp_z || !divu_x || !divu_y || !divu_z || !vol || !vf) return NGAP_ERR_ARG;
    Box b{};
    b.nxo = g->imaxo_ - g->imino_ + 1;
    b.nyo = g->jmaxo_ - g->jmino_ + 1;
    b.nzo = g->kmaxo_ - g->kmino_ + 1;
    b.no_x = g->imin_ - g->imino_;
    b.no_y = g->jmin_ - g->jmino_;
    b.no_z = g->kmin_ - g->kmino_;
    b.nx = g->imax_ - g->imin_ + 1;
    b.ny = g->jmax_ - g->jmin_ + 1;
    b.nz = g->kmax_ - g->kmin_ + 1;
    b.px = g->xper;
    b.py = g->yper;
    b.pz = g->zper;
    if (b.nx < 1 || b.ny < 1 || b.nz < 1 || b.no_x < 0 || b.no_y < 0 || b.no_z < 0) return NGAP_ERR_ARG;
    if (b.nx != g->nx || b.ny != g->ny || b.nz != g->nz) return NGAP_ERR_ARG;  // one GPU holds the whole grid
    if (b.nxo != b.nx + 2 * b.no_x || b.nyo != b.ny + 2 * b.no_y || b.nzo != b.nz + 2 * b.no_z) return NGAP_ERR_ARG;
    // the stencils read one cell past the interior in each direction
    if (b.no_x < 1 || b.no_y < 1 || b.no_z < 1) return NGAP_ERR_ARG;
    b.No = (long long)b.nxo * b.nyo * b.nzo;
    if (7 * b.No >= (1LL << 40)) return NGAP_ERR_ARG;
    ngap_proj *P = new ngap_proj();
    P->b = b;
    int rc = hipGetDevice(&P->device) == hipSuccess ? 0 : NGAP_ERR_HIP;
    if (!rc) rc = hipStreamCreateWithFlags(&P->s, hipStreamNonBlocking) == hipSuccess ? 0 : NGAP_ERR_HIP;
    const double *hp[3] = {divp_x, divp_y, divp_z}, *hu[3] = {divu_x, divu_y, divu_z};
    for (int ax = 0; ax < 3 && !rc; ax++) {
        rc = up(P, &P->dp[ax], hp[ax], 2 * (size_t)b.No);
        if (!rc) rc = up(P, &P->du[ax], hu[ax], 2 * (size_t)b.No);
    }
    if (!rc) rc = up(P, &P->vol, vol, (size_t)b.No);
    if (!rc) rc = up(P, &P->vf, vf, (size_t)b.No);
    P->nb = 1024;
    if (!rc) rc = up(P, &P->part, nullptr, (size_t)P->nb);
    if (!rc) rc = up(P, &P->num, nullptr, 1);
    if (!rc) {
        hipLaunchKernelGGL(k_integrate, dim3(P->nb), dim3(PT), 0, P->s, b, (const double *)P->vol,
                           (const double *)P->vf, (const double *)nullptr, P->part);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(PT), 0, P->s, (const double *)P->part, P->nb, P->num);
        rc = fin(P);
    }
    if (!rc && hipMemcpy(&P->fvol, P->num, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = NGAP_ERR_HIP;
    if (rc) {
        delete P;
        return rc;
    }
    *pp = P;
    return 0;
}

int ngap_proj_destroy(ngap_proj *p) {
    if (!p) return NGAP_ERR_ARG;
    if (p->s) (void)hipStreamSynchronize(p->s);
    delete p;
    return 0;
}

int ngap_proj_laplacian(ngap_proj *p, const double *ru, const double *rv, const double *rw, double *opr) {
    if (!p || !opr) return NGAP_ERR_ARG;
    const bool rho = ru || rv || rw;
    if (rho && !(ru && rv && rw)) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    if (rho)
        hipLaunchKernelGGL(k_laplacian<true>, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->dp[0],
                           (const double *)p->dp[1], (const double *)p->dp[2], (const double *)p->du[0],
                           (const double *)p->du[1], (const double *)p->du[2], (const double *)p->vol, ru, rv, rw, opr);
    else
        hipLaunchKernelGGL(k_laplacian<false>, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->dp[0],
                           (const double *)p->dp[1], (const double *)p->dp[2], (const double *)p->du[0],
                           (const double *)p->du[1], (const double *)p->du[2], (const double *)p->vol,
                           (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, opr);
    return fin(p);
}

int ngap_proj_div(ngap_proj *p, const double *u, const double *v, const double *w, double *div) {
    if (!p || !u || !v || !w || !div) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    hipLaunchKernelGGL(k_div, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->dp[0],
                       (const double *)p->dp[1], (const double *)p->dp[2], u, v, w, div);
    sync_all(p, div);
    return fin(p);
}

int ngap_proj_pgrad(ngap_proj *p, const double *pres, double *px, double *py, double *pz) {
    if (!p || !pres || !px || !py || !pz) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    double *out[3] = {px, py, pz};
    const long long st[3] = {1, (long long)b.nxo, (long long)b.nxo * b.nyo};
    for (int ax = 0; ax < 3; ax++) {
        hipLaunchKernelGGL(k_pgrad, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->du[ax], st[ax],
                           pres, out[ax]);
        sync_all(p, out[ax]);
    }
    return fin(p);
}

int ngap_proj_shift_p(ngap_proj *p, double *pres, double *mean) {
    if (!p || !pres) return NGAP_ERR_ARG;
    if (!(p->fvol > 0.0)) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    hipLaunchKernelGGL(k_integrate, dim3(p->nb), dim3(PT), 0, p->s, b, (const double *)p->vol, (const double *)p->vf,
                       (const double *)pres, p->part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(PT), 0, p->s, (const double *)p->part, p->nb, p->num);
    hipLaunchKernelGGL(k_shift, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->vf,
                       (const double *)p->num, p->fvol, pres);
    sync_all(p, pres);
    int rc = fin(p);
    if (rc) return rc;
    if (mean) {
        double num = 0.0;
        if (hipMemcpy(&num, p->num, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return NGAP_ERR_HIP;
        *mean = num / p->fvol;
    }
    return 0;
}

int ngap_proj_sync(ngap_proj *p, double *a) {
    if (!p || !a) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    sync_all(p, a);
    return fin(p);
}

}  // extern "C"
