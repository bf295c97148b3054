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
#include <mutex>
#include <vector>

#include "ngapproj.h"

#pragma clang fp contract(off)

namespace {

constexpr int PT = 256;
constexpr int MFR_PARTS = 64;  // partial sums per boundary condition (fixed order)

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


// rhs = -(((vol * div) * rho) / dt) over every cell: the driver's
// -cfg%vol*div*rho/dt (channel/src/simulation.f90:318); rho = 1 gives the
// two-phase form -vol*div/dt (falling_drop/src/simulation.f90:302) exactly
__global__ void __launch_bounds__(PT) k_rhs(long long n, const double *__restrict__ vol,
                                            const double *__restrict__ div, double rho, double dt,
                                            double *__restrict__ rhs) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < n; q += (long long)gridDim.x * PT)
        rhs[q] = -(((vol[q] * div[q]) * rho) / dt);
}

// U -= (dt * Px) / rho_face over every cell (channel/src/simulation.f90:325-
// 327 with the constant rho, falling_drop/src/simulation.f90:308-310 with
// rho_U); P += sol when both are given (:324, :307)
__global__ void __launch_bounds__(PT) k_correct(long long n, double dt, double rho, const double *__restrict__ ru,
                                                const double *__restrict__ rv, const double *__restrict__ rw,
                                                const double *__restrict__ px, const double *__restrict__ py,
                                                const double *__restrict__ pz, double *__restrict__ u,
                                                double *__restrict__ v, double *__restrict__ w,
                                                double *__restrict__ pres, const double *__restrict__ sol) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < n; q += (long long)gridDim.x * PT) {
        if (pres) pres[q] = pres[q] + sol[q];
        u[q] = u[q] - (dt * px[q]) / (ru ? ru[q] : rho);
        v[q] = v[q] - (dt * py[q]) / (rv ? rv[q] : rho);
        w[q] = w[q] - (dt * pz[q]) / (rw ? rw[q] : rho);
    }
}

// two-phase pressure gradient along one axis: divu(-1) P(c - e) + divu(0) P(c)
// - dPj(c) on interior cells, 0 on ghosts (tpns_class.f90:1695-1716)
__global__ void __launch_bounds__(PT) k_pgrad_jump(Box b, const double *__restrict__ du, long long stride,
                                                   const double *__restrict__ pres, const double *__restrict__ dpj,
                                                   double *__restrict__ g) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (!interior(b, i, j, k)) {
            g[q] = 0.0;
            continue;
        }
        g[q] = (du[2 * q] * pres[q - stride] + du[2 * q + 1] * pres[q]) - dpj[q];
    }
}

// CSF pressure jump on the faces of cells imin_..imax_+1 (and j, k alike),
// tpns_class.f90:1633-1663; dpj = new jump - old jump over every cell
// (:1676-1679, which leaves 0 where the jump is not recomputed)
__global__ void __launch_bounds__(PT) k_csf_pj(Box b, double sigma, const double *__restrict__ dux,
                                               const double *__restrict__ duy, const double *__restrict__ duz,
                                               const double *__restrict__ vol, const double *__restrict__ sd,
                                               const double *__restrict__ curv, const double *__restrict__ vf,
                                               double *__restrict__ pjx, double *__restrict__ pjy,
                                               double *__restrict__ pjz, double *__restrict__ dpjx,
                                               double *__restrict__ dpjy, double *__restrict__ dpjz) {
    const long long st[3] = {1, (long long)b.nxo, (long long)b.nxo * b.nyo};
    const double *du[3] = {dux, duy, duz};
    double *pj[3] = {pjx, pjy, pjz};
    double *dpj[3] = {dpjx, dpjy, dpjz};
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        const bool in = i >= b.no_x && i <= b.no_x + b.nx && j >= b.no_y && j <= b.no_y + b.ny && k >= b.no_z &&
                        k <= b.no_z + b.nz;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const double old = pj[ax][q];
            double nw = old;
            if (in) {
                const long long m = q - st[ax];
                const double surf = sd[m] * vol[m] + sd[q] * vol[q];
                double cv = 0.0;
                if (surf > 0.0) cv = ((sd[m] * curv[m]) * vol[m] + (sd[q] * curv[q]) * vol[q]) / surf;
                nw = (sigma * cv) * (du[ax][2 * q] * vf[m] + du[ax][2 * q + 1] * vf[q]);
                pj[ax][q] = nw;
            }
            dpj[ax][q] = nw - old;
        }
    }
}

// div += dt * (sum(divp_x*DPjx(i:i+1)/rho_U(i:i+1)) + ...) on interior cells
// (tpns_class.f90:1682-1688)
__global__ void __launch_bounds__(PT) k_csf_div(Box b, double dt, const double *__restrict__ dpx,
                                                const double *__restrict__ dpy, const double *__restrict__ dpz,
                                                const double *__restrict__ dpjx, const double *__restrict__ dpjy,
                                                const double *__restrict__ dpjz, const double *__restrict__ ru,
                                                const double *__restrict__ rv, const double *__restrict__ rw,
                                                double *__restrict__ div) {
    const long long sy = b.nxo, sz = (long long)b.nxo * b.nyo;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (!interior(b, i, j, k)) continue;
        const double tx = (dpx[2 * q] * dpjx[q]) / ru[q] + (dpx[2 * q + 1] * dpjx[q + 1]) / ru[q + 1];
        const double ty = (dpy[2 * q] * dpjy[q]) / rv[q] + (dpy[2 * q + 1] * dpjy[q + sy]) / rv[q + sy];
        const double tz = (dpz[2 * q] * dpjz[q]) / rw[q] + (dpz[2 * q + 1] * dpjz[q + sz]) / rw[q + sz];
        div[q] = div[q] + dt * (tx + ty + tz);
    }
}

// outgoing flow rate of one boundary condition: per-block partials of
// ((rdir * rho) * U) * d1 * d2 over its interior cells (incomp_class.f90:
// 1629-1646, tpns_class.f90:2175-2192 with rho = 1), in a fixed order
__global__ void __launch_bounds__(PT) k_mfr_part(int n, const long long *__restrict__ cell,
                                                 const double *__restrict__ d1, const double *__restrict__ d2,
                                                 double rdir_rho, const double *__restrict__ a,
                                                 double *__restrict__ part) {
    __shared__ double sm[PT / 64];
    double acc = 0.0;
    for (int t = blockIdx.x * PT + threadIdx.x; t < n; t += gridDim.x * PT) acc += ((rdir_rho * a[cell[t]]) * d1[t]) * d2[t];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < PT / 64; w++) s += sm[w];
        part[blockIdx.x] = s;
    }
}

// mfr_error = sum of the flow rates in boundary order; the velocity
// correction -mfr_error / (rho * correctable area), or 0 below the
// reference's thresholds (incomp_class.f90:1678-1681)
__global__ void k_mfr_fin(int nbc, const double *__restrict__ mfr, double rho, double carea, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double e = 0.0;
    for (int c = 0; c < nbc; c++) e += mfr[c];
    const double tiny = 10.0 * 2.220446049250313e-16;
    out[0] = e;
    out[1] = (fabs(e) < tiny || fabs(carea) < tiny) ? 0.0 : -e / (rho * carea);
    out[2] = (fabs(e) < tiny || fabs(carea) < tiny) ? 0.0 : 1.0;
}

// U(cell) += rdir * vel_correction on every cell of a correctable boundary
// condition, overlap included (incomp_class.f90:1692-1708)
__global__ void __launch_bounds__(PT) k_mfr_apply(int n, const long long *__restrict__ cell, double rdir,
                                                  const double *__restrict__ corr, double *__restrict__ a) {
    if (corr[2] == 0.0) return;
    const double vc = corr[1];
    for (int t = blockIdx.x * PT + threadIdx.x; t < n; t += gridDim.x * PT) a[cell[t]] = a[cell[t]] + rdir * vc;
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
    double *mfr_part = nullptr;   // flow-rate partials (MFR_PARTS per boundary condition)
    int imino = 0, jmino = 0, kmino = 0;  // Fortran lower bounds of the ghosted arrays
    int nb = 0;
    double fvol = 0.0;
    // boundary conditions for correct_mfr (ngap_proj_set_bconds)
    struct Bc {
        int face, can_correct, n_in, n_all;
        double rdir;
        long long *cell;    // ghosted flat index of each cell (interior cells first)
        double *d1, *d2;    // the two face lengths of each interior cell, in the reference's product order
    };
    std::vector<Bc> bcs;
    double carea = 0.0;     // correctable area (get_mfr)
    double *mfr = nullptr;  // per boundary flow rate, then [error, correction, apply] scalars
    std::vector<void *> bc_allocs;
    std::vector<void *> allocs;
    ~ngap_proj() {
        for (void *q : bc_allocs) (void)hipFree(q);
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

// extents of a ghosted grid on one rank; NGAP_ERR_ARG for an inconsistent
// grid or one of 2^40 / 7 cells or more
int box_of(const ngap_grid *g, Box &b) {
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
    b.No = (long long)b.nxo * b.nyo * b.nzo;
    if (7 * b.No >= (1LL << 40)) return NGAP_ERR_ARG;
    return 0;
}

int sync_box(const Box &b, hipStream_t st, double *a) {
    const int per[3] = {b.px, b.py, b.pz};
    const int no[3] = {b.no_x, b.no_y, b.no_z};
    for (int ax = 0; ax < 3; ax++) {
        if (!per[ax] || no[ax] == 0) continue;
        const long long e0 = ax == 0 ? b.nyo : b.nxo;
        const long long e1 = ax == 2 ? b.nyo : b.nzo;
        hipLaunchKernelGGL(k_sync, dim3(blocks_for(e0 * e1 * 2 * no[ax])), dim3(PT), 0, st, b, ax, a);
    }
    return 0;
}

int sync_all(ngap_proj *P, double *a) { return sync_box(P->b, P->s, a); }

}  // namespace

extern "C" {

int ngap_proj_create(ngap_proj **pp, const ngap_grid *g, const double *divp_x, const double *divp_y,
                     const double *divp_z, const double *divu_x, const double *divu_y, const double *divu_z,
                     const double *vol, const double *vf) {
    if (!pp) return NGAP_ERR_ARG;
    *pp = nullptr;
    if (!g || !divp_x || !divp_y || !divp_z || !divu_x || !divu_y || !divu_z || !vol || !vf) return NGAP_ERR_ARG;
    Box b{};
    const int rc0 = box_of(g, b);
    if (rc0) return rc0;
    // the stencils read one cell past the interior in each direction
    if (b.no_x < 1 || b.no_y < 1 || b.no_z < 1) return NGAP_ERR_ARG;
    ngap_proj *P = new ngap_proj();
    P->b = b;
    P->imino = g->imino_;
    P->jmino = g->jmino_;
    P->kmino = g->kmino_;
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


int ngap_proj_rhs(ngap_proj *p, const double *div, double rho, double dt, double *rhs) {
    if (!p || !div || !rhs) return NGAP_ERR_ARG;
    if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(rho)) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    hipLaunchKernelGGL(k_rhs, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b.No, (const double *)p->vol, div, rho, dt,
                       rhs);
    return fin(p);
}

int ngap_proj_correct(ngap_proj *p, double dt, double rho, const double *rho_u, const double *rho_v,
                      const double *rho_w, const double *px, const double *py, const double *pz, double *u, double *v,
                      double *w, double *pres, const double *sol) {
    if (!p || !px || !py || !pz || !u || !v || !w) return NGAP_ERR_ARG;
    const bool faces = rho_u || rho_v || rho_w;
    if (faces && !(rho_u && rho_v && rho_w)) return NGAP_ERR_ARG;
    if (!faces && !(rho != 0.0)) return NGAP_ERR_ARG;
    if ((pres == nullptr) != (sol == nullptr)) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    hipLaunchKernelGGL(k_correct, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b.No, dt, rho, rho_u, rho_v, rho_w, px, py,
                       pz, u, v, w, pres, sol);
    return fin(p);
}

int ngap_proj_pgrad_jump(ngap_proj *p, const double *pres, const double *dpjx, const double *dpjy,
                         const double *dpjz, double *px, double *py, double *pz) {
    if (!p || !pres || !dpjx || !dpjy || !dpjz || !px || !py || !pz) return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    double *out[3] = {px, py, pz};
    const double *dj[3] = {dpjx, dpjy, dpjz};
    const long long st[3] = {1, (long long)b.nxo, (long long)b.nxo * b.nyo};
    for (int ax = 0; ax < 3; ax++) {
        hipLaunchKernelGGL(k_pgrad_jump, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, (const double *)p->du[ax],
                           st[ax], pres, dj[ax], out[ax]);
        sync_all(p, out[ax]);
    }
    return fin(p);
}

int ngap_proj_csf_jump(ngap_proj *p, double sigma, double dt, const double *sd, const double *curv, const double *vf,
                       const double *rho_u, const double *rho_v, const double *rho_w, double *pjx, double *pjy,
                       double *pjz, double *dpjx, double *dpjy, double *dpjz, double *div) {
    if (!p || !sd || !curv || !vf || !rho_u || !rho_v || !rho_w || !pjx || !pjy || !pjz || !dpjx || !dpjy || !dpjz ||
        !div)
        return NGAP_ERR_ARG;
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    const Box &b = p->b;
    hipLaunchKernelGGL(k_csf_pj, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, sigma, (const double *)p->du[0],
                       (const double *)p->du[1], (const double *)p->du[2], (const double *)p->vol, sd, curv, vf, pjx,
                       pjy, pjz, dpjx, dpjy, dpjz);
    hipLaunchKernelGGL(k_csf_div, dim3(blocks_for(b.No)), dim3(PT), 0, p->s, b, dt, (const double *)p->dp[0],
                       (const double *)p->dp[1], (const double *)p->dp[2], (const double *)dpjx, (const double *)dpjy,
                       (const double *)dpjz, rho_u, rho_v, rho_w, div);
    return fin(p);
}

int ngap_proj_set_bconds(ngap_proj *p, int nbc, const ngap_bcond *bc, const double *dx, const double *dy,
                         const double *dz) {
    if (!p || nbc < 0 || (nbc > 0 && (!bc || !dx || !dy || !dz))) return NGAP_ERR_ARG;
    const Box &b = p->b;
    // validate everything before replacing the current set
    for (int c = 0; c < nbc; c++) {
        const ngap_bcond &B = bc[c];
        if (B.face < 0 || B.face > 2 || (B.rdir != 1 && B.rdir != -1) || B.n_in < 0 || B.n_all < B.n_in ||
            (B.n_all > 0 && !B.ijk))
            return NGAP_ERR_ARG;
        for (int t = 0; t < B.n_all; t++) {
            const int i = B.ijk[3 * t] - p->imino, j = B.ijk[3 * t + 1] - p->jmino, k = B.ijk[3 * t + 2] - p->kmino;
            if (i < 0 || i >= b.nxo || j < 0 || j >= b.nyo || k < 0 || k >= b.nzo) return NGAP_ERR_ARG;
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    for (void *q : p->bc_allocs) (void)hipFree(q);
    p->bc_allocs.clear();
    p->bcs.clear();
    p->mfr = nullptr;
    p->carea = 0.0;
    auto dup = [&](void **d, const void *h, size_t bytes) -> int {
        *d = nullptr;
        if (hipMalloc(d, bytes ? bytes : 8) != hipSuccess) return NGAP_ERR_ALLOC;
        p->bc_allocs.push_back(*d);
        if (h && bytes && hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return NGAP_ERR_HIP;
        return 0;
    };
    double carea = 0.0;
    for (int c = 0; c < nbc; c++) {
        const ngap_bcond &B = bc[c];
        std::vector<long long> cell(B.n_all);
        std::vector<double> d1(B.n_in), d2(B.n_in);
        double area = 0.0;
        for (int t = 0; t < B.n_all; t++) {
            const int fi = B.ijk[3 * t], fj = B.ijk[3 * t + 1], fk = B.ijk[3 * t + 2];
            cell[t] = (fi - p->imino) + (long long)b.nxo * ((fj - p->jmino) + (long long)b.nyo * (fk - p->kmino));
            if (t >= B.n_in) continue;
            const double ddx = dx[fi - p->imino], ddy = dy[fj - p->jmino], ddz = dz[fk - p->kmino];
            // the reference's product order: x face dy*dz, y face dz*dx, z face dx*dy (incomp_class.f90:1631-1645)
            d1[t] = B.face == 0 ? ddy : (B.face == 1 ? ddz : ddx);
            d2[t] = B.face == 0 ? ddz : (B.face == 1 ? ddx : ddy);
            area = area + d1[t] * d2[t];
        }
        ngap_proj::Bc E{};
        E.face = B.face;
        E.can_correct = B.can_correct ? 1 : 0;
        E.n_in = B.n_in;
        E.n_all = B.n_all;
        E.rdir = (double)B.rdir;
        int rc = dup((void **)&E.cell, cell.data(), sizeof(long long) * cell.size());
        if (!rc) rc = dup((void **)&E.d1, d1.data(), sizeof(double) * d1.size());
        if (!rc) rc = dup((void **)&E.d2, d2.data(), sizeof(double) * d2.size());
        if (rc) return rc;
        p->bcs.push_back(E);
        // correctable_area = sum(area * canCorrect) (incomp_class.f90:1660)
        carea = carea + area * (B.can_correct ? 1.0 : 0.0);
    }
    p->carea = carea;
    int rc = dup((void **)&p->mfr, nullptr, sizeof(double) * (nbc + 3));
    if (rc) return rc;
    if (p->nb > 0 && (size_t)nbc * MFR_PARTS > (size_t)p->nb) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(double) * nbc * MFR_PARTS) != hipSuccess) return NGAP_ERR_ALLOC;
        p->bc_allocs.push_back(q);
        p->mfr_part = (double *)q;
    } else {
        p->mfr_part = p->part;
    }
    return 0;
}

int ngap_proj_correct_mfr(ngap_proj *p, double rho, double *u, double *v, double *w, double *mfr_error) {
    if (!p || !u || !v || !w || !(rho != 0.0)) return NGAP_ERR_ARG;
    if (p->bcs.empty()) {  // no boundary condition: nothing flows in or out (mfr_error = sum of nothing = 0)
        if (mfr_error) *mfr_error = 0.0;
        return 0;
    }
    if (hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    double *fld[3] = {u, v, w};
    const int nbc = (int)p->bcs.size();
    for (int c = 0; c < nbc; c++) {
        const ngap_proj::Bc &B = p->bcs[c];
        double *part = p->mfr_part + (size_t)c * MFR_PARTS;
        hipLaunchKernelGGL(k_mfr_part, dim3(MFR_PARTS), dim3(PT), 0, p->s, B.n_in, (const long long *)B.cell,
                           (const double *)B.d1, (const double *)B.d2, B.rdir * rho, (const double *)fld[B.face],
                           part);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(PT), 0, p->s, (const double *)part, (int)MFR_PARTS, p->mfr + c);
    }
    hipLaunchKernelGGL(k_mfr_fin, dim3(1), dim3(64), 0, p->s, nbc, (const double *)p->mfr, rho, p->carea,
                       p->mfr + nbc);
    for (int c = 0; c < nbc; c++) {
        const ngap_proj::Bc &B = p->bcs[c];
        if (!B.can_correct || B.n_all == 0) continue;
        hipLaunchKernelGGL(k_mfr_apply, dim3(blocks_for(B.n_all)), dim3(PT), 0, p->s, B.n_all,
                           (const long long *)B.cell, B.rdir, (const double *)(p->mfr + nbc), fld[B.face]);
    }
    int rc = fin(p);
    if (rc) return rc;
    if (mfr_error && hipMemcpy(mfr_error, p->mfr + nbc, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return NGAP_ERR_HIP;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- mean removal without allocation
// ngap_proj_shift_p_noalloc keeps its partial sums in a file-scope device
// array instead of memory of an ngap_proj handle: 1024 partials of the fluid
// volume, 1024 of the weighted sum of P, then the two totals (16.4 KB). The
// partial count equals that of ngap_proj_create / ngap_proj_shift_p, so the
// sums, and the result, are the same bits. Every slot is written before it
// is read within one call, so nothing needs clearing between calls; calls
// that overlap on one device share the array and race.
namespace {
constexpr int NOALLOC_PARTS = 1024;
constexpr int SCRATCH_DOUBLES = 2 * NOALLOC_PARTS + 2;
}  // namespace

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[MAX_DEVICES] = {};

// device address of `scratch` on the current device, looked up once per device
// (each device holds its own copy of the array at its own address)
int scratch_ptr(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return NGAP_ERR_HIP;
    if (dev < 0 || dev >= MAX_DEVICES) return NGAP_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return NGAP_ERR_HIP;
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return 0;
}

// P -= tot[1] / tot[0] on interior cells with VF > 0, tot[0] the fluid
// volume; mean (optional) = that value. A fluid volume that is not positive
// leaves P as it is and the mean NaN (ngap_proj_create refuses such a grid).
__global__ void __launch_bounds__(PT) k_shift_dev(Box b, const double *__restrict__ vf,
                                                  const double *__restrict__ tot, double *__restrict__ pres,
                                                  double *__restrict__ mean) {
    const double fvol = tot[0];
    const bool ok = fvol > 0.0;
    const double m = ok ? tot[1] / fvol : NAN;
    if (mean && blockIdx.x == 0 && threadIdx.x == 0) mean[0] = m;
    if (!ok) return;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < b.No; q += (long long)gridDim.x * PT) {
        int i, j, k;
        split(b, q, i, j, k);
        if (interior(b, i, j, k) && vf[q] > 0.0) pres[q] = pres[q] - m;
    }
}

}  // namespace

extern "C" int ngap_proj_shift_p_noalloc(const ngap_grid *g, const double *vol, const double *vf, double *pres,
                                         double *d_mean, void *stream) {
    if (!g || !vol || !vf || !pres) return NGAP_ERR_ARG;
    Box b{};
    int rc = box_of(g, b);
    if (rc) return rc;
    double *sp = nullptr;
    if ((rc = scratch_ptr(&sp))) return rc;
    hipStream_t st = (hipStream_t)stream;
    double *tot = sp + 2 * NOALLOC_PARTS;
    hipLaunchKernelGGL(k_integrate, dim3(NOALLOC_PARTS), dim3(PT), 0, st, b, vol, vf, (const double *)nullptr, sp);
    hipLaunchKernelGGL(k_integrate, dim3(NOALLOC_PARTS), dim3(PT), 0, st, b, vol, vf, (const double *)pres,
                       sp + NOALLOC_PARTS);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(PT), 0, st, (const double *)sp, NOALLOC_PARTS, tot);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(PT), 0, st, (const double *)(sp + NOALLOC_PARTS), NOALLOC_PARTS, tot + 1);
    hipLaunchKernelGGL(k_shift_dev, dim3(blocks_for(b.No)), dim3(PT), 0, st, b, vf, (const double *)tot, pres, d_mean);
    sync_box(b, st, pres);
    if (hipGetLastError() != hipSuccess) return NGAP_ERR_HIP;
    if (!stream && hipStreamSynchronize(nullptr) != hipSuccess) return NGAP_ERR_HIP;
    return 0;
}
