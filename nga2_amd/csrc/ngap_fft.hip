// fourier3d on the GPU: the FFT direct solve of a circulant stencil operator
// on a periodic uniform grid (fourier3d_class.f90:1103-1202), with rocFFT
// (through hipFFT) for the transforms and HIP kernels for everything around
// them: the circulance check, the placement of the stencil, the inversion of
// the transformed operator, the spectral divide, the copy of the solution into
// sol and its ghost fill.
//
// The transforms are real-to-complex over the directions with more than one
// cell, slowest first (z, y, x). The forward one reads the interior of the
// caller's ghosted rhs in place (strided plan layout); the inverse one writes
// a dense array (rocFFT's two-dimensional complex-to-real plans with a
// strided output fail to execute), which one pass copies into the interior
// of sol. A solve is: forward transform, one pass over the half spectrum
// (zero mode, 1/F, normalisation), inverse transform, copy, ghost fill.
// The half spectrum (last transformed direction n/2+1 long) holds every mode
// of the reference's full complex transform, since F(-k) = conj F(k) for a
// real stencil and a real right-hand side.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "ngapfft.h"

namespace {

constexpr int PT = 256;
constexpr int NB = 1024;  // blocks of the reductions (partials summed on the host)

struct Grid {
    int nx, ny, nz;          // interior
    int nxo, nyo, nzo;       // ghosted
    int gx, gy, gz;          // ghost widths
    int px, py, pz;          // periodic
    long long N, No;         // interior and ghosted cell counts
};

int blocks_for(long long n) {
    long long v = (n + PT - 1) / PT;
    if (v > 65535LL * 16) v = 65535LL * 16;
    return (int)(v < 1 ? 1 : v);
}

__device__ __forceinline__ long long ghosted(const Grid &g, int i, int j, int k) {
    return ((long long)(k + g.gz) * g.nyo + (j + g.gy)) * g.nxo + (i + g.gx);
}

// count the interior cells whose stencil differs from the first interior
// cell's by more than tol in any entry (fourier3d_class.f90:1118-1126)
__global__ void __launch_bounds__(PT) k_circulant(Grid g, int nst, const double *__restrict__ opr, double tol,
                                                  int *__restrict__ part) {
    __shared__ int sm[PT];
    const double *ref = opr + (size_t)nst * ghosted(g, 0, 0, 0);
    int bad = 0;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < g.N; q += (long long)gridDim.x * PT) {
        const int i = (int)(q % g.nx);
        const long long r = q / g.nx;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        const double *o = opr + (size_t)nst * ghosted(g, i, j, k);
        for (int s = 0; s < nst; s++)
            if (fabs(o[s] - ref[s]) > tol) {
                bad = 1;
                break;
            }
    }
    sm[threadIdx.x] = bad;
    __syncthreads();
    for (int w = PT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(PT) k_zero(double *__restrict__ a, long long n) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < n; q += (long long)gridDim.x * PT) a[q] = 0.0;
}

// the first interior cell's stencil at its offsets modulo the grid
// (fourier3d_class.f90:1128-1137), into the interior of a ghosted-layout
// array; one thread, in stencil order
__global__ void k_place(Grid g, int nst, const int *__restrict__ stc, const double *__restrict__ opr,
                        double *__restrict__ kern) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double *ref = opr + (size_t)nst * ghosted(g, 0, 0, 0);
    for (int s = 0; s < nst; s++) {
        int i = stc[s] % g.nx, j = stc[s + nst] % g.ny, k = stc[s + 2 * nst] % g.nz;
        if (i < 0) i += g.nx;
        if (j < 0) j += g.ny;
        if (k < 0) k += g.nz;
        kern[ghosted(g, i, j, k)] += ref[s];
    }
}

// zero mode to 1, count |F| < 1000 eps, invert, count NaN
// (fourier3d_class.f90:1142-1160)
__global__ void __launch_bounds__(PT) k_invert(long long M, double2 *__restrict__ F, int *__restrict__ part_small,
                                               int *__restrict__ part_nan) {
    __shared__ int s0[PT], s1[PT];
    int small = 0, nan = 0;
    for (long long m = blockIdx.x * (long long)PT + threadIdx.x; m < M; m += (long long)gridDim.x * PT) {
        double2 f = F[m];
        if (m == 0) f = make_double2(1.0, 0.0);
        if (hypot(f.x, f.y) < 1000.0 * DBL_EPSILON) small++;
        const double d = f.x * f.x + f.y * f.y;
        const double2 r = make_double2(f.x / d, -f.y / d);
        if (isnan(hypot(r.x, r.y))) nan++;
        F[m] = r;
    }
    s0[threadIdx.x] = small;
    s1[threadIdx.x] = nan;
    __syncthreads();
    for (int w = PT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_small[blockIdx.x] = s0[0];
        part_nan[blockIdx.x] = s1[0];
    }
}

// G = G / F with the zero mode of G set to 0 (fourier3d_class.f90:1039-1040,
// 1186-1187) and the three 1/n of the inverse transforms folded in
__global__ void __launch_bounds__(PT) k_divide(long long M, double2 *__restrict__ G, const double2 *__restrict__ Finv,
                                               double scale) {
    for (long long m = blockIdx.x * (long long)PT + threadIdx.x; m < M; m += (long long)gridDim.x * PT) {
        const double2 a = m == 0 ? make_double2(0.0, 0.0) : G[m];
        const double2 b = Finv[m];
        G[m] = make_double2((a.x * b.x - a.y * b.y) * scale, (a.x * b.y + a.y * b.x) * scale);
    }
}

// interior of a ghosted array from a dense x-fastest one
__global__ void __launch_bounds__(PT) k_put(Grid g, const double *__restrict__ d, double *__restrict__ a) {
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < g.N; q += (long long)gridDim.x * PT) {
        const int i = (int)(q % g.nx);
        const long long r = q / g.nx;
        a[ghosted(g, i, (int)(r % g.ny), (int)(r / g.ny))] = d[q];
    }
}

// ghost fill of one direction, as cfg%sync does on one rank
// (pgrid_class.f90:459-516): the 2*w ghost layers of direction ax take the
// periodic image, over the whole ghosted extent of the other two directions
// (so the passes x, then y, then z also fill edges and corners). The cells
// read lie inside the box in direction ax and are not written by the pass.
__global__ void __launch_bounds__(PT) k_sync_dir(Grid g, int ax, double *a) {
    const int n = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
    const int w = ax == 0 ? g.gx : (ax == 1 ? g.gy : g.gz);
    const int e0 = ax == 0 ? g.nyo : g.nxo;  // the other two ghosted extents
    const int e1 = ax == 2 ? g.nyo : g.nzo;
    const long long total = (long long)e0 * e1 * 2 * w;
    for (long long q = blockIdx.x * (long long)PT + threadIdx.x; q < total; q += (long long)gridDim.x * PT) {
        const int l = (int)(q % (2 * w));
        const long long r = q / (2 * w);
        const int u = (int)(r % e0), v = (int)(r / e0);
        // ghost layer l: ghosted index 0..w-1 (below) or n+w..n+2w-1 (above),
        // box coordinate l-w or n+l-w; its image is the box coordinate modulo n
        const int gidx = l < w ? l : n + l;
        const int src = w + ((l - w) % n + n) % n;
        long long dst_off, src_off;
        if (ax == 0) {
            dst_off = ((long long)v * g.nyo + u) * g.nxo + gidx;
            src_off = ((long long)v * g.nyo + u) * g.nxo + src;
        } else if (ax == 1) {
            dst_off = ((long long)v * g.nyo + gidx) * g.nxo + u;
            src_off = ((long long)v * g.nyo + src) * g.nxo + u;
        } else {
            dst_off = ((long long)gidx * g.nyo + v) * g.nxo + u;
            src_off = ((long long)src * g.nyo + v) * g.nxo + u;
        }
        a[dst_off] = a[src_off];
    }
}

int sync_all(hipStream_t s, const Grid &g, double *a) {
    const int per[3] = {g.px, g.py, g.pz}, w[3] = {g.gx, g.gy, g.gz};
    const long long e[3] = {(long long)g.nyo * g.nzo, (long long)g.nxo * g.nzo, (long long)g.nxo * g.nyo};
    for (int ax = 0; ax < 3; ax++) {
        if (!per[ax] || w[ax] == 0) continue;
        hipLaunchKernelGGL(k_sync_dir, dim3(blocks_for(e[ax] * 2 * w[ax])), dim3(PT), 0, s, g, ax, a);
    }
    return 0;
}

}  // namespace

struct ngap_fft {
    Grid g{};
    int nst = 0;
    int device = 0;
    int rank = 0;       // directions with more than one cell
    long long M = 1;    // complex modes of the half spectrum
    hipStream_t s = nullptr;
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false, setup_done = false;
    int *d_stc = nullptr, *d_part = nullptr;
    double *d_kern = nullptr;  // scratch: the stencil in ghosted layout (setup), the dense solution (solve)
    long long off = 0;         // offset of the first interior cell in a ghosted array
    double2 *d_spec = nullptr, *d_finv = nullptr;
    double *d_io = nullptr;  // ghosted staging for the host entry points
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    std::vector<void *> allocs;
    ~ngap_fft() {
        if (have_fwd) (void)hipfftDestroy(fwd);
        if (have_inv) (void)hipfftDestroy(inv);
        for (void *q : allocs) (void)hipFree(q);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

template <typename T>
int dalloc(ngap_fft *F, T **p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess) return NGAP_ERR_ALLOC;
    F->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

// a failed runtime or transform call: name it on stderr, return NGAP_ERR_HIP
int fail(const char *where, int code) {
    std::fprintf(stderr, "[ngap_fft] %s failed (%d)\n", where, code);
    return NGAP_ERR_HIP;
}

int fin(ngap_fft *F) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(hipGetErrorString(e), (int)e);
    e = hipStreamSynchronize(F->s);
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e), (int)e);
}

int sum_parts(ngap_fft *F, const int *d, long long *out) {
    std::vector<int> h(NB);
    if (hipMemcpy(h.data(), d, sizeof(int) * NB, hipMemcpyDeviceToHost) != hipSuccess) return NGAP_ERR_HIP;
    long long t = 0;
    for (int v : h) t += v;
    *out = t;
    return 0;
}

int setup_from(ngap_fft *F, const double *d_opr, double tol) {
    const Grid &g = F->g;
    F->setup_done = false;
    hipLaunchKernelGGL(k_circulant, dim3(NB), dim3(PT), 0, F->s, g, F->nst, d_opr, tol, F->d_part);
    int rc = fin(F);
    long long bad = 0;
    if (!rc) rc = sum_parts(F, F->d_part, &bad);
    if (rc) return rc;
    if (bad) return NGAP_ERR_NOT_CIRCULANT;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(g.No)), dim3(PT), 0, F->s, F->d_kern, g.No);
    hipLaunchKernelGGL(k_place, dim3(1), dim3(64), 0, F->s, g, F->nst, (const int *)F->d_stc, d_opr, F->d_kern);
    if (F->rank > 0) {
        const hipfftResult r = hipfftExecD2Z(F->fwd, F->d_kern + F->off, (hipfftDoubleComplex *)F->d_finv);
        if (r != HIPFFT_SUCCESS) return fail("hipfftExecD2Z (setup)", (int)r);
    } else if (hipMemcpyAsync(F->d_finv, F->d_kern + F->off, sizeof(double), hipMemcpyDeviceToDevice, F->s) !=
               hipSuccess) {
        return NGAP_ERR_HIP;  // one cell: the transform is the value itself (imaginary part set below)
    }
    if (F->rank == 0) hipLaunchKernelGGL(k_zero, dim3(1), dim3(PT), 0, F->s, (double *)F->d_finv + 1, 1LL);
    hipLaunchKernelGGL(k_invert, dim3(NB), dim3(PT), 0, F->s, F->M, F->d_finv, F->d_part, F->d_part + NB);
    rc = fin(F);
    long long small = 0, nan = 0;
    if (!rc) rc = sum_parts(F, F->d_part, &small);
    if (!rc) rc = sum_parts(F, F->d_part + NB, &nan);
    if (rc) return rc;
    if (small || nan) return NGAP_ERR_SINGULAR;
    F->setup_done = true;
    return 0;
}

int solve_on(ngap_fft *F, const double *d_rhs, double *d_sol) {
    const Grid &g = F->g;
    if (hipEventRecord(F->e0, F->s) != hipSuccess) return fail("hipEventRecord", 0);
    if (F->rank > 0) {
        // the plans read the interior of rhs and write the interior of sol in place
        hipfftResult r = hipfftExecD2Z(F->fwd, const_cast<double *>(d_rhs) + F->off, (hipfftDoubleComplex *)F->d_spec);
        if (r != HIPFFT_SUCCESS) return fail("hipfftExecD2Z", (int)r);
        hipLaunchKernelGGL(k_divide, dim3(blocks_for(F->M)), dim3(PT), 0, F->s, F->M, F->d_spec,
                           (const double2 *)F->d_finv, 1.0 / ((double)g.nx * g.ny * g.nz));
        r = hipfftExecZ2D(F->inv, (hipfftDoubleComplex *)F->d_spec, F->d_kern);
        if (r != HIPFFT_SUCCESS) return fail("hipfftExecZ2D", (int)r);
        hipLaunchKernelGGL(k_put, dim3(blocks_for(g.N)), dim3(PT), 0, F->s, g, (const double *)F->d_kern, d_sol);
    } else {
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(PT), 0, F->s, d_sol + F->off, 1LL);  // the only mode is the zero mode
    }
    sync_all(F->s, g, d_sol);
    if (hipEventRecord(F->e1, F->s) != hipSuccess) return fail("hipEventRecord", 0);
    int rc = fin(F);
    if (rc) return rc;
    float ms = 0.0f;
    const hipError_t e = hipEventElapsedTime(&ms, F->e0, F->e1);
    if (e != hipSuccess) return fail("hipEventElapsedTime", (int)e);
    F->last_ms = ms;
    return 0;
}

}  // namespace

extern "C" {

int ngap_fft_create(ngap_fft **fp, const ngap_grid *gr, int nst, const int *stc) {
    if (!fp) return NGAP_ERR_ARG;
    *fp = nullptr;
    if (!gr || !stc || nst < 1) return NGAP_ERR_ARG;
    Grid g{};
    g.nx = gr->imax_ - gr->imin_ + 1;
    g.ny = gr->jmax_ - gr->jmin_ + 1;
    g.nz = gr->kmax_ - gr->kmin_ + 1;
    g.nxo = gr->imaxo_ - gr->imino_ + 1;
    g.nyo = gr->jmaxo_ - gr->jmino_ + 1;
    g.nzo = gr->kmaxo_ - gr->kmino_ + 1;
    g.gx = gr->imin_ - gr->imino_;
    g.gy = gr->jmin_ - gr->jmino_;
    g.gz = gr->kmin_ - gr->kmino_;
    g.px = gr->xper;
    g.py = gr->yper;
    g.pz = gr->zper;
    if (g.nx < 1 || g.ny < 1 || g.nz < 1 || g.gx < 0 || g.gy < 0 || g.gz < 0) return NGAP_ERR_ARG;
    if (g.nx != gr->nx || g.ny != gr->ny || g.nz != gr->nz) return NGAP_ERR_ARG;  // one GPU holds the whole grid
    if (g.nxo != g.nx + 2 * g.gx || g.nyo != g.ny + 2 * g.gy || g.nzo != g.nz + 2 * g.gz) return NGAP_ERR_ARG;
    // fourier3d_class.f90:146-148: periodic wherever there is more than one cell
    if ((g.nx > 1 && !g.px) || (g.ny > 1 && !g.py) || (g.nz > 1 && !g.pz)) return NGAP_ERR_UNSUPPORTED;
    g.N = (long long)g.nx * g.ny * g.nz;
    g.No = (long long)g.nxo * g.nyo * g.nzo;
    if (nst * g.No >= (1LL << 40)) return NGAP_ERR_ARG;
    if (g.No >= (1LL << 31)) return NGAP_ERR_UNSUPPORTED;  // the plans' strides and distances are int
    ngap_fft *F = new ngap_fft();
    F->g = g;
    F->nst = nst;
    // transformed directions, slowest first, with their strides in the ghosted layout
    int dims[3], r = 0;
    long long str[3];
    if (g.nz > 1) dims[r] = g.nz, str[r++] = (long long)g.nxo * g.nyo;
    if (g.ny > 1) dims[r] = g.ny, str[r++] = g.nxo;
    if (g.nx > 1) dims[r] = g.nx, str[r++] = 1;
    F->rank = r;
    F->off = ((long long)g.gz * g.nyo + g.gy) * g.nxo + g.gx;
    F->M = 1;
    for (int d = 0; d + 1 < r; d++) F->M *= dims[d];
    if (r > 0) F->M *= dims[r - 1] / 2 + 1;
    // real side: the ghosted array (embed[d] = stride ratio of consecutive
    // directions, the fastest stride as the element stride); complex side: the
    // dense half spectrum
    int rembed[3], cembed[3], dembed[3], rstride = r > 0 ? (int)str[r - 1] : 1;
    for (int d = 0; d < r; d++) {
        rembed[d] = d == 0 ? dims[0] : (int)(str[d - 1] / str[d]);
        cembed[d] = d == r - 1 ? dims[d] / 2 + 1 : dims[d];
        dembed[d] = dims[d];
    }
    int rc = hipGetDevice(&F->device) == hipSuccess ? 0 : NGAP_ERR_HIP;
    if (!rc) rc = hipStreamCreateWithFlags(&F->s, hipStreamNonBlocking) == hipSuccess ? 0 : NGAP_ERR_HIP;
    if (!rc) rc = hipEventCreate(&F->e0) == hipSuccess && hipEventCreate(&F->e1) == hipSuccess ? 0 : NGAP_ERR_HIP;
    if (!rc) rc = dalloc(F, &F->d_stc, 3 * (size_t)nst);
    if (!rc) rc = dalloc(F, &F->d_part, 2 * (size_t)NB);
    if (!rc) rc = dalloc(F, &F->d_kern, (size_t)g.No);
    if (!rc) rc = dalloc(F, &F->d_spec, (size_t)F->M);
    if (!rc) rc = dalloc(F, &F->d_finv, (size_t)F->M);
    if (!rc && hipMemcpy(F->d_stc, stc, sizeof(int) * 3 * nst, hipMemcpyHostToDevice) != hipSuccess) rc = NGAP_ERR_HIP;
    if (!rc && r > 0) {
        if (hipfftPlanMany(&F->fwd, r, dims, rembed, rstride, (int)g.No, cembed, 1, (int)F->M, HIPFFT_D2Z, 1) !=
            HIPFFT_SUCCESS)
            rc = NGAP_ERR_HIP;
        else
            F->have_fwd = true;
        if (!rc && hipfftPlanMany(&F->inv, r, dims, cembed, 1, (int)F->M, dembed, 1, (int)g.N, HIPFFT_Z2D, 1) !=
                       HIPFFT_SUCCESS)
            rc = NGAP_ERR_HIP;
        else if (!rc)
            F->have_inv = true;
        if (!rc && (hipfftSetStream(F->fwd, F->s) != HIPFFT_SUCCESS || hipfftSetStream(F->inv, F->s) != HIPFFT_SUCCESS))
            rc = NGAP_ERR_HIP;
    }
    if (rc) {
        delete F;
        return rc;
    }
    *fp = F;
    return 0;
}

int ngap_fft_setup_device(ngap_fft *f, const double *d_opr, double circ_tol) {
    if (!f || !d_opr || !(circ_tol >= 0.0)) return NGAP_ERR_ARG;
    if (hipSetDevice(f->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    return setup_from(f, d_opr, circ_tol);
}

int ngap_fft_setup(ngap_fft *f, const double *opr, double circ_tol) {
    if (!f || !opr || !(circ_tol >= 0.0)) return NGAP_ERR_ARG;
    if (hipSetDevice(f->device) != hipSuccess) return NGAP_ERR_HIP;
    const size_t n = (size_t)f->nst * f->g.No;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * n) != hipSuccess) return NGAP_ERR_ALLOC;
    int rc = hipMemcpy(d, opr, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess ? 0 : NGAP_ERR_HIP;
    if (!rc) rc = setup_from(f, d, circ_tol);
    (void)hipFree(d);
    return rc;
}

int ngap_fft_solve_device(ngap_fft *f, const double *d_rhs, double *d_sol) {
    if (!f || !d_rhs || !d_sol) return NGAP_ERR_ARG;
    if (!f->setup_done) return NGAP_ERR_NOT_SETUP;  // fourier3d_class.f90:1177
    if (hipSetDevice(f->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return NGAP_ERR_HIP;
    return solve_on(f, d_rhs, d_sol);
}

int ngap_fft_solve(ngap_fft *f, const double *rhs, double *sol) {
    if (!f || !rhs || !sol) return NGAP_ERR_ARG;
    if (!f->setup_done) return NGAP_ERR_NOT_SETUP;
    if (hipSetDevice(f->device) != hipSuccess) return NGAP_ERR_HIP;
    const size_t n = (size_t)f->g.No;
    if (!f->d_io) {
        int rc = dalloc(f, &f->d_io, 2 * n);
        if (rc) return rc;
    }
    double *dr = f->d_io, *ds = f->d_io + n;
    // sol goes up too: ghosts of a non-periodic (one-cell) direction keep the caller's values
    if (hipMemcpy(dr, rhs, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ds, sol, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail("hipMemcpy (upload)", 0);
    int rc = solve_on(f, dr, ds);
    if (!rc && hipMemcpy(sol, ds, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail("hipMemcpy (download)", 0);
    return rc;
}

int ngap_fft_destroy(ngap_fft *f) {
    if (!f) return NGAP_ERR_ARG;
    if (f->s) (void)hipStreamSynchronize(f->s);
    delete f;
    return 0;
}

int ngap_fft_last_ms(const ngap_fft *f, double *ms) {
    if (!f || !ms) return NGAP_ERR_ARG;
    *ms = f->last_ms;
    return 0;
}

}  // extern "C"
