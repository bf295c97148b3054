// Tiled level kernels (gfx950): SpMV, fused red/black Gauss-Seidel sweep,
// residual + restriction, prolongation + correction.
//
// Geometry: a level is n0 x n1 x n2 cells, p = i + n0*(j + n1*k) (int32 is
// enough for every level this solver allocates: N < 2^31 is checked at setup).
// Blocks tile the x-y plane; the SpMV and the sweep march through a chunk of
// z planes so each array element is read once and z neighbours come from
// registers.
#include <type_traits>

#include "ngap_internal.h"

namespace ngap {

namespace {

// coordinate wrapped into [0,n) for periodic directions; -1 outside a
// non-periodic box
__device__ __forceinline__ int wrapc(int i, int n, int per) {
    if (i >= 0 && i < n) return i;
    if (!per) return -1;
    i %= n;
    return i < 0 ? i + n : i;
}

// XCD-aware tile order. The dispatcher deals workgroups round-robin to the 8
// XCDs (workgroup b runs on XCD b % 8), each with its own L2. Remap the
// linear workgroup id so every XCD walks one contiguous range of tiles: tiles
// adjacent in y and z then share an L2, and their halo rows are hits.
// Bijective on [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, m = b >> 3;
    return x < r ? x * (q + 1) + m : r * (q + 1) + (x - r) * q + m;
}

// tile coordinates of this workgroup for a 1-D launch of gx*gy*gz groups
__device__ __forceinline__ void tile_of(int gx, int gy, int &bx, int &by, int &bz) {
    const int nb = gridDim.x;
    const int t = xcd_tile(blockIdx.x, nb);
    bx = t % gx;
    const int r = t / gx;
    by = r % gy;
    bz = r / gy;
}

}  // namespace

// ------------------------------------------------------------------ SpMV
// y = A x with the plane march; part (optional) gets one FP64 partial of
// <x, y> per block. Block (TX, TY) tiles x-y; blockIdx.z selects a z chunk.
// TX: the type x is stored in (FP32 search directions of the mixed solve
// with the FP64 operator); the product is formed in T.
template <typename T, typename TX = T>
__global__ void __launch_bounds__(256) k_apply_t(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                 const T *__restrict__ a1, const T *__restrict__ a2,
                                                 const TX *__restrict__ x, T *__restrict__ y,
                                                 double *__restrict__ part, int kc, int gx, int gy) {
    __shared__ double sm[4];
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    double acc = 0.0;
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);
    if (i < n0 && j < n1 && k0 < k1) {
        const int im = nbm(i, n0, g.p0), ip = nbp(i, n0, g.p0);
        const int jm = nbm(j, n1, g.p1), jp = nbp(j, n1, g.p1);
        const int row = j * n0;
        const int om = im - i, op = ip - i;            // x offsets
        const int ym = (jm - j) * n0, yp = (jp - j) * n0;  // y offsets
        int km = nbm(k0, n2, g.p2);
        const int col = row + i;
        T xm = km >= 0 ? T(x[km * plane + col]) : T(0);
        T xc = T(x[k0 * plane + col]);
        T a2c = a2[k0 * plane + col];
        for (int k = k0; k < k1; k++) {
            const int p = k * plane + col;
            const int kp = nbp(k, n2, g.p2);
            const T xp = kp >= 0 ? T(x[kp * plane + col]) : T(0);
            const T a2p = kp >= 0 ? a2[kp * plane + col] : T(0);
            T s = c[p] * xc;
            if (im >= 0) s += a0[p] * T(x[p + om]);
            if (ip >= 0) s += a0[p + op] * T(x[p + op]);
            if (jm >= 0) s += a1[p] * T(x[p + ym]);
            if (jp >= 0) s += a1[p + yp] * T(x[p + yp]);
            if (km >= 0) s += a2c * xm;
            if (kp >= 0) s += a2p * xp;
            y[p] = s;
            acc += (double)xc * (double)s;
            xm = xc;
            xc = xp;
            a2c = a2p;
            km = k;
        }
    }
    if (part) {
        const int t = threadIdx.y * blockDim.x + threadIdx.x;
        const int nt = blockDim.x * blockDim.y;
        double v = wave_sum(acc);
        if ((t & 63) == 0) sm[t >> 6] = v;
        __syncthreads();
        if (t == 0) {
            double s = 0.0;
            for (int q = 0; q < (nt + 63) / 64; q++) s += sm[q];
            part[blockIdx.x] = s;
        }
    }
}

// ------------------------------------------------------------------ SpMV, row waves
// Row-wave layout for levels whose rows are long enough: a wave owns one
// x-row segment of 64*CPL cells, CPL consecutive cells per lane, so every
// array is read with one 16-byte load per lane (1 KB per wave instruction).
// The x neighbours come from the neighbouring lanes (lane shuffles) and, at
// the two ends of the segment, from memory. Four waves (rows) per workgroup,
// a z chunk of kc planes per workgroup, y neighbours read directly (they are
// L2 hits: adjacent rows belong to the same or the XCD-neighbouring group).
template <typename T, int CPL>
struct alignas(16) Vec {
    T v[CPL];
};

template <typename T, int CPL>
__device__ __forceinline__ Vec<T, CPL> vld(const T *__restrict__ p) {
    return *reinterpret_cast<const Vec<T, CPL> *>(p);
}

template <typename T, int CPL>
__device__ __forceinline__ Vec<T, CPL> vzero() {
    Vec<T, CPL> r;
#pragma unroll
    for (int e = 0; e < CPL; e++) r.v[e] = T(0);
    return r;
}

// The three stencils that share this layout:
//   RW_APPLY    out = A x, with one FP64 partial of <x, A x> per workgroup
//   RW_RESID    out = f - A x
//   RW_SMOOTH0  first sweep from a zero iterate: with u = f / c, first colour
//               out = u, second colour out = (f - sum_nb a_nb u_nb) / c
// Sums are accumulated in the same order as the tiled kernels.
enum { RW_APPLY = 0, RW_RESID = 1, RW_SMOOTH0 = 2 };

// TX: the type x is stored in (RW_APPLY of the mixed solve: FP32 search
// directions, FP64 operator and product).
template <typename T, int CPL, int MODE, typename TX = T>
__global__ void __launch_bounds__(256) k_stencil_rw(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                    const T *__restrict__ a1, const T *__restrict__ a2,
                                                    const T *__restrict__ f, const TX *__restrict__ x,
                                                    T *__restrict__ y, double *__restrict__ part, int c1, int kc,
                                                    int gx, int gy) {
    __shared__ double sm[4];
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    const int i0 = (bx * 64 + lane) * CPL;
    const int j = by * 4 + w;
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);
    double acc = 0.0;
    typedef Vec<T, CPL> V;
    // the operand at a row/plane offset: x, or u = f / c for the zero-iterate sweep
    auto opnd = [&](int q) -> V {
        if (MODE == RW_SMOOTH0) {
            const V cq = vld<T, CPL>(c + q), fq = vld<T, CPL>(f + q);
            V u;
#pragma unroll
            for (int e = 0; e < CPL; e++) u.v[e] = fq.v[e] / cq.v[e];
            return u;
        } else if constexpr (std::is_same<TX, T>::value) {
            return vld<T, CPL>(x + q);
        } else {
            // CPL narrower values in one load of CPL * sizeof(TX) bytes
            typedef TX xv_t __attribute__((ext_vector_type(CPL)));
            const xv_t t = *reinterpret_cast<const xv_t *>(x + q);
            V r;
#pragma unroll
            for (int e = 0; e < CPL; e++) r.v[e] = T(t[e]);
            return r;
        }
    };
    auto opnd1 = [&](int q) -> T { return MODE == RW_SMOOTH0 ? f[q] / c[q] : T(x[q]); };
    if (i0 < n0 && j < n1 && k0 < k1) {
        const int jm = nbm(j, n1, g.p1), jp = nbp(j, n1, g.p1);
        const int il = nbm(i0, n0, g.p0), ir = nbp(i0 + CPL - 1, n0, g.p0);
        const bool lsh = lane > 0;                    // cell i0-1 is held by lane-1
        const bool rsh = lane < 63 && i0 + CPL < n0;  // cell i0+CPL is held by lane+1
        const bool hasl = lsh || il >= 0, hasr = rsh || ir >= 0;
        const int col = j * n0 + i0;
        const int row0 = j * n0;
        const int dym = (jm - j) * n0, dyp = (jp - j) * n0;
        int km = nbm(k0, n2, g.p2);
        V xm = km >= 0 ? opnd(km * plane + col) : vzero<T, CPL>();
        V xc = opnd(k0 * plane + col);
        V a2c = vld<T, CPL>(a2 + k0 * plane + col);
        for (int k = k0; k < k1; k++) {
            const int p = k * plane + col;
            const int kp = nbp(k, n2, g.p2);
            const V xp = kp >= 0 ? opnd(kp * plane + col) : vzero<T, CPL>();
            const V a2p = kp >= 0 ? vld<T, CPL>(a2 + kp * plane + col) : vzero<T, CPL>();
            const V cc = vld<T, CPL>(c + p), a0c = vld<T, CPL>(a0 + p), a1c = vld<T, CPL>(a1 + p);
            const V xjm = jm >= 0 ? opnd(p + dym) : vzero<T, CPL>();
            const V xjp = jp >= 0 ? opnd(p + dyp) : vzero<T, CPL>();
            const V a1p = jp >= 0 ? vld<T, CPL>(a1 + p + dyp) : vzero<T, CPL>();
            V fc;
            if (MODE != RW_APPLY) fc = vld<T, CPL>(f + p);
            const T xl_s = __shfl(xc.v[CPL - 1], (lane + 63) & 63, 64);
            const T xr_s = __shfl(xc.v[0], (lane + 1) & 63, 64);
            const T ar_s = __shfl(a0c.v[0], (lane + 1) & 63, 64);
            const int rk = k * plane + row0;
            const T xl = lsh ? xl_s : (il >= 0 ? opnd1(rk + il) : T(0));
            const T xr = rsh ? xr_s : (ir >= 0 ? opnd1(rk + ir) : T(0));
            const T ar = rsh ? ar_s : (ir >= 0 ? a0[rk + ir] : T(0));
            V out;
#pragma unroll
            for (int e = 0; e < CPL; e++) {
                const T left = e > 0 ? xc.v[e - 1] : xl;
                const T right = e < CPL - 1 ? xc.v[e + 1] : xr;
                const T aright = e < CPL - 1 ? a0c.v[e + 1] : ar;
                if (MODE == RW_SMOOTH0) {
                    if (((i0 + e + j + k) & 1) == c1) {
                        out.v[e] = xc.v[e];
                    } else {
                        T sum = fc.v[e];
                        if (e > 0 || hasl) sum -= a0c.v[e] * left;
                        if (e < CPL - 1 || hasr) sum -= aright * right;
                        if (jm >= 0) sum -= a1c.v[e] * xjm.v[e];
                        if (jp >= 0) sum -= a1p.v[e] * xjp.v[e];
                        if (km >= 0) sum -= a2c.v[e] * xm.v[e];
                        if (kp >= 0) sum -= a2p.v[e] * xp.v[e];
                        out.v[e] = sum / cc.v[e];
                    }
                } else {
                    T t = cc.v[e] * xc.v[e];
                    if (e > 0 || hasl) t += a0c.v[e] * left;
                    if (e < CPL - 1 || hasr) t += aright * right;
                    if (jm >= 0) t += a1c.v[e] * xjm.v[e];
                    if (jp >= 0) t += a1p.v[e] * xjp.v[e];
                    if (km >= 0) t += a2c.v[e] * xm.v[e];
                    if (kp >= 0) t += a2p.v[e] * xp.v[e];
                    if (MODE == RW_APPLY) {
                        out.v[e] = t;
                        acc += (double)xc.v[e] * (double)t;
                    } else {
                        out.v[e] = fc.v[e] - t;
                    }
                }
            }
            *reinterpret_cast<V *>(y + p) = out;
            xm = xc;
            xc = xp;
            a2c = a2p;
            km = k;
        }
    }
    if (MODE == RW_APPLY && part) {
        double v = wave_sum(acc);
        if (lane == 0) sm[w] = v;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    }
}

// ------------------------------------------------------------------ stencils from zero
// First symmetric sweep from a zero iterate, as a plain stencil:
//   first colour c1:  v = f / c
//   second colour:    v = (f - sum_nb a_nb f_nb / c_nb) / c
// (the second colour reads the first colour's new values, which are f/c).
template <typename T>
__global__ void __launch_bounds__(256) k_smooth0_t(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                   const T *__restrict__ a1, const T *__restrict__ a2,
                                                   const T *__restrict__ f, T *__restrict__ v, int c1, int kc, int gx, int gy) {
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);
    if (i >= n0 || j >= n1 || k0 >= k1) return;
    const int im = nbm(i, n0, g.p0), ip = nbp(i, n0, g.p0);
    const int jm = nbm(j, n1, g.p1), jp = nbp(j, n1, g.p1);
    const int om = im - i, op = ip - i;
    const int ym = (jm - j) * n0, yp = (jp - j) * n0;
    const int col = j * n0 + i;
    int km = nbm(k0, n2, g.p2);
    // f/c of the own column at k-1, k, k+1 and a2 at k
    T um = km >= 0 ? f[km * plane + col] / c[km * plane + col] : T(0);
    T cc = c[k0 * plane + col];
    T fc = f[k0 * plane + col];
    T a2c = a2[k0 * plane + col];
    for (int k = k0; k < k1; k++) {
        const int p = k * plane + col;
        const int kp = nbp(k, n2, g.p2);
        T cp = T(1), fp = T(0), a2p = T(0);
        if (kp >= 0) {
            cp = c[kp * plane + col];
            fp = f[kp * plane + col];
            a2p = a2[kp * plane + col];
        }
        const T up = kp >= 0 ? fp / cp : T(0);
        T out;
        if (((i + j + k) & 1) == c1) {
            out = fc / cc;
        } else {
            T sum = fc;
            if (im >= 0) sum -= a0[p] * (f[p + om] / c[p + om]);
            if (ip >= 0) sum -= a0[p + op] * (f[p + op] / c[p + op]);
            if (jm >= 0) sum -= a1[p] * (f[p + ym] / c[p + ym]);
            if (jp >= 0) sum -= a1[p + yp] * (f[p + yp] / c[p + yp]);
            if (km >= 0) sum -= a2c * um;
            if (kp >= 0) sum -= a2p * up;
            out = sum / cc;
        }
        v[p] = out;
        um = fc / cc;
        cc = cp;
        fc = fp;
        a2c = a2p;
        km = k;
    }
}

// r = f - A v (the residual handed to the restriction)
template <typename T>
__global__ void __launch_bounds__(256) k_resid_t(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                 const T *__restrict__ a1, const T *__restrict__ a2,
                                                 const T *__restrict__ f, const T *__restrict__ x,
                                                 T *__restrict__ r, int kc, int gx, int gy) {
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x;
    const int j = by * blockDim.y + threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);
    if (i >= n0 || j >= n1 || k0 >= k1) return;
    const int im = nbm(i, n0, g.p0), ip = nbp(i, n0, g.p0);
    const int jm = nbm(j, n1, g.p1), jp = nbp(j, n1, g.p1);
    const int om = im - i, op = ip - i;
    const int ym = (jm - j) * n0, yp = (jp - j) * n0;
    const int col = j * n0 + i;
    int km = nbm(k0, n2, g.p2);
    T xm = km >= 0 ? x[km * plane + col] : T(0);
    T xc = x[k0 * plane + col];
    T a2c = a2[k0 * plane + col];
    for (int k = k0; k < k1; k++) {
        const int p = k * plane + col;
        const int kp = nbp(k, n2, g.p2);
        const T xp = kp >= 0 ? x[kp * plane + col] : T(0);
        const T a2p = kp >= 0 ? a2[kp * plane + col] : T(0);
        T t = c[p] * xc;
        if (im >= 0) t += a0[p] * x[p + om];
        if (ip >= 0) t += a0[p + op] * x[p + op];
        if (jm >= 0) t += a1[p] * x[p + ym];
        if (jp >= 0) t += a1[p + yp] * x[p + yp];
        if (km >= 0) t += a2c * xm;
        if (kp >= 0) t += a2p * xp;
        r[p] = f[p] - t;
        xm = xc;
        xc = xp;
        a2c = a2p;
        km = k;
    }
}

// fc = P^T r : one thread per coarse cell. The transfer kernels take their
// multiply-adds explicitly (fma_t) with contraction off, so that the scalar
// and vector forms round alike.
template <typename T, int D>
__global__ void __launch_bounds__(256) k_gather_t(Geo F, Geo C, const T *__restrict__ r, const T *__restrict__ pm,
                                                  const T *__restrict__ pp, T *__restrict__ fc) {
#pragma clang fp contract(off)
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= C.n0 || J >= C.n1) return;
    const int pc = (K * C.n1 + J) * C.n0 + I;
    const int i = D == 0 ? 2 * I : I, j = D == 1 ? 2 * J : J, k = D == 2 ? 2 * K : K;
    const int e = D == 0 ? i : (D == 1 ? j : k);
    const int nd = D == 0 ? F.n0 : (D == 1 ? F.n1 : F.n2);
    const int per = D == 0 ? F.p0 : (D == 1 ? F.p1 : F.p2);
    const int stride = D == 0 ? 1 : (D == 1 ? F.n0 : F.n0 * F.n1);
    const int pf = (k * F.n1 + j) * F.n0 + i;
    const int l = nbm(e, nd, per);
    T s = r[pf];
    if (e + 1 < nd) s = fma_t(pm[pc], r[pf + stride], s);
    if (l >= 0) {
        const int Jd = D == 0 ? I : (D == 1 ? J : K);
        const int nc = D == 0 ? C.n0 : (D == 1 ? C.n1 : C.n2);
        const int cper = D == 0 ? C.p0 : (D == 1 ? C.p1 : C.p2);
        const int jl = nbm(Jd, nc, cper);
        const int pl = D == 0 ? (K * C.n1 + J) * C.n0 + jl : (D == 1 ? (K * C.n1 + jl) * C.n0 + I
                                                                      : (jl * C.n1 + J) * C.n0 + I);
        s = fma_t(pp[pl], r[pf + (l - e) * stride], s);
    }
    fc[pc] = s;
}

// The same restriction for coarsening in y (D = 1) or z (D = 2) on levels
// whose x rows hold a multiple of CPL cells: x is not coarsened, so CPL
// consecutive coarse cells of a thread sit over CPL consecutive fine cells,
// and every array is read with 16-byte loads. Per cell the arithmetic of
// k_gather_t, in the same order.
template <typename T, int CPL, int D>
__global__ void __launch_bounds__(256) k_gather_v(Geo F, Geo C, const T *__restrict__ r, const T *__restrict__ pm,
                                                  const T *__restrict__ pp, T *__restrict__ fc) {
#pragma clang fp contract(off)
    static_assert(D == 1 || D == 2, "x coarsening has no vector form");
    typedef Vec<T, CPL> V;
    const int I = (blockIdx.x * blockDim.x + threadIdx.x) * CPL;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= C.n0 || J >= C.n1) return;
    const int pc = (K * C.n1 + J) * C.n0 + I;
    const int j = D == 1 ? 2 * J : J, k = D == 2 ? 2 * K : K;
    const int e = D == 1 ? j : k;
    const int nd = D == 1 ? F.n1 : F.n2;
    const int per = D == 1 ? F.p1 : F.p2;
    const int stride = D == 1 ? F.n0 : F.n0 * F.n1;
    const int pf = (k * F.n1 + j) * F.n0 + I;
    const int l = nbm(e, nd, per);
    V s = vld<T, CPL>(r + pf);
    if (e + 1 < nd) {
        const V w = vld<T, CPL>(pm + pc), x = vld<T, CPL>(r + pf + stride);
#pragma unroll
        for (int q = 0; q < CPL; q++) s.v[q] = fma_t(w.v[q], x.v[q], s.v[q]);
    }
    if (l >= 0) {
        const int Jd = D == 1 ? J : K;
        const int jl = nbm(Jd, D == 1 ? C.n1 : C.n2, D == 1 ? C.p1 : C.p2);
        const int pl = D == 1 ? (K * C.n1 + jl) * C.n0 + I : (jl * C.n1 + J) * C.n0 + I;
        const V w = vld<T, CPL>(pp + pl), x = vld<T, CPL>(r + pf + (l - e) * stride);
#pragma unroll
        for (int q = 0; q < CPL; q++) s.v[q] = fma_t(w.v[q], x.v[q], s.v[q]);
    }
    *reinterpret_cast<V *>(fc + pc) = s;
}

// The restriction for coarsening in x on levels whose rows hold a multiple of
// 2 CPL fine cells: CPL coarse cells per thread, their 2 CPL fine residuals
// read with two 16-byte loads, the one fine cell to the left (and its pp
// weight) with scalar loads. Per cell the arithmetic of k_gather_t<T, 0>.
template <typename T, int CPL>
__global__ void __launch_bounds__(256) k_gather_xv(Geo F, Geo C, const T *__restrict__ r, const T *__restrict__ pm,
                                                   const T *__restrict__ pp, T *__restrict__ fc) {
#pragma clang fp contract(off)
    typedef Vec<T, CPL> V;
    const int I0 = (blockIdx.x * blockDim.x + threadIdx.x) * CPL;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I0 >= C.n0 || J >= C.n1) return;
    const int rowc = (K * C.n1 + J) * C.n0, rowf = (K * F.n1 + J) * F.n0;
    const V r0 = vld<T, CPL>(r + rowf + 2 * I0), r1 = vld<T, CPL>(r + rowf + 2 * I0 + CPL);
    const V w = vld<T, CPL>(pm + rowc + I0), u = vld<T, CPL>(pp + rowc + I0);
    T rr[2 * CPL];
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        rr[q] = r0.v[q];
        rr[CPL + q] = r1.v[q];
    }
    V s;
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int I = I0 + q, e = 2 * I;
        T t = rr[2 * q];
        if (e + 1 < F.n0) t = fma_t(w.v[q], rr[2 * q + 1], t);
        const int l = nbm(e, F.n0, F.p0);
        if (l >= 0) {
            T ppl, rl;
            if (q > 0) {
                ppl = u.v[q - 1];
                rl = rr[2 * q - 1];
            } else {
                ppl = pp[rowc + nbm(I, C.n0, C.p0)];
                rl = r[rowf + l];
            }
            t = fma_t(ppl, rl, t);
        }
        s.v[q] = t;
    }
    *reinterpret_cast<V *>(fc + rowc + I0) = s;
}

// ------------------------------------------------------------------ fused sweep
// One symmetric-order Gauss-Seidel sweep, out of place:
//   first colour c1:  v1 = (f - sum a vin) / c        (vin: previous iterate)
//   second colour:    v2 = (f - sum a v1) / c
//   vout = v1 on c1 cells, v2 on the others.
// Identical to the in-place two-pass red/black sweep, since each colour reads
// only the other colour. ZERO_IN: vin == 0 (first sweep of the V-cycle).
// A block covers (SX+2) x (SY+2) cell positions; the outer ring only computes
// first-colour values needed by the core's second colour. The block marches
// over planes s: at step s it computes v1 on plane s+1 and v2 on plane s.
constexpr int SX = 64;

// Workgroup barrier that waits only for LDS traffic, so global loads issued
// for the next plane stay in flight across it (__syncthreads would also drain
// them). gfx9 s_waitcnt: vmcnt = 63, expcnt = 7, lgkmcnt = 0.
__device__ __forceinline__ void lds_barrier() {
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

template <typename T, bool ZERO_IN, int SY>
__global__ void __launch_bounds__(SX *SY) k_sweep(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                  const T *__restrict__ a1, const T *__restrict__ a2,
                                                  const T *__restrict__ f, const T *__restrict__ vin,
                                                  T *__restrict__ vout, int c1, int kc, int gx, int gy) {
    __shared__ T s_vin[3][SY][SX];
    __shared__ T s_a0[3][SY][SX];
    __shared__ T s_a1[3][SY][SX];
    __shared__ T s_v1[2][SY][SX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int iu = bx * (SX - 2) - 1 + tx;
    const int ju = by * (SY - 2) - 1 + ty;
    const int ii = wrapc(iu, n0, g.p0);
    const int jj = wrapc(ju, n1, g.p1);
    const bool valid = ii >= 0 && jj >= 0;
    const int col = valid ? jj * n0 + ii : 0;
    const bool left = tx == 0, right = tx == SX - 1, bot = ty == 0, top = ty == SY - 1;
    const bool corner = (left || right) && (bot || top);
    const bool core = !left && !right && !bot && !top && iu < n0 && ju < n1;
    // the one cell outside the tile a ring position couples to (-1: none)
    int cout = -1;
    if (valid && !corner) {
        if (left) { int q = wrapc(ii - 1, n0, g.p0); cout = q >= 0 ? jj * n0 + q : -1; }
        if (right) { int q = wrapc(ii + 1, n0, g.p0); cout = q >= 0 ? jj * n0 + q : -1; }
        if (bot) { int q = wrapc(jj - 1, n1, g.p1); cout = q >= 0 ? q * n0 + ii : -1; }
        if (top) { int q = wrapc(jj + 1, n1, g.p1); cout = q >= 0 ? q * n0 + ii : -1; }
    }
    // coupling coefficient to that outside cell lives there for +x/+y, here for -x/-y
    const T *__restrict__ aout = right ? a0 : a1;
    const bool out_owns = right || top;
    const int par = (ii + jj) & 1;  // colour of the cell on plane k is (par + k) & 1
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);

    auto kwrap = [&](int k) { return wrapc(k, n2, g.p2); };
    auto ld = [&](const T *__restrict__ a, int kk) -> T { return (valid && kk >= 0) ? a[kk * plane + col] : T(0); };
    auto ldo = [&](const T *__restrict__ a, int kk) -> T { return (cout >= 0 && kk >= 0) ? a[kk * plane + cout] : T(0); };

    int s = k0 - 2;
    // data of plane s+1 (coefficients, vin, outside values) and of plane s+2 (vin, a2)
    int kA = kwrap(s + 1);
    T vin_s = ZERO_IN ? T(0) : ld(vin, kwrap(s));
    T c_A = valid && kA >= 0 ? c[kA * plane + col] : T(1);
    T f_A = ld(f, kA), a0_A = ld(a0, kA), a1_A = ld(a1, kA), a2_A = ld(a2, kA);
    T vin_A = ZERO_IN ? T(0) : ld(vin, kA);
    T vo_A = ZERO_IN ? T(0) : ldo(vin, kA);
    T ao_A = out_owns ? ldo(aout, kA) : T(0);
    int kB = kwrap(s + 2);
    T vin_B = ZERO_IN ? T(0) : ld(vin, kB);
    T a2_B = ld(a2, kB);
    T c_s = T(1), f_s = T(0), a0_s = T(0), a1_s = T(0), a2_s = T(0);
    T v1_m = T(0), v1_0 = T(0);
    int slot3 = 0;  // slot of plane s+1 in the 3-deep ring
    for (; s < k1; s++) {
        // prefetch: plane s+2 coefficients and outside values, plane s+3 vin and a2
        const int kN = kwrap(s + 2), kN2 = kwrap(s + 3);
        const T c_N = valid && kN >= 0 ? c[kN * plane + col] : T(1);
        const T f_N = ld(f, kN), a0_N = ld(a0, kN), a1_N = ld(a1, kN);
        const T vo_N = ZERO_IN ? T(0) : ldo(vin, kN);
        const T ao_N = out_owns ? ldo(aout, kN) : T(0);
        const T vin_N2 = ZERO_IN ? T(0) : ld(vin, kN2);
        const T a2_N2 = ld(a2, kN2);
        // publish plane s+1
        if (!ZERO_IN) s_vin[slot3][ty][tx] = vin_A;
        s_a0[slot3][ty][tx] = a0_A;
        s_a1[slot3][ty][tx] = a1_A;
        lds_barrier();
        const int ks = kwrap(s);
        // first colour on plane s+1
        T v1_p = T(0);
        if (valid && kA >= 0 && !corner && ((par + kA) & 1) == c1) {
            T sum = f_A;
            if (!ZERO_IN) {
                const T xm = left ? vo_A : s_vin[slot3][ty][tx - 1];
                const T xp = right ? vo_A : s_vin[slot3][ty][tx + 1];
                const T ym = bot ? vo_A : s_vin[slot3][ty - 1][tx];
                const T yp = top ? vo_A : s_vin[slot3][ty + 1][tx];
                const T a0p = right ? ao_A : s_a0[slot3][ty][tx + 1];
                const T a1p = top ? ao_A : s_a1[slot3][ty + 1][tx];
                sum -= a0_A * xm;
                sum -= a0p * xp;
                sum -= a1_A * ym;
                sum -= a1p * yp;
                sum -= a2_A * vin_s;
                sum -= a2_B * vin_B;
            }
            v1_p = sum / c_A;
        }
        s_v1[(s + 1) & 1][ty][tx] = v1_p;
        // second colour on plane s (core only)
        if (s >= k0 && core && ks >= 0) {
            T out;
            if (((par + ks) & 1) == c1) {
                out = v1_0;
            } else {
                const int sl = (slot3 + 2) % 3;  // slot of plane s
                T sum = f_s;
                sum -= a0_s * s_v1[s & 1][ty][tx - 1];
                sum -= s_a0[sl][ty][tx + 1] * s_v1[s & 1][ty][tx + 1];
                sum -= a1_s * s_v1[s & 1][ty - 1][tx];
                sum -= s_a1[sl][ty + 1][tx] * s_v1[s & 1][ty + 1][tx];
                sum -= a2_s * v1_m;
                sum -= a2_A * v1_p;
                out = sum / c_s;
            }
            vout[ks * plane + col] = out;
        }
        // rotate: plane s+1 -> s, prefetched plane s+2 -> s+1
        vin_s = vin_A;
        c_s = c_A;
        f_s = f_A;
        a0_s = a0_A;
        a1_s = a1_A;
        a2_s = a2_A;
        kA = kN;
        c_A = c_N;
        f_A = f_N;
        a0_A = a0_N;
        a1_A = a1_N;
        a2_A = a2_B;
        vin_A = vin_B;
        vo_A = vo_N;
        ao_A = ao_N;
        vin_B = vin_N2;
        a2_B = a2_N2;
        v1_m = v1_0;
        v1_0 = v1_p;
        slot3 = (slot3 + 1) % 3;
    }
}

// ------------------------------------------------------------------ sweep over whole rows
// The sweep of k_sweep on levels whose x rows fit one workgroup row
// (n0 <= NX, NX = 32, 64 or 128): a workgroup of NX x NY threads (NX NY =
// 1024) holds whole rows, so every x neighbour is in the tile (across the row
// end for periodic x) and only y has a recompute ring (NY - 2 output rows).
// On 128-cell rows k_sweep's 62-column tiles need three tile columns per row
// and half of the threads do no output; here a quarter do none. Per cell the
// arithmetic of k_sweep in the same order; a neighbour beyond a wall
// contributes 0 * 0 there and here, so the two kernels are bit-identical.
template <typename T, bool ZERO_IN, int NX>
__global__ void __launch_bounds__(1024) k_sweep_row(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                    const T *__restrict__ a1, const T *__restrict__ a2,
                                                    const T *__restrict__ f, const T *__restrict__ vin,
                                                    T *__restrict__ vout, int c1, int kc, int gy) {
    constexpr int NY = 1024 / NX;
    __shared__ T s_vin[3][NY][NX];
    __shared__ T s_a0[3][NY][NX];
    __shared__ T s_a1[3][NY][NX];
    __shared__ T s_v1[2][NY][NX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    int bx, by, bz;
    tile_of(1, gy, bx, by, bz);
    const int ju = by * (NY - 2) - 1 + ty;
    const int jj = wrapc(ju, n1, g.p1);
    const bool inrow = tx < n0;  // lanes past the row end hold zeros
    const bool valid = inrow && jj >= 0;
    const int col = valid ? jj * n0 + tx : 0;
    // in-row neighbours in the tile (-1: beyond a wall, or a lane past the row)
    const int txm = inrow ? nbm(tx, n0, g.p0) : -1;
    const int txp = inrow ? nbp(tx, n0, g.p0) : -1;
    const bool bot = ty == 0, top = ty == NY - 1;
    const bool core = !bot && !top && ju < n1 && inrow;
    // the one cell outside the tile a ring row couples to (-1: none)
    int cout = -1;
    if (valid) {
        if (bot) { int q = wrapc(jj - 1, n1, g.p1); cout = q >= 0 ? q * n0 + tx : -1; }
        if (top) { int q = wrapc(jj + 1, n1, g.p1); cout = q >= 0 ? q * n0 + tx : -1; }
    }
    const int par = (tx + jj) & 1;  // colour of the cell on plane k is (par + k) & 1
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);

    auto kwrap = [&](int k) { return wrapc(k, n2, g.p2); };
    auto ld = [&](const T *__restrict__ a, int kk) -> T { return (valid && kk >= 0) ? a[kk * plane + col] : T(0); };
    auto ldo = [&](const T *__restrict__ a, int kk) -> T { return (cout >= 0 && kk >= 0) ? a[kk * plane + cout] : T(0); };
    // an x neighbour's entry of an LDS plane (0 beyond a wall)
    auto lx = [&](const T (*S)[NX], int q) -> T { return q >= 0 ? S[ty][q] : T(0); };

    int s = k0 - 2;
    int kA = kwrap(s + 1);
    T vin_s = ZERO_IN ? T(0) : ld(vin, kwrap(s));
    T c_A = valid && kA >= 0 ? c[kA * plane + col] : T(1);
    T f_A = ld(f, kA), a0_A = ld(a0, kA), a1_A = ld(a1, kA), a2_A = ld(a2, kA);
    T vin_A = ZERO_IN ? T(0) : ld(vin, kA);
    T vo_A = ZERO_IN ? T(0) : ldo(vin, kA);
    T ao_A = top ? ldo(a1, kA) : T(0);
    int kB = kwrap(s + 2);
    T vin_B = ZERO_IN ? T(0) : ld(vin, kB);
    T a2_B = ld(a2, kB);
    T c_s = T(1), f_s = T(0), a0_s = T(0), a1_s = T(0), a2_s = T(0);
    T v1_m = T(0), v1_0 = T(0);
    int slot3 = 0;
    for (; s < k1; s++) {
        const int kN = kwrap(s + 2), kN2 = kwrap(s + 3);
        const T c_N = valid && kN >= 0 ? c[kN * plane + col] : T(1);
        const T f_N = ld(f, kN), a0_N = ld(a0, kN), a1_N = ld(a1, kN);
        const T vo_N = ZERO_IN ? T(0) : ldo(vin, kN);
        const T ao_N = top ? ldo(a1, kN) : T(0);
        const T vin_N2 = ZERO_IN ? T(0) : ld(vin, kN2);
        const T a2_N2 = ld(a2, kN2);
        if (!ZERO_IN) s_vin[slot3][ty][tx] = vin_A;
        s_a0[slot3][ty][tx] = a0_A;
        s_a1[slot3][ty][tx] = a1_A;
        lds_barrier();
        const int ks = kwrap(s);
        // first colour on plane s+1 (ring rows included)
        T v1_p = T(0);
        if (valid && kA >= 0 && ((par + kA) & 1) == c1) {
            T sum = f_A;
            if (!ZERO_IN) {
                const T xm = lx(s_vin[slot3], txm);
                const T xp = lx(s_vin[slot3], txp);
                const T ym = bot ? vo_A : s_vin[slot3][ty - 1][tx];
                const T yp = top ? vo_A : s_vin[slot3][ty + 1][tx];
                const T a0p = lx(s_a0[slot3], txp);
                const T a1p = top ? ao_A : s_a1[slot3][ty + 1][tx];
                sum -= a0_A * xm;
                sum -= a0p * xp;
                sum -= a1_A * ym;
                sum -= a1p * yp;
                sum -= a2_A * vin_s;
                sum -= a2_B * vin_B;
            }
            v1_p = sum / c_A;
        }
        s_v1[(s + 1) & 1][ty][tx] = v1_p;
        // second colour on plane s (core rows only)
        if (s >= k0 && core && ks >= 0) {
            T out;
            if (((par + ks) & 1) == c1) {
                out = v1_0;
            } else {
                const int sl = (slot3 + 2) % 3;
                T sum = f_s;
                sum -= a0_s * lx(s_v1[s & 1], txm);
                sum -= lx(s_a0[sl], txp) * lx(s_v1[s & 1], txp);
                sum -= a1_s * s_v1[s & 1][ty - 1][tx];
                sum -= s_a1[sl][ty + 1][tx] * s_v1[s & 1][ty + 1][tx];
                sum -= a2_s * v1_m;
                sum -= a2_A * v1_p;
                out = sum / c_s;
            }
            vout[ks * plane + col] = out;
        }
        vin_s = vin_A;
        c_s = c_A;
        f_s = f_A;
        a0_s = a0_A;
        a1_s = a1_A;
        a2_s = a2_A;
        kA = kN;
        c_A = c_N;
        f_A = f_N;
        a0_A = a0_N;
        a1_A = a1_N;
        a2_A = a2_B;
        vin_A = vin_B;
        vo_A = vo_N;
        ao_A = ao_N;
        vin_B = vin_N2;
        a2_B = a2_N2;
        v1_m = v1_0;
        v1_0 = v1_p;
        slot3 = (slot3 + 1) % 3;
    }
}

// ------------------------------------------------------------------ two fused sweeps
// Two consecutive sweeps of the same colour order c1 (the post-smoothing of a
// V(., 2) cycle) in one launch: four colour half-steps
//   h1: colour c1   from vin       h2: colour c2 from h1
//   h3: colour c1   from h2        h4: colour c2 from h3
// vout = h3 on c1 cells, h4 on c2 cells: the result of two k_sweep launches,
// with the same arithmetic in the same order, so bit-identical to them.
// A 64 x 16 workgroup recomputes a 3-cell ring (output core 58 x 10): h1 runs
// on every position (the outer ring reads the cells just outside the tile),
// h2 from depth 1, h3 from depth 2, h4 on the core. The planes are pipelined:
// at step s, h1 on plane s+3, h2 on s+2, h3 on s+1, h4 on s. Each field
// lives in LDS for the in-plane neighbours of the next half-step and in
// registers (three planes) for its z neighbours; a_x, a_y are in LDS for the
// four planes in use. One LDS barrier per step.
constexpr int R2 = 3;  // recompute ring of the fused pair

// ZERO_IN: vin == 0 (the first two sweeps of the pre-smoothing): h1 is f / c,
// vin is never read, and the pair equals the zero-iterate sweep followed by a
// sweep of the same order.
template <typename T, int SY, bool ZERO_IN = false>
__global__ void __launch_bounds__(SX *SY) k_sweep2(Geo g, const T *__restrict__ c, const T *__restrict__ a0,
                                                   const T *__restrict__ a1, const T *__restrict__ a2,
                                                   const T *__restrict__ f, const T *__restrict__ vin,
                                                   T *__restrict__ vout, int c1, int kc, int gx, int gy) {
    __shared__ T s_vin[2][SY][SX];
    __shared__ T s_a0[5][SY][SX];
    __shared__ T s_a1[5][SY][SX];
    __shared__ T s_h[3][2][SY][SX];  // fields h1, h2, h3, two planes each
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int n0 = g.n0, n1 = g.n1, n2 = g.n2;
    const int plane = n0 * n1;
    int bx, by, bz;
    tile_of(gx, gy, bx, by, bz);
    const int iu = bx * (SX - 2 * R2) - R2 + tx;
    const int ju = by * (SY - 2 * R2) - R2 + ty;
    const int ii = wrapc(iu, n0, g.p0);
    const int jj = wrapc(ju, n1, g.p1);
    const bool valid = ii >= 0 && jj >= 0;
    const int col = valid ? jj * n0 + ii : 0;
    const int depth = min(min(tx, SX - 1 - tx), min(ty, SY - 1 - ty));
    const bool left = tx == 0, right = tx == SX - 1, bot = ty == 0, top = ty == SY - 1;
    const bool corner = (left || right) && (bot || top);
    const bool core = depth >= R2 && iu < n0 && ju < n1;
    int cout = -1;
    if (valid && !corner) {
        if (left) { int q = wrapc(ii - 1, n0, g.p0); cout = q >= 0 ? jj * n0 + q : -1; }
        if (right) { int q = wrapc(ii + 1, n0, g.p0); cout = q >= 0 ? jj * n0 + q : -1; }
        if (bot) { int q = wrapc(jj - 1, n1, g.p1); cout = q >= 0 ? q * n0 + ii : -1; }
        if (top) { int q = wrapc(jj + 1, n1, g.p1); cout = q >= 0 ? q * n0 + ii : -1; }
    }
    const T *__restrict__ aout = right ? a0 : a1;
    const bool out_owns = right || top;
    const int par = (ii + jj) & 1;
    const int c2 = c1 ^ 1;
    const int k0 = bz * kc;
    const int k1 = min(n2, k0 + kc);

    auto kwrap = [&](int k) { return wrapc(k, n2, g.p2); };
    auto ld = [&](const T *__restrict__ a, int kk) -> T { return (valid && kk >= 0) ? a[kk * plane + col] : T(0); };
    auto ldo = [&](const T *__restrict__ a, int kk) -> T { return (cout >= 0 && kk >= 0) ? a[kk * plane + cout] : T(0); };
    auto ldc = [&](int kk) -> T { return valid && kk >= 0 ? c[kk * plane + col] : T(1); };
    // colour of this cell on (unwrapped) plane k, from its wrapped index
    auto colour = [&](int kk) { return (par + kk) & 1; };

    // registers: index 0 = plane P = s+3 (h1's plane), 1 = P-1, 2 = P-2, 3 = P-3
    T cq[4], fq[4];
    T a2q[5];        // a2 of planes P+1, P, P-1, P-2, P-3
    T vin_m, vin_p;  // vin of planes P-1 and P+1
    // the fields' values on the two planes below the one computed this step:
    // h1 of P-1, P-2; h2 of P-2, P-3; h3 of P-3, P-4
    T h1[2], h2[2], h3[2];
    // data of plane P for the first step
    int s = k0 - 6;
    int kP = kwrap(s + 3);
    for (int q = 0; q < 4; q++) {
        const int kq = kwrap(s + 3 - q);
        cq[q] = ldc(kq);
        fq[q] = ld(f, kq);
    }
    for (int q = 0; q < 5; q++) a2q[q] = ld(a2, kwrap(s + 4 - q));
    vin_m = ZERO_IN ? T(0) : ld(vin, kwrap(s + 2));
    vin_p = ZERO_IN ? T(0) : ld(vin, kwrap(s + 4));
    T vin_P = ZERO_IN ? T(0) : ld(vin, kP), a0_P = ld(a0, kP), a1_P = ld(a1, kP);
    T vo_P = ZERO_IN ? T(0) : ldo(vin, kP), ao_P = out_owns ? ldo(aout, kP) : T(0);
    for (int q = 0; q < 2; q++) h1[q] = h2[q] = h3[q] = T(0);
    int sa = 0;  // slot of plane P in the 5-deep a0/a1 ring
    for (; s < k1; s++) {
        const int P = s + 3;
        // prefetch plane P+1 (and vin, a2 of P+2)
        const int kN = kwrap(P + 1), kN2 = kwrap(P + 2);
        const T c_N = ldc(kN), f_N = ld(f, kN), a0_N = ld(a0, kN), a1_N = ld(a1, kN);
        const T vo_N = ZERO_IN ? T(0) : ldo(vin, kN), ao_N = out_owns ? ldo(aout, kN) : T(0);
        const T vin_N2 = ZERO_IN ? T(0) : ld(vin, kN2), a2_N2 = ld(a2, kN2);
        // publish plane P
        const int sv = P & 1;
        if (!ZERO_IN) s_vin[sv][ty][tx] = vin_P;
        s_a0[sa][ty][tx] = a0_P;
        s_a1[sa][ty][tx] = a1_P;
        lds_barrier();
        const int sa1 = sa == 0 ? 4 : sa - 1, sa2 = sa1 == 0 ? 4 : sa1 - 1, sa3 = sa2 == 0 ? 4 : sa2 - 1;
        // one half-step at plane p (ring slot sl of a_x, a_y): in-plane neighbours
        // from field slab F, z neighbours zm / zp, couplings a2 below / above
        auto half = [&](int q, int sl, const T (*F)[SX], T zm, T zp, T a2m, T a2p, bool ring) -> T {
            T sum = fq[q];
            const T a0s = s_a0[sl][ty][tx], a1s = s_a1[sl][ty][tx];
            const T xm = (ring && left) ? vo_P : F[ty][tx - 1];
            const T xp = (ring && right) ? vo_P : F[ty][tx + 1];
            const T ym = (ring && bot) ? vo_P : F[ty - 1][tx];
            const T yp = (ring && top) ? vo_P : F[ty + 1][tx];
            const T a0p = (ring && right) ? ao_P : s_a0[sl][ty][tx + 1];
            const T a1p = (ring && top) ? ao_P : s_a1[sl][ty + 1][tx];
            sum -= a0s * xm;
            sum -= a0p * xp;
            sum -= a1s * ym;
            sum -= a1p * yp;
            sum -= a2m * zm;
            sum -= a2p * zp;
            return sum / cq[q];
        };
        // h1 on P (every position but the corners; the outer ring reads outside values)
        T n1v = T(0);
        if (valid && kP >= 0 && !corner && colour(kP) == c1)
            n1v = ZERO_IN ? fq[0] / cq[0] : half(0, sa, s_vin[sv], vin_m, vin_p, a2q[1], a2q[0], true);
        // h2 on P-1 from h1 (depth >= 1)
        const int k_1 = kwrap(P - 1);
        T n2v = T(0);
        if (valid && k_1 >= 0 && depth >= 1 && colour(k_1) == c2)
            n2v = half(1, sa1, s_h[0][(P - 1) & 1], h1[1], n1v, a2q[2], a2q[1], false);
        // h3 on P-2 from h2 (depth >= 2)
        const int k_2 = kwrap(P - 2);
        T n3v = T(0);
        if (valid && k_2 >= 0 && depth >= 2 && colour(k_2) == c1)
            n3v = half(2, sa2, s_h[1][(P - 2) & 1], h2[1], n2v, a2q[3], a2q[2], false);
        // h4 on P-3 = s from h3, stored on the core
        const int ks = kwrap(s);
        if (s >= k0 && core && ks >= 0) {
            T out;
            if (colour(ks) == c1) {
                out = h3[0];
            } else {
                out = half(3, sa3, s_h[2][s & 1], h3[1], n3v, a2q[4], a2q[3], false);
            }
            vout[ks * plane + col] = out;
        }
        // publish the new field planes (read at the next step, after its barrier)
        s_h[0][P & 1][ty][tx] = n1v;
        s_h[1][(P - 1) & 1][ty][tx] = n2v;
        s_h[2][(P - 2) & 1][ty][tx] = n3v;
        // rotate: plane P+1 becomes P
        h3[1] = h3[0];
        h3[0] = n3v;
        h2[1] = h2[0];
        h2[0] = n2v;
        h1[1] = h1[0];
        h1[0] = n1v;
        for (int q = 3; q > 0; q--) {
            cq[q] = cq[q - 1];
            fq[q] = fq[q - 1];
        }
        cq[0] = c_N;
        fq[0] = f_N;
        for (int q = 4; q > 0; q--) a2q[q] = a2q[q - 1];
        a2q[0] = a2_N2;
        vin_m = vin_P;
        vin_P = vin_p;
        vin_p = vin_N2;
        a0_P = a0_N;
        a1_P = a1_N;
        vo_P = vo_N;
        ao_P = ao_N;
        kP = kN;
        sa = sa == 4 ? 0 : sa + 1;
    }
}

// ------------------------------------------------------------------ restriction
__device__ __forceinline__ int gdim_(const Geo &g, int d) { return d == 0 ? g.n0 : (d == 1 ? g.n1 : g.n2); }
__device__ __forceinline__ int gper_(const Geo &g, int d) { return d == 0 ? g.p0 : (d == 1 ? g.p1 : g.p2); }

// vout = vin + P vc: one thread per fine cell (vout may alias vin)
template <typename T, int D>
__global__ void __launch_bounds__(256) k_prolong_t(Geo F, Geo C, const T *__restrict__ pm, const T *__restrict__ pp,
                                                   const T *__restrict__ vc, const T *vin, T *vout) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= F.n0 || j >= F.n1) return;
    const int p = (k * F.n1 + j) * F.n0 + i;
    const int fd = D == 0 ? i : (D == 1 ? j : k);
    const int Jd = fd >> 1;
    const int I = D == 0 ? Jd : i, J = D == 1 ? Jd : j, K = D == 2 ? Jd : k;
    const int pl = (K * C.n1 + J) * C.n0 + I;
    T add;
    if ((fd & 1) == 0) {
        add = vc[pl];
    } else {
        add = pm[pl] * vc[pl];
        const int nc = D == 0 ? C.n0 : (D == 1 ? C.n1 : C.n2);
        const int cper = D == 0 ? C.p0 : (D == 1 ? C.p1 : C.p2);
        const int jr = nbp(Jd, nc, cper);
        if (jr >= 0) {
            const int pr = D == 0 ? (K * C.n1 + J) * C.n0 + jr : (D == 1 ? (K * C.n1 + jr) * C.n0 + I
                                                                          : (jr * C.n1 + J) * C.n0 + I);
            add = fma_t(pp[pl], vc[pr], add);
        }
    }
    vout[p] = (vin ? vin[p] : T(0)) + add;
}

// The same prolongation for coarsening in y (D = 1) or z (D = 2) on levels
// whose x rows hold a multiple of CPL cells: CPL consecutive fine cells per
// thread, 16-byte accesses, per cell the arithmetic of k_prolong_t.
template <typename T, int CPL, int D>
__global__ void __launch_bounds__(256) k_prolong_v(Geo F, Geo C, const T *__restrict__ pm, const T *__restrict__ pp,
                                                   const T *__restrict__ vc, const T *vin, T *vout) {
#pragma clang fp contract(off)
    static_assert(D == 1 || D == 2, "x coarsening has no vector form");
    typedef Vec<T, CPL> V;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) * CPL;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= F.n0 || j >= F.n1) return;
    const int p = (k * F.n1 + j) * F.n0 + i;
    const int fd = D == 1 ? j : k;
    const int Jd = fd >> 1;
    const int J = D == 1 ? Jd : j, K = D == 2 ? Jd : k;
    const int pl = (K * C.n1 + J) * C.n0 + i;
    V add;
    if ((fd & 1) == 0) {
        add = vld<T, CPL>(vc + pl);
    } else {
        const V w = vld<T, CPL>(pm + pl), x = vld<T, CPL>(vc + pl);
#pragma unroll
        for (int q = 0; q < CPL; q++) add.v[q] = w.v[q] * x.v[q];
        const int jr = nbp(Jd, D == 1 ? C.n1 : C.n2, D == 1 ? C.p1 : C.p2);
        if (jr >= 0) {
            const int pr = D == 1 ? (K * C.n1 + jr) * C.n0 + i : (jr * C.n1 + J) * C.n0 + i;
            const V w2 = vld<T, CPL>(pp + pl), x2 = vld<T, CPL>(vc + pr);
#pragma unroll
            for (int q = 0; q < CPL; q++) add.v[q] = fma_t(w2.v[q], x2.v[q], add.v[q]);
        }
    }
    // vin may be vout: this thread reads its cells before it writes them
    V a;
    if (vin) {
        a = *reinterpret_cast<const V *>(vin + p);
    } else {
#pragma unroll
        for (int q = 0; q < CPL; q++) a.v[q] = T(0);
    }
    V o;
#pragma unroll
    for (int q = 0; q < CPL; q++) o.v[q] = a.v[q] + add.v[q];
    *reinterpret_cast<V *>(vout + p) = o;
}

// The prolongation for coarsening in x on levels whose rows hold a multiple
// of 2 CPL fine cells: CPL fine cells per thread (16-byte accesses to vin and
// vout), over CPL / 2 coarse cells and the next one to the right. Per cell
// the arithmetic of k_prolong_t<T, 0>.
template <typename T, int CPL>
__global__ void __launch_bounds__(256) k_prolong_xv(Geo F, Geo C, const T *__restrict__ pm, const T *__restrict__ pp,
                                                    const T *__restrict__ vc, const T *vin, T *vout) {
#pragma clang fp contract(off)
    typedef Vec<T, CPL> V;
    constexpr int H = CPL / 2;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * CPL;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i0 >= F.n0 || j >= F.n1) return;
    const int rowc = (k * C.n1 + j) * C.n0, p = (k * F.n1 + j) * F.n0 + i0;
    const int J0 = i0 >> 1;
    T cv[H + 1], cm[H], cp[H];
#pragma unroll
    for (int q = 0; q < H; q++) {
        cv[q] = vc[rowc + J0 + q];
        cm[q] = pm[rowc + J0 + q];
        cp[q] = pp[rowc + J0 + q];
    }
    const int jr = nbp(J0 + H - 1, C.n0, C.p0);
    cv[H] = jr >= 0 ? vc[rowc + jr] : T(0);
    V a;  // vin may be vout: read before the write below
    if (vin) {
        a = *reinterpret_cast<const V *>(vin + p);
    } else {
#pragma unroll
        for (int q = 0; q < CPL; q++) a.v[q] = T(0);
    }
    V o;
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        const int h = q >> 1;
        T add;
        if ((q & 1) == 0) {
            add = cv[h];
        } else {
            add = cm[h] * cv[h];
            // the right neighbour exists for every coarse cell but a wall-side last one
            if (h < H - 1 || jr >= 0) add = fma_t(cp[h], cv[h + 1], add);
        }
        o.v[q] = a.v[q] + add;
    }
    *reinterpret_cast<V *>(vout + p) = o;
}

// coarsest level: v = Ainv f, one wavefront per row (Ainv symmetric)
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_dense_t(int n, const double *__restrict__ Ainv, const Tin *__restrict__ f,
                                                 Tout *__restrict__ v) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += Ainv[(long long)row * n + j] * (double)f[j];
    s = wave_sum(s);
    if (lane == 0) v[row] = (Tout)s;
}

// ------------------------------------------------------------------ bottom of the V-cycle
// The levels of at most BOT_MAX_CELLS cells, from the pre-smoothing of the
// first of them through the dense coarsest solve back to its post-smoothing,
// in one workgroup of 1024 threads with a barrier between phases. On these
// levels every launch of the per-level kernels costs a few microseconds of
// latency for a few thousand cells of work. Sweeps are in place (each colour
// reads only the other colour, so this equals the out-of-place sweeps of the
// large levels), and every sum keeps the order of the per-level kernels.
// A phase gives each thread at most one cell (BOT_MAX_CELLS = 1024): out of
// range threads are clamped to the last cell and do not store, so the code
// has no branch around the loads. (Taking 4096-cell levels here, with four
// cells per thread, measured slower than their per-level kernels.) Global
// memory written before a barrier is read after it by other threads of the
// same workgroup (one CU, one L1), which __syncthreads orders.
namespace {

constexpr int BT_THREADS = 1024;
constexpr int BT_CPT = 1;  // cells per thread and phase

// q = p / d for 0 <= p < 2^24 from a float reciprocal, corrected to exact
__device__ __forceinline__ int bt_div(int p, int d, float inv) {
    int q = (int)((float)p * inv);
    if (q * d > p) q--;
    if ((q + 1) * d <= p) q++;
    return q;
}

__device__ __forceinline__ void bt_ijk(const Geo &g, int p, int &i, int &j, int &k) {
    const float i0 = 1.0f / (float)g.n0, i1 = 1.0f / (float)g.n1;
    const int r = bt_div(p, g.n0, i0);
    i = p - r * g.n0;
    k = bt_div(r, g.n1, i1);
    j = r - k * g.n1;
}

// neighbour offsets of cell p (0 where there is no neighbour; `has` bits
// -x +x -y +y -z +z)
struct BtNb {
    int xm, xp, ym, yp, zm, zp;
    int has;
};

__device__ __forceinline__ BtNb bt_nb(const Geo &g, int p, int &parity) {
    int i, j, k;
    bt_ijk(g, p, i, j, k);
    parity = (i + j + k) & 1;
    const int n0 = g.n0, plane = g.n0 * g.n1;
    const int im = nbm(i, g.n0, g.p0), ip = nbp(i, g.n0, g.p0);
    const int jm = nbm(j, g.n1, g.p1), jp = nbp(j, g.n1, g.p1);
    const int km = nbm(k, g.n2, g.p2), kp = nbp(k, g.n2, g.p2);
    BtNb b;
    b.has = (im >= 0) | ((ip >= 0) << 1) | ((jm >= 0) << 2) | ((jp >= 0) << 3) | ((km >= 0) << 4) | ((kp >= 0) << 5);
    b.xm = im >= 0 ? p - i + im : p;
    b.xp = ip >= 0 ? p - i + ip : p;
    b.ym = jm >= 0 ? p + (jm - j) * n0 : p;
    b.yp = jp >= 0 ? p + (jp - j) * n0 : p;
    b.zm = km >= 0 ? p + (km - k) * plane : p;
    b.zp = kp >= 0 ? p + (kp - k) * plane : p;
    return b;
}

// in-place red/black phase: cells of `colour` get (f - sum a v_nb) / c
// (zero: the iterate is zero, so the cell gets f / c)
template <typename T>
__device__ void bt_colour(const BotLevel<T> &L, T *v, int colour, bool zero) {
    const int N = (int)L.g.N;
    for (int base = 0; base < N; base += BT_THREADS * BT_CPT) {
        T res[BT_CPT];
        int pp[BT_CPT];
        bool st[BT_CPT];
#pragma unroll
        for (int m = 0; m < BT_CPT; m++) {
            const int q = base + m * BT_THREADS + (int)threadIdx.x;
            const int p = q < N ? q : N - 1;
            int par;
            const BtNb b = bt_nb(L.g, p, par);
            T sum = L.f[p];
            if (!zero) {
                const T vxm = v[b.xm], vxp = v[b.xp], vym = v[b.ym], vyp = v[b.yp], vzm = v[b.zm], vzp = v[b.zp];
                const T axm = L.a0[p], axp = L.a0[b.xp], aym = L.a1[p], ayp = L.a1[b.yp], azm = L.a2[p],
                        azp = L.a2[b.zp];
                if (b.has & 1) sum -= axm * vxm;
                if (b.has & 2) sum -= axp * vxp;
                if (b.has & 4) sum -= aym * vym;
                if (b.has & 8) sum -= ayp * vyp;
                if (b.has & 16) sum -= azm * vzm;
                if (b.has & 32) sum -= azp * vzp;
            }
            res[m] = sum / L.c[p];
            pp[m] = p;
            st[m] = q < N && par == colour;
        }
#pragma unroll
        for (int m = 0; m < BT_CPT; m++)
            if (st[m]) v[pp[m]] = res[m];
    }
}

// w = f - A v
template <typename T>
__device__ void bt_resid(const BotLevel<T> &L, const T *v, T *w) {
    const int N = (int)L.g.N;
    for (int base = 0; base < N; base += BT_THREADS * BT_CPT) {
#pragma unroll
        for (int m = 0; m < BT_CPT; m++) {
            const int q = base + m * BT_THREADS + (int)threadIdx.x;
            const int p = q < N ? q : N - 1;
            int par;
            const BtNb b = bt_nb(L.g, p, par);
            const T vc = v[p], vxm = v[b.xm], vxp = v[b.xp], vym = v[b.ym], vyp = v[b.yp], vzm = v[b.zm],
                    vzp = v[b.zp];
            T t = L.c[p] * vc;
            if (b.has & 1) t += L.a0[p] * vxm;
            if (b.has & 2) t += L.a0[b.xp] * vxp;
            if (b.has & 4) t += L.a1[p] * vym;
            if (b.has & 8) t += L.a1[b.yp] * vyp;
            if (b.has & 16) t += L.a2[p] * vzm;
            if (b.has & 32) t += L.a2[b.zp] * vzp;
            if (q < N) w[p] = L.f[p] - t;
        }
    }
}

__device__ __forceinline__ int bt_dim(const Geo &g, int d) { return d == 0 ? g.n0 : (d == 1 ? g.n1 : g.n2); }
__device__ __forceinline__ int bt_per(const Geo &g, int d) { return d == 0 ? g.p0 : (d == 1 ? g.p1 : g.p2); }

// fc = P^T r along direction d (as k_gather_t)
template <typename T>
__device__ void bt_gather(const Geo &F, const Geo &C, int d, const T *r, const T *__restrict__ pm,
                          const T *__restrict__ pp, T *fc) {
    const int N = (int)C.N;
    for (int base = 0; base < N; base += BT_THREADS * BT_CPT) {
#pragma unroll
        for (int m = 0; m < BT_CPT; m++) {
            const int q = base + m * BT_THREADS + (int)threadIdx.x;
            const int pc = q < N ? q : N - 1;
            int I, J, K;
            bt_ijk(C, pc, I, J, K);
            const int i = d == 0 ? 2 * I : I, j = d == 1 ? 2 * J : J, k = d == 2 ? 2 * K : K;
            const int e = d == 0 ? i : (d == 1 ? j : k);
            const int nd = bt_dim(F, d), per = bt_per(F, d);
            const int stride = d == 0 ? 1 : (d == 1 ? F.n0 : F.n0 * F.n1);
            const int pf = (k * F.n1 + j) * F.n0 + i;
            const int l = nbm(e, nd, per);
            const bool hasr = e + 1 < nd;
            const int Jd = d == 0 ? I : (d == 1 ? J : K);
            const int jl = l >= 0 ? nbm(Jd, bt_dim(C, d), bt_per(C, d)) : Jd;
            const int pl = d == 0 ? (K * C.n1 + J) * C.n0 + jl
                                  : (d == 1 ? (K * C.n1 + jl) * C.n0 + I : (jl * C.n1 + J) * C.n0 + I);
            const T r0 = r[pf], rr = r[hasr ? pf + stride : pf], rl = r[l >= 0 ? pf + (l - e) * stride : pf];
            const T wm = pm[pc], wp = pp[pl];
            T s = r0;
            if (hasr) s += wm * rr;
            if (l >= 0) s += wp * rl;
            if (q < N) fc[pc] = s;
        }
    }
}

// v += P vc along direction d (as k_prolong_t)
template <typename T>
__device__ void bt_prolong(const Geo &F, const Geo &C, int d, const T *__restrict__ pm, const T *__restrict__ pp,
                           const T *vc, T *v) {
    const int N = (int)F.N;
    for (int base = 0; base < N; base += BT_THREADS * BT_CPT) {
#pragma unroll
        for (int m = 0; m < BT_CPT; m++) {
            const int q = base + m * BT_THREADS + (int)threadIdx.x;
            const int p = q < N ? q : N - 1;
            int i, j, k;
            bt_ijk(F, p, i, j, k);
            const int fd = d == 0 ? i : (d == 1 ? j : k);
            const int Jd = fd >> 1;
            const int I = d == 0 ? Jd : i, J = d == 1 ? Jd : j, K = d == 2 ? Jd : k;
            const int pl = (K * C.n1 + J) * C.n0 + I;
            const int jr = nbp(Jd, bt_dim(C, d), bt_per(C, d));
            const bool odd = fd & 1, hasr = jr >= 0;
            const int pr = !hasr ? pl
                                 : (d == 0 ? (K * C.n1 + J) * C.n0 + jr
                                           : (d == 1 ? (K * C.n1 + jr) * C.n0 + I : (jr * C.n1 + J) * C.n0 + I));
            const T v0 = v[p], c0 = vc[pl], c1 = vc[pr], wm = pm[pl], wp = pp[pl];
            T add;
            if (!odd) {
                add = c0;
            } else {
                add = wm * c0;
                if (hasr) add += wp * c1;
            }
            if (q < N) v[p] = v0 + add;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(BT_THREADS) k_bottom(BotArgs<T> a) {
    const int nl = a.nlev;
    // down: pre-smoothing, residual, restriction
    for (int l = 0; l < nl - 1; l++) {
        const BotLevel<T> &L = a.L[l];
        if (!L.relax) {
            // skipped level: v = 0 (the prolongation adds to it), f restricted as it is
            for (int p = threadIdx.x; p < (int)L.g.N; p += BT_THREADS) L.v[p] = T(0);
            bt_gather<T>(L.g, a.L[l + 1].g, L.cdir, L.f, a.L[l + 1].pm, a.L[l + 1].pp, a.L[l + 1].f);
            __syncthreads();
            continue;
        }
        for (int sw = 0; sw < a.npre; sw++) {
            bt_colour<T>(L, L.v, 0, sw == 0);
            __syncthreads();
            bt_colour<T>(L, L.v, 1, false);
            __syncthreads();
        }
        bt_resid<T>(L, L.v, L.w);
        __syncthreads();
        bt_gather<T>(L.g, a.L[l + 1].g, L.cdir, L.w, a.L[l + 1].pm, a.L[l + 1].pp, a.L[l + 1].f);
        __syncthreads();
    }
    // coarsest: v = Ainv f, one wavefront per row (lane-strided columns, as k_dense_t)
    {
        const BotLevel<T> &C = a.L[nl - 1];
        const int n = (int)C.g.N;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int row = wave; row < n; row += BT_THREADS / 64) {
            double s = 0.0;
            for (int jj = lane; jj < n; jj += 64) s += a.Ainv[(long long)row * n + jj] * (double)C.f[jj];
            s = wave_sum(s);
            if (lane == 0) C.v[row] = (T)s;
        }
        __syncthreads();
    }
    // up: prolongation and post-smoothing (the other colour first)
    for (int l = nl - 2; l >= 0; l--) {
        const BotLevel<T> &L = a.L[l];
        bt_prolong<T>(L.g, a.L[l + 1].g, L.cdir, a.L[l + 1].pm, a.L[l + 1].pp, a.L[l + 1].v, L.v);
        __syncthreads();
        if (!L.relax) continue;
        for (int sw = 0; sw < a.npost; sw++) {
            bt_colour<T>(L, L.v, 1, false);
            __syncthreads();
            bt_colour<T>(L, L.v, 0, false);
            __syncthreads();
        }
    }
}

}  // namespace

template <typename T>
void launch_bottom(hipStream_t s, const BotArgs<T> &a) {
    hipLaunchKernelGGL(k_bottom<T>, dim3(1), dim3(BT_THREADS), 0, s, a);
}

// ------------------------------------------------------------------ launchers

namespace {
int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
}  // namespace

// x-y tile for 256-thread blocks adapted to small planes
void tile_dims(const Geo &g, int &tx, int &ty) {
    tx = std::min(64, pow2ceil(g.n0));
    ty = std::min(256 / tx, pow2ceil(g.n1));
}

// 256-thread blocks for the vector transfers: x threads own cpl cells each
static void vec_tile_dims(const Geo &g, int cpl, int &tx, int &ty) {
    tx = std::min(64, pow2ceil(g.n0 / cpl));
    ty = std::min(256 / tx, pow2ceil(g.n1));
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// z chunk for the tiled march kernels (256-thread blocks)
int march_blocks(const Geo &g, int &kc) {
    int tx, ty;
    tile_dims(g, tx, ty);
    const int gx = (g.n0 + tx - 1) / tx, gy = (g.n1 + ty - 1) / ty;
    long long tiles = (long long)gx * gy;
    long long want = 4096;  // 256-thread blocks; z chunks of ~16 planes at 256^3
    kc = (int)std::max<long long>(1, (g.n2 * tiles + want - 1) / want);
    if (kc > g.n2) kc = g.n2;
    int gz = (g.n2 + kc - 1) / kc;
    while ((long long)gz * tiles > MAX_PARTS && kc < g.n2) {
        kc++;
        gz = (g.n2 + kc - 1) / kc;
    }
    return (int)(tiles * gz);
}

// launch shape of the SpMV: row waves (CPL cells of T per lane, 16 B) when
// the rows are long enough and a multiple of CPL, else the tiled march
struct ApplyCfg {
    bool rw;
    int kc, gx, gy, gz;
    int blocks;
};

template <typename T>
ApplyCfg apply_cfg(const Geo &g) {
    constexpr int CPL = 16 / sizeof(T);
    ApplyCfg a{};
    if (g.n0 % CPL == 0 && g.n0 >= 32 * CPL) {
        a.rw = true;
        a.gx = (g.n0 + 64 * CPL - 1) / (64 * CPL);
        a.gy = (g.n1 + 3) / 4;
        const long long tiles = (long long)a.gx * a.gy;
        // z chunks of ~4 planes at 256^3 (measured best among 4, 8, 16 for both
        // precisions: 5.6 TB/s FP32, 5.1 TB/s FP64 in a standalone SpMV)
        const long long want = 1024 * (long long)sizeof(T);
        a.kc = (int)std::max<long long>(1, (g.n2 * tiles + want - 1) / want);
        if (a.kc > g.n2) a.kc = g.n2;
        a.gz = (g.n2 + a.kc - 1) / a.kc;
        while ((long long)a.gz * tiles > MAX_PARTS && a.kc < g.n2) {
            a.kc++;
            a.gz = (g.n2 + a.kc - 1) / a.kc;
        }
        a.blocks = (int)(tiles * a.gz);
        return a;
    }
    a.rw = false;
    int tx, ty;
    tile_dims(g, tx, ty);
    a.blocks = march_blocks(g, a.kc);
    a.gx = (g.n0 + tx - 1) / tx;
    a.gy = (g.n1 + ty - 1) / ty;
    a.gz = (g.n2 + a.kc - 1) / a.kc;
    return a;
}

// workgroups of the FP64 SpMV (one partial of <x, Ax> each)
int apply_blocks(const Geo &g, int &kc) {
    const ApplyCfg a = apply_cfg<double>(g);
    kc = a.kc;
    return a.blocks;
}

template <typename T>
void launch_apply(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *x, T *y,
                  double *part) {
    const ApplyCfg a = apply_cfg<T>(g);
    if (a.rw) {
        hipLaunchKernelGGL((k_stencil_rw<T, 16 / sizeof(T), RW_APPLY>), dim3(a.blocks), dim3(256), 0, s, g, c, a0, a1,
                           a2, (const T *)nullptr, x, y, part, 0, a.kc, a.gx, a.gy);
        return;
    }
    int tx, ty;
    tile_dims(g, tx, ty);
    hipLaunchKernelGGL(k_apply_t<T>, dim3(a.blocks), dim3(tx, ty), 0, s, g, c, a0, a1, a2, x, y, part, a.kc, a.gx,
                       a.gy);
}

void launch_apply_x32(hipStream_t s, const Geo &g, const double *c, const double *a0, const double *a1,
                      const double *a2, const float *x, double *y, double *part) {
    const ApplyCfg a = apply_cfg<double>(g);
    if (a.rw) {
        hipLaunchKernelGGL((k_stencil_rw<double, 2, RW_APPLY, float>), dim3(a.blocks), dim3(256), 0, s, g, c, a0, a1,
                           a2, (const double *)nullptr, x, y, part, 0, a.kc, a.gx, a.gy);
        return;
    }
    int tx, ty;
    tile_dims(g, tx, ty);
    hipLaunchKernelGGL((k_apply_t<double, float>), dim3(a.blocks), dim3(tx, ty), 0, s, g, c, a0, a1, a2, x, y, part,
                       a.kc, a.gx, a.gy);
}

template <typename T>
void launch_smooth0(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f, T *v,
                    int c1) {
    const ApplyCfg a = apply_cfg<T>(g);
    if (a.rw) {
        hipLaunchKernelGGL((k_stencil_rw<T, 16 / sizeof(T), RW_SMOOTH0>), dim3(a.blocks), dim3(256), 0, s, g, c, a0,
                           a1, a2, f, (const T *)nullptr, v, (double *)nullptr, c1, a.kc, a.gx, a.gy);
        return;
    }
    int tx, ty, kc;
    tile_dims(g, tx, ty);
    march_blocks(g, kc);
    const int gx = (g.n0 + tx - 1) / tx, gy = (g.n1 + ty - 1) / ty, gz = (g.n2 + kc - 1) / kc;
    hipLaunchKernelGGL(k_smooth0_t<T>, dim3(gx * gy * gz), dim3(tx, ty), 0, s, g, c, a0, a1, a2, f, v, c1, kc, gx,
                       gy);
}

template <typename T>
void launch_resid(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                  const T *x, T *r) {
    const ApplyCfg a = apply_cfg<T>(g);
    if (a.rw) {
        hipLaunchKernelGGL((k_stencil_rw<T, 16 / sizeof(T), RW_RESID>), dim3(a.blocks), dim3(256), 0, s, g, c, a0, a1,
                           a2, f, x, r, (double *)nullptr, 0, a.kc, a.gx, a.gy);
        return;
    }
    int tx, ty, kc;
    tile_dims(g, tx, ty);
    march_blocks(g, kc);
    const int gx = (g.n0 + tx - 1) / tx, gy = (g.n1 + ty - 1) / ty, gz = (g.n2 + kc - 1) / kc;
    hipLaunchKernelGGL(k_resid_t<T>, dim3(gx * gy * gz), dim3(tx, ty), 0, s, g, c, a0, a1, a2, f, x, r, kc, gx,
                       gy);
}

template <typename T>
void launch_gather(hipStream_t s, const Geo &F, const Geo &C, int d, const T *r, const T *pm, const T *pp, T *fc,
                   bool vec) {
    constexpr int CPL = 16 / sizeof(T);
    if (vec && d == 0 && F.n0 % (2 * CPL) == 0 && aligned16(r) && aligned16(pm) && aligned16(pp) && aligned16(fc)) {
        int tx, ty;
        vec_tile_dims(C, CPL, tx, ty);
        dim3 grid((C.n0 / CPL + tx - 1) / tx, (C.n1 + ty - 1) / ty, C.n2);
        hipLaunchKernelGGL((k_gather_xv<T, CPL>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
        return;
    }
    if (vec && d != 0 && F.n0 % CPL == 0 && aligned16(r) && aligned16(pm) && aligned16(pp) && aligned16(fc)) {
        int tx, ty;
        vec_tile_dims(C, CPL, tx, ty);
        dim3 grid((C.n0 / CPL + tx - 1) / tx, (C.n1 + ty - 1) / ty, C.n2);
        if (d == 1)
            hipLaunchKernelGGL((k_gather_v<T, CPL, 1>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
        else
            hipLaunchKernelGGL((k_gather_v<T, CPL, 2>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
        return;
    }
    int tx, ty;
    tile_dims(C, tx, ty);
    dim3 grid((C.n0 + tx - 1) / tx, (C.n1 + ty - 1) / ty, C.n2);
    if (d == 0)
        hipLaunchKernelGGL((k_gather_t<T, 0>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
    else if (d == 1)
        hipLaunchKernelGGL((k_gather_t<T, 1>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
    else
        hipLaunchKernelGGL((k_gather_t<T, 2>), grid, dim3(tx, ty), 0, s, F, C, r, pm, pp, fc);
}

template <typename T, int SY>
static void launch_sweep_sy(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2,
                            const T *f, const T *vin, T *vout, int c1, long long target) {
    const int gx = (g.n0 + SX - 3) / (SX - 2), gy = (g.n1 + SY - 3) / (SY - 2);
    const long long tiles = (long long)gx * gy;
    // about `target` workgroups. Each z chunk recomputes 3 planes, so on large
    // levels (bandwidth-bound) chunks stay >= 16 planes when possible; small
    // levels are latency-bound (one barrier per plane), so there the chunks
    // get as short as the target allows.
    long long gz = std::max<long long>(1, target / tiles);
    if (g.N >= (1 << 22) && gz > 1 && (g.n2 + gz - 1) / gz < 16) gz = std::max<long long>(1, g.n2 / 16);
    if (gz > g.n2) gz = g.n2;
    int kc = (int)((g.n2 + gz - 1) / gz);
    if (kc < 1) kc = 1;
    const int nz = (g.n2 + kc - 1) / kc;
    dim3 grid((unsigned)(gx * gy * nz));
    if (vin)
        hipLaunchKernelGGL((k_sweep<T, false, SY>), grid, dim3(SX, SY), 0, s, g, c, a0, a1, a2, f, vin, vout, c1, kc,
                           gx, gy);
    else
        hipLaunchKernelGGL((k_sweep<T, true, SY>), grid, dim3(SX, SY), 0, s, g, c, a0, a1, a2, f,
                           (const T *)nullptr, vout, c1, kc, gx, gy);
}

// 64 x 16 workgroups (1024 threads, two per CU); ~512 of them per launch at
// 256^3 measured best among 64 x {4, 8, 16} tiles and 512..4096 groups. (A
// row-wave variant with 16-byte loads and 14-row workgroups measured slower:
// 183 against 137 us at 256^3, FP32.)
// Whole-row tiles (k_sweep_row) on levels of rows <= 128 cells: about
// `target` workgroups, z chunks of >= ROW_MIN_KC planes on large levels. With
// fewer tiles per plane than k_sweep the same workgroup count leaves shorter
// z chunks, hence fewer plane steps per workgroup.
constexpr int ROW_MIN_KC = 8;

template <typename T, int NX>
static void launch_sweep_row(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2,
                             const T *f, const T *vin, T *vout, int c1, long long target) {
    constexpr int NY = 1024 / NX;
    const int gy = (g.n1 + NY - 3) / (NY - 2);
    long long gz = std::max<long long>(1, target / gy);
    if (g.N >= (1 << 22) && gz > 1 && (g.n2 + gz - 1) / gz < ROW_MIN_KC)
        gz = std::max<long long>(1, g.n2 / ROW_MIN_KC);
    if (gz > g.n2) gz = g.n2;
    int kc = (int)((g.n2 + gz - 1) / gz);
    if (kc < 1) kc = 1;
    const int nz = (g.n2 + kc - 1) / kc;
    dim3 grid((unsigned)(gy * nz)), block(NX, NY);
    if (vin)
        hipLaunchKernelGGL((k_sweep_row<T, false, NX>), grid, block, 0, s, g, c, a0, a1, a2, f, vin, vout, c1, kc, gy);
    else
        hipLaunchKernelGGL((k_sweep_row<T, true, NX>), grid, block, 0, s, g, c, a0, a1, a2, f, (const T *)nullptr,
                           vout, c1, kc, gy);
}

template <typename T>
void launch_sweep(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                  const T *vin, T *vout, int c1, bool tiles) {
    if (!tiles && g.n0 <= 32)
        launch_sweep_row<T, 32>(s, g, c, a0, a1, a2, f, vin, vout, c1, 512);
    else if (!tiles && g.n0 <= 64)
        launch_sweep_row<T, 64>(s, g, c, a0, a1, a2, f, vin, vout, c1, 512);
    else if (!tiles && g.n0 <= 128)
        launch_sweep_row<T, 128>(s, g, c, a0, a1, a2, f, vin, vout, c1, 512);
    else
        launch_sweep_sy<T, 16>(s, g, c, a0, a1, a2, f, vin, vout, c1, 512);
}

// Two sweeps of colour order c1 in one launch (k_sweep2): 58 x 10 output
// tiles, about 512 workgroups, z chunks of >= 16 planes on large levels (each
// chunk recomputes 6 planes). vin and vout must differ; vin == nullptr: the
// first sweep starts from a zero iterate.
template <typename T>
void launch_sweep2(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                   const T *vin, T *vout, int c1) {
    constexpr int SY = 16;
    const int gx = (g.n0 + SX - 2 * R2 - 1) / (SX - 2 * R2), gy = (g.n1 + SY - 2 * R2 - 1) / (SY - 2 * R2);
    const long long tiles = (long long)gx * gy;
    long long gz = std::max<long long>(1, 512 / tiles);
    if (g.N >= (1 << 22) && gz > 1 && (g.n2 + gz - 1) / gz < 16) gz = std::max<long long>(1, g.n2 / 16);
    if (gz > g.n2) gz = g.n2;
    int kc = (int)((g.n2 + gz - 1) / gz);
    if (kc < 1) kc = 1;
    const int nz = (g.n2 + kc - 1) / kc;
    if (vin)
        hipLaunchKernelGGL((k_sweep2<T, SY, false>), dim3((unsigned)(gx * gy * nz)), dim3(SX, SY), 0, s, g, c, a0, a1,
                           a2, f, vin, vout, c1, kc, gx, gy);
    else
        hipLaunchKernelGGL((k_sweep2<T, SY, true>), dim3((unsigned)(gx * gy * nz)), dim3(SX, SY), 0, s, g, c, a0, a1,
                           a2, f, vin, vout, c1, kc, gx, gy);
}

template <typename T>
void launch_prolong_to(hipStream_t s, const Geo &F, const Geo &C, int d, const T *pm, const T *pp, const T *vc,
                       const T *vin, T *vout, bool vec) {
    constexpr int CPL = 16 / sizeof(T);
    if (vec && d == 0 && F.n0 % (2 * CPL) == 0 && aligned16(vin) && aligned16(vout)) {
        int tx, ty;
        vec_tile_dims(F, CPL, tx, ty);
        dim3 grid((F.n0 / CPL + tx - 1) / tx, (F.n1 + ty - 1) / ty, F.n2);
        hipLaunchKernelGGL((k_prolong_xv<T, CPL>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
        return;
    }
    if (vec && d != 0 && F.n0 % CPL == 0 && aligned16(pm) && aligned16(pp) && aligned16(vc) && aligned16(vin) &&
        aligned16(vout)) {
        int tx, ty;
        vec_tile_dims(F, CPL, tx, ty);
        dim3 grid((F.n0 / CPL + tx - 1) / tx, (F.n1 + ty - 1) / ty, F.n2);
        if (d == 1)
            hipLaunchKernelGGL((k_prolong_v<T, CPL, 1>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
        else
            hipLaunchKernelGGL((k_prolong_v<T, CPL, 2>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
        return;
    }
    int tx, ty;
    tile_dims(F, tx, ty);
    dim3 grid((F.n0 + tx - 1) / tx, (F.n1 + ty - 1) / ty, F.n2);
    if (d == 0)
        hipLaunchKernelGGL((k_prolong_t<T, 0>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
    else if (d == 1)
        hipLaunchKernelGGL((k_prolong_t<T, 1>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
    else
        hipLaunchKernelGGL((k_prolong_t<T, 2>), grid, dim3(tx, ty), 0, s, F, C, pm, pp, vc, vin, vout);
}

template <typename Tin, typename Tout>
void launch_dense(hipStream_t s, int n, const double *Ainv, const Tin *f, Tout *v) {
    hipLaunchKernelGGL((k_dense_t<Tin, Tout>), dim3((n + 3) / 4), dim3(256), 0, s, n, Ainv, f, v);
}

template void launch_bottom<double>(hipStream_t, const BotArgs<double> &);
template void launch_bottom<float>(hipStream_t, const BotArgs<float> &);

#define INST(T)                                                                                                    \
    template void launch_apply<T>(hipStream_t, const Geo &, const T *, const T *, const T *, const T *, const T *, \
                                  T *, double *);                                                                  \
    template void launch_sweep<T>(hipStream_t, const Geo &, const T *, const T *, const T *, const T *, const T *, \
                                  const T *, T *, int, bool);                                                      \
    template void launch_sweep2<T>(hipStream_t, const Geo &, const T *, const T *, const T *, const T *,          \
                                   const T *, const T *, T *, int);                                                \
    template void launch_smooth0<T>(hipStream_t, const Geo &, const T *, const T *, const T *, const T *,          \
                                    const T *, T *, int);                                                          \
    template void launch_resid<T>(hipStream_t, const Geo &, const T *, const T *, const T *, const T *, const T *, \
                                  const T *, T *);                                                                 \
    template void launch_gather<T>(hipStream_t, const Geo &, const Geo &, int, const T *, const T *, const T *,    \
                                   T *, bool);                                                                     \
    template void launch_prolong_to<T>(hipStream_t, const Geo &, const Geo &, int, const T *, const T *,           \
                                       const T *, const T *, T *, bool);                                           \
    template void launch_dense<T, T>(hipStream_t, int, const double *, const T *, T *);
INST(double)
INST(float)
#undef INST

}  // namespace ngap
