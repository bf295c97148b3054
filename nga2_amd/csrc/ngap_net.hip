// Kernels of the small 3-D U-Net that supplies initial guesses and a
// flexible-PCG stage to the pressure solver (gfx950).
//
// Activations are channels-last: a[v][c], v = x + nx (y + ny z), BF16 for the
// convolution inputs, FP32 elsewhere. Padding is periodic (the workloads the
// network serves are periodic boxes). 3x3x3 convolutions are implicit GEMMs
// on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation:
//   out[v][co] = sum_tap sum_ci W[tap][co][ci] * in[v + off(tap)][ci],
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1), off = (dx, dy, dz) in {-1,0,1}^3.
// Weight layout [27][Cout][Cin] (ci fastest) makes both MFMA operands 16-byte
// LDS reads. Channel counts of MFMA layers are multiples of 32; the first
// layer (1-2 input channels) and the 1x1 output layer use plain FP32 kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ngapnet.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// output tile of one workgroup: 16 x 4 x 4 voxels, 4 waves (one z-slice each)
constexpr int TXV = 16, TYV = 4, TZV = 4;
constexpr int HX = TXV + 2, HY = TYV + 2, HZ = TZV + 2;
constexpr int HVOX = HX * HY * HZ;  // 648 halo voxels
constexpr int CK = 32;              // channel chunk
constexpr int IN_STRIDE = 40;       // bf16 per halo voxel row in LDS (80 B: 16-B aligned, fewer bank conflicts)
constexpr int W_STRIDE = 40;        // bf16 per weight row (co) in LDS

// ---------------------------------------------------------------- forward 3x3x3
// grid.x = tiles, grid.y = Cout/32. Writes FP32 pre-activations (out32, may be
// null) and/or BF16 ReLU(out) (outb, may be null).
__global__ void __launch_bounds__(256) k_conv3_fwd(const __bf16 *__restrict__ in, const __bf16 *__restrict__ wt,
                                                   float *__restrict__ out32, __bf16 *__restrict__ outb, int nx,
                                                   int ny, int nz, int cin, int cout, int relu) {
    // weights are staged one plane of taps (9 of 27) at a time, so the tile
    // takes 41 + 18 KB of LDS and two workgroups fit on a CU
    __shared__ __attribute__((aligned(16))) __bf16 s_in[HVOX * IN_STRIDE];
    __shared__ __attribute__((aligned(16))) __bf16 s_w[9 * CK * W_STRIDE];
    const int tilesx = nx / TXV, tilesy = ny / TYV;
    const int t = blockIdx.x;
    const int x0 = (t % tilesx) * TXV;
    const int y0 = ((t / tilesx) % tilesy) * TYV;
    const int z0 = (t / (tilesx * tilesy)) * TZV;
    const int co0 = blockIdx.y * CK;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    f32x4 acc[TYV][2];
    for (int a = 0; a < TYV; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lrow = lane & 15, lk = (lane >> 4) * 8;
    for (int ci0 = 0; ci0 < cin; ci0 += CK) {
        // stage the halo tile: 648 voxels x 32 channels, 16 B per piece
        for (int q = tid; q < HVOX * 4; q += 256) {
            const int hv = q >> 2, part = q & 3;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
            const long long gv = (long long)gx + (long long)nx * (gy + (long long)ny * gz);
            const uint4 val = *reinterpret_cast<const uint4 *>(in + gv * cin + ci0 + part * 8);
            *reinterpret_cast<uint4 *>(s_in + hv * IN_STRIDE + part * 8) = val;
        }
        for (int dz = 0; dz < 3; dz++) {
            // stage the weights of taps dz*9 .. dz*9+8 of this ci chunk: 9 taps x 32 co x 32 ci
            for (int q = tid; q < 9 * CK * 4; q += 256) {
                const int row = q >> 2, part = q & 3;  // row = (tap - 9 dz)*32 + co
                const int tap = dz * 9 + row / CK, co = row % CK;
                const uint4 val =
                    *reinterpret_cast<const uint4 *>(wt + ((long long)tap * cout + co0 + co) * cin + ci0 + part * 8);
                *reinterpret_cast<uint4 *>(s_w + row * W_STRIDE + part * 8) = val;
            }
            __syncthreads();
            for (int t = 0; t < 9; t++) {
                const int dx = t % 3, dy = t / 3;  // halo offsets (0..2)
                bf16x8 bfr[2];
                for (int nb = 0; nb < 2; nb++)
                    bfr[nb] = *reinterpret_cast<const bf16x8 *>(s_w + (t * CK + nb * 16 + lrow) * W_STRIDE + lk);
                for (int mb = 0; mb < TYV; mb++) {
                    const int hv = (lrow + dx) + HX * ((mb + dy) + HY * (wave + dz));
                    const bf16x8 afr = *reinterpret_cast<const bf16x8 *>(s_in + hv * IN_STRIDE + lk);
                    for (int nb = 0; nb < 2; nb++)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nb], acc[mb][nb], 0, 0, 0);
                }
            }
            __syncthreads();
        }
    }
    // epilogue: lane holds D[m = 4*(lane/16)+q][n = lane%16] of each 16x16 block
    for (int mb = 0; mb < TYV; mb++)
        for (int nb = 0; nb < 2; nb++)
            for (int q = 0; q < 4; q++) {
                const int vx = x0 + 4 * (lane >> 4) + q, vy = y0 + mb, vz = z0 + wave;
                const long long v = (long long)vx + (long long)nx * (vy + (long long)ny * vz);
                const int co = co0 + nb * 16 + lrow;
                const float val = acc[mb][nb][q];
                if (out32) out32[v * cout + co] = val;
                if (outb) outb[v * cout + co] = (__bf16)(relu ? fmaxf(val, 0.f) : val);
            }
}

// ---------------------------------------------------------------- weight gradient
// dW[tap][co][ci] = sum_v dout[v][co] * in[v + off(tap)][ci], split over voxel
// tiles: workgroup (g, cob, cib) sums tiles g, g + G, g + 2G, ... into
// part[g][27][32][32] (co-major within the block). Four waves share the taps.
constexpr int KT = TXV * TYV * TZV;  // 256 voxels per tile
constexpr int DT_STRIDE = KT + 8;    // bf16 per co row of the transposed dout tile
// The transposed input tile is held once per x shift dx = 0, 1, 2 of the taps:
// copy dx holds halo columns dx .. dx+15 of every (y, z) halo row, so the 8
// consecutive x of a B fragment are one aligned 16-byte LDS read.
constexpr int WVOX = TXV * HY * HZ;  // 576 voxels of one shifted copy
constexpr int IT_STRIDE = WVOX + 8;  // bf16 per ci row of a shifted copy

__global__ void __launch_bounds__(256) k_conv3_wgrad(const __bf16 *__restrict__ in, const __bf16 *__restrict__ dout,
                                                     float *__restrict__ part, int nx, int ny, int nz, int cin,
                                                     int cout, int G) {
    __shared__ __attribute__((aligned(16))) __bf16 s_inT[3 * CK * IT_STRIDE];
    __shared__ __attribute__((aligned(16))) __bf16 s_dT[CK * DT_STRIDE];
    const int tilesx = nx / TXV, tilesy = ny / TYV, tilesz = nz / TZV;
    const int ntiles = tilesx * tilesy * tilesz;
    const int g = blockIdx.x, cob = blockIdx.y, cib = blockIdx.z;
    const int co0 = cob * CK, ci0 = cib * CK;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lq = lane >> 4;
    f32x4 acc[7][2][2];
    for (int a = 0; a < 7; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) acc[a][b][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = g; t < ntiles; t += G) {
        const int x0 = (t % tilesx) * TXV;
        const int y0 = ((t / tilesx) % tilesy) * TYV;
        const int z0 = (t / (tilesx * tilesy)) * TZV;
        // transposed input halo tile in its three x-shifted copies:
        // s_inT[dx][ci][w], w = (hx - dx) + TXV * (hy + HY * hz) for hx - dx in [0, TXV)
        for (int q = tid; q < HVOX * 4; q += 256) {
            const int hv = q >> 2, part = q & 3;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
            const long long gv = (long long)gx + (long long)nx * (gy + (long long)ny * gz);
            const bf16x8 val = *reinterpret_cast<const bf16x8 *>(in + gv * cin + ci0 + part * 8);
            for (int dx = 0; dx < 3; dx++) {
                const int wx = hx - dx;
                if (wx < 0 || wx >= TXV) continue;
                const int w = wx + TXV * (hy + HY * hz);
                for (int e = 0; e < 8; e++) s_inT[(dx * CK + part * 8 + e) * IT_STRIDE + w] = val[e];
            }
        }
        // transposed dout tile: s_dT[co][v]
        for (int q = tid; q < KT * 4; q += 256) {
            const int v = q >> 2, part = q & 3;
            const int vx = x0 + v % TXV, vy = y0 + (v / TXV) % TYV, vz = z0 + v / (TXV * TYV);
            const long long gv = (long long)vx + (long long)nx * (vy + (long long)ny * vz);
            const bf16x8 val = *reinterpret_cast<const bf16x8 *>(dout + gv * cout + co0 + part * 8);
            for (int e = 0; e < 8; e++) s_dT[(part * 8 + e) * DT_STRIDE + v] = val[e];
        }
        __syncthreads();
        for (int kc = 0; kc < KT / 32; kc++) {
            // voxels of this K chunk: v = 32*kc + 8*lq + e, e = 0..7 (consecutive x)
            const int vb = 32 * kc + 8 * lq;
            const int vx = vb % TXV, vy = (vb / TXV) % TYV, vz = vb / (TXV * TYV);
            bf16x8 afr[2];
            for (int mb = 0; mb < 2; mb++)
                afr[mb] = *reinterpret_cast<const bf16x8 *>(s_dT + (mb * 16 + lrow) * DT_STRIDE + vb);
            for (int tt = 0; tt < 7; tt++) {
                const int tap = wave + 4 * tt;
                if (tap >= 27) break;
                const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
                const int w = vx + TXV * ((vy + dy) + HY * (vz + dz));  // vx is 0 or 8
                for (int nb = 0; nb < 2; nb++) {
                    const bf16x8 bfr =
                        *reinterpret_cast<const bf16x8 *>(s_inT + (dx * CK + nb * 16 + lrow) * IT_STRIDE + w);
                    for (int mb = 0; mb < 2; mb++)
                        acc[tt][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[mb], bfr, acc[tt][mb][nb], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D[m = co][n = ci]: lane holds m = 4*lq + q, n = lrow
    float *dst = part + (long long)((g * gridDim.z + cib) * gridDim.y + cob) * 27 * CK * CK;
    for (int tt = 0; tt < 7; tt++) {
        const int tap = wave + 4 * tt;
        if (tap >= 27) break;
        for (int mb = 0; mb < 2; mb++)
            for (int nb = 0; nb < 2; nb++)
                for (int q = 0; q < 4; q++) {
                    const int co = mb * 16 + 4 * lq + q, ci = nb * 16 + lrow;
                    dst[(tap * CK + co) * CK + ci] = acc[tt][mb][nb][q];
                }
    }
}

// dW[tap][co][ci] = sum_g part[g][cib][cob][tap][co%32][ci%32] in fixed order
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float *__restrict__ part, float *__restrict__ dw, int cin,
                                                      int cout, int G) {
    const long long n = 27LL * cout * cin;
    const int ncb = cin / CK, nob = cout / CK;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ci = (int)(i % cin);
        const int co = (int)((i / cin) % cout);
        const int tap = (int)(i / ((long long)cin * cout));
        const int cib = ci / CK, cob = co / CK;
        const long long off = (long long)(tap * CK + co % CK) * CK + ci % CK;
        float s = 0.f;
        for (int g = 0; g < G; g++) s += part[(long long)((g * ncb + cib) * nob + cob) * 27 * CK * CK + off];
        dw[i] = s;
    }
}

// weights of the data gradient: wd[tap'][ci][co] = wt[26 - tap'][co][ci]
__global__ void k_flip_transpose(const __bf16 *__restrict__ wt, __bf16 *__restrict__ wd, int cin, int cout) {
    const long long n = 27LL * cin * cout;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int co = (int)(i % cout);
        const int ci = (int)((i / cout) % cin);
        const int tp = (int)(i / ((long long)cin * cout));
        wd[i] = wt[((long long)(26 - tp) * cout + co) * cin + ci];
    }
}

// ---------------------------------------------------------------- small-channel layers (FP32)
// out[v][co] = sum_tap sum_ci w[tap][co][ci] in[v+off][ci]   (cin small)
__global__ void k_conv3_small_fwd(const float *__restrict__ in, const float *__restrict__ w, float *__restrict__ out32,
                                  __bf16 *__restrict__ outb, int nx, int ny, int nz, int cin, int cout, int relu) {
    const long long nv = (long long)nx * ny * nz;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= nv * cout) return;
    const int co = (int)(i % cout);
    const long long v = i / cout;
    const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
    float s = 0.f;
    for (int tap = 0; tap < 27; tap++) {
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
        const long long u = wrapi(x + dx, nx) + (long long)nx * (wrapi(y + dy, ny) + (long long)ny * wrapi(z + dz, nz));
        for (int ci = 0; ci < cin; ci++) s += w[(tap * cout + co) * cin + ci] * in[u * cin + ci];
    }
    if (out32) out32[i] = s;
    if (outb) outb[i] = (__bf16)(relu ? fmaxf(s, 0.f) : s);
}

// The first layer (one input channel): one thread per voxel computes all COUT
// outputs from its 27 input values, with the weights in LDS (a broadcast read
// per tap and channel), and writes the voxel's COUT channels as 16-byte
// stores. Same summation order per output as k_conv3_small_fwd (taps in
// order). The per-(voxel, channel) form re-read the 27 inputs for every
// channel and spent its time in index arithmetic.
template <int COUT>
__global__ void __launch_bounds__(256) k_conv3_c1_fwd(const float *__restrict__ in, const float *__restrict__ w,
                                                      float *__restrict__ out32, __bf16 *__restrict__ outb, int nx,
                                                      int ny, int nz, int relu) {
    __shared__ float sw[27 * COUT];
    for (int q = threadIdx.x; q < 27 * COUT; q += 256) sw[q] = w[q];
    __syncthreads();
    const int nv = nx * ny * nz;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
    float acc[COUT];
#pragma unroll
    for (int co = 0; co < COUT; co++) acc[co] = 0.f;
#pragma unroll
    for (int tap = 0; tap < 27; tap++) {
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
        const float iv = in[wrapi(x + dx, nx) + nx * (wrapi(y + dy, ny) + ny * wrapi(z + dz, nz))];
#pragma unroll
        for (int co = 0; co < COUT; co++) acc[co] += sw[tap * COUT + co] * iv;
    }
    const long long o = (long long)v * COUT;
    if (out32) {
#pragma unroll
        for (int co = 0; co < COUT; co += 4)
            *reinterpret_cast<float4 *>(out32 + o + co) = make_float4(acc[co], acc[co + 1], acc[co + 2], acc[co + 3]);
    }
    if (outb) {
#pragma unroll
        for (int co = 0; co < COUT; co += 8) {
            bf16x8 b;
#pragma unroll
            for (int q = 0; q < 8; q++) b[q] = (__bf16)(relu ? fmaxf(acc[co + q], 0.f) : acc[co + q]);
            *reinterpret_cast<bf16x8 *>(outb + o + co) = b;
        }
    }
}

// dW of the small layer: one block per weight, fixed-order block reduction
__global__ void __launch_bounds__(256) k_conv3_small_wgrad(const float *__restrict__ in, const float *__restrict__ dout,
                                                           float *__restrict__ dw, int nx, int ny, int nz, int cin,
                                                           int cout) {
    __shared__ float sm[4];
    const int wi = blockIdx.x;  // tap*cout*cin + co*cin + ci
    const int ci = wi % cin, co = (wi / cin) % cout, tap = wi / (cin * cout);
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const long long nv = (long long)nx * ny * nz;
    float s = 0.f;
    for (long long v = threadIdx.x; v < nv; v += 256) {
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const long long u = wrapi(x + dx, nx) + (long long)nx * (wrapi(y + dy, ny) + (long long)ny * wrapi(z + dz, nz));
        s += dout[v * cout + co] * in[u * cin + ci];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) dw[wi] = sm[0] + sm[1] + sm[2] + sm[3];
}

// dW of a 1-input-channel layer over bricks of 32 x 8 x 8 voxels: the brick's
// input with a one-voxel halo sits in LDS, thread (voxel column vx, channel
// quad q) accumulates all 27 taps of its 4 channels (108 registers), so dout
// is read once, as one 16-byte load per thread and voxel. Partial bricks at
// the upper faces skip the voxels beyond the grid. Partials per workgroup,
// part[b][tap][co], summed in fixed order by k_colsum.
constexpr int SW_BLOCKS = 1024;
constexpr int SB_X = 32, SB_Y = 8, SB_Z = 8;
constexpr int SB_HX = SB_X + 2, SB_HY = SB_Y + 2, SB_HZ = SB_Z + 2;

template <int COUT>
__global__ void __launch_bounds__(256) k_small_wgrad_brick(const float *__restrict__ in, const float *__restrict__ dout,
                                                           float *__restrict__ part, int nx, int ny, int nz) {
    constexpr int NQ = COUT / 4;
    static_assert(256 / NQ == SB_X, "one voxel column per thread group");
    __shared__ float s_in[SB_HX * SB_HY * SB_HZ];
    __shared__ float s_red[4][27 * COUT];
    const int tid = threadIdx.x, q = tid % NQ, vx = tid / NQ;
    const int nbx = (nx + SB_X - 1) / SB_X, nby = (ny + SB_Y - 1) / SB_Y, nbz = (nz + SB_Z - 1) / SB_Z;
    const int nb = nbx * nby * nbz;
    float acc[27][4];
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int k = 0; k < 4; k++) acc[t][k] = 0.f;
    for (int brick = blockIdx.x; brick < nb; brick += gridDim.x) {
        const int x0 = (brick % nbx) * SB_X, y0 = ((brick / nbx) % nby) * SB_Y, z0 = (brick / (nbx * nby)) * SB_Z;
        __syncthreads();
        // halo coordinates beyond n are clamped to n: only voxels inside the
        // grid are accumulated, and they read at most one past the last
        for (int i = tid; i < SB_HX * SB_HY * SB_HZ; i += 256) {
            const int hx = i % SB_HX, hy = (i / SB_HX) % SB_HY, hz = i / (SB_HX * SB_HY);
            const int gx = wrapi(min(x0 - 1 + hx, nx), nx), gy = wrapi(min(y0 - 1 + hy, ny), ny),
                      gz = wrapi(min(z0 - 1 + hz, nz), nz);
            s_in[i] = in[gx + (long long)nx * (gy + (long long)ny * gz)];
        }
        __syncthreads();
        if (x0 + vx >= nx) continue;
        const int ly1 = min(SB_Y, ny - y0), lz1 = min(SB_Z, nz - z0);
        for (int lz = 0; lz < lz1; lz++)
            for (int ly = 0; ly < ly1; ly++) {
                const long long v = (x0 + vx) + (long long)nx * ((y0 + ly) + (long long)ny * (z0 + lz));
                const float4 d = reinterpret_cast<const float4 *>(dout + v * COUT)[q];
#pragma unroll
                for (int t = 0; t < 27; t++) {
                    const float x = s_in[(vx + t % 3) + SB_HX * ((ly + (t / 3) % 3) + SB_HY * (lz + t / 9))];
                    acc[t][0] += d.x * x;
                    acc[t][1] += d.y * x;
                    acc[t][2] += d.z * x;
                    acc[t][3] += d.w * x;
                }
            }
    }
    // lanes with the same channel quad: lane % NQ; sum over the wave, then the 4 waves
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float r = acc[t][k];
            for (int off = NQ; off < 64; off <<= 1) r += __shfl_xor(r, off, 64);
            if (lane < NQ) s_red[wave][t * COUT + 4 * q + k] = r;
        }
    __syncthreads();
    for (int i = tid; i < 27 * COUT; i += 256)
        part[(long long)blockIdx.x * 27 * COUT + i] = s_red[0][i] + s_red[1][i] + s_red[2][i] + s_red[3][i];
}

// out[i] = sum_b part[b][i], i < n: workgroup of 16 columns x 16 slices, slice
// s sums b = s, s + 16, ... and the 16 slice sums are added in slice order
__global__ void __launch_bounds__(256) k_colsum(const float *__restrict__ part, float *__restrict__ out, int nb, int n) {
    __shared__ float sm[16][17];
    const int c = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + c;
    float s = 0.f;
    if (i < n)
        for (int b = sl; b < nb; b += 16) s += part[(long long)b * n + i];
    sm[sl][c] = s;
    __syncthreads();
    if (sl == 0 && i < n) {
        float t = 0.f;
        for (int k = 0; k < 16; k++) t += sm[k][c];
        out[i] = t;
    }
}

// dx[v][c] = w[c] dy[v]; dW partials per block (C known at compile time, registers only)
template <int C>
__global__ void __launch_bounds__(256) k_conv1_bwd_t(const __bf16 *__restrict__ x, const float *__restrict__ w,
                                                     const float *__restrict__ dy, float *__restrict__ dx,
                                                     float *__restrict__ part, long long nv) {
    __shared__ float sm[4][C];
    float accw[C];
    float wr[C];
#pragma unroll
    for (int ch = 0; ch < C; ch++) {
        accw[ch] = 0.f;
        wr[ch] = w[ch];
    }
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < nv; v += (long long)gridDim.x * 256) {
        const float g = dy[v];
#pragma unroll
        for (int ch = 0; ch < C; ch++) {
            dx[v * C + ch] = wr[ch] * g;
            accw[ch] += g * (float)x[v * C + ch];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int ch = 0; ch < C; ch++) {
        float t = accw[ch];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) sm[wave][ch] = t;
    }
    __syncthreads();
    if (threadIdx.x < C)
        part[blockIdx.x * C + threadIdx.x] = sm[0][threadIdx.x] + sm[1][threadIdx.x] + sm[2][threadIdx.x] + sm[3][threadIdx.x];
}

// ---------------------------------------------------------------- elementwise
__global__ void k_relu_bwd(const float *__restrict__ dy, const __bf16 *__restrict__ y, float *__restrict__ dx,
                           __bf16 *__restrict__ dxb, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float g = (float)y[i] > 0.f ? dy[i] : 0.f;
        if (dx) dx[i] = g;
        if (dxb) dxb[i] = (__bf16)g;
    }
}

// The same in pieces of 4 (n % 4 == 0, dy/dx 16-byte and y/dxb 8-byte aligned)
__global__ void k_relu_bwd_v4(const float *__restrict__ dy, const __bf16 *__restrict__ y, float *__restrict__ dx,
                              __bf16 *__restrict__ dxb, long long n4) {
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        const float4 d = reinterpret_cast<const float4 *>(dy)[i];
        const bf16x4 yv = reinterpret_cast<const bf16x4 *>(y)[i];
        const float4 g = make_float4((float)yv[0] > 0.f ? d.x : 0.f, (float)yv[1] > 0.f ? d.y : 0.f,
                                     (float)yv[2] > 0.f ? d.z : 0.f, (float)yv[3] > 0.f ? d.w : 0.f);
        if (dx) reinterpret_cast<float4 *>(dx)[i] = g;
        if (dxb) {
            bf16x4 b;
            b[0] = (__bf16)g.x;
            b[1] = (__bf16)g.y;
            b[2] = (__bf16)g.z;
            b[3] = (__bf16)g.w;
            reinterpret_cast<bf16x4 *>(dxb)[i] = b;
        }
    }
}

__global__ void k_to_bf16(const float *__restrict__ s, __bf16 *__restrict__ d, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) d[i] = (__bf16)s[i];
}

// average pool 2x2x2 (bf16 -> bf16), fine dims nx,ny,nz
__global__ void k_pool2(const __bf16 *__restrict__ in, __bf16 *__restrict__ out, int nx, int ny, int nz, int c) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    const long long n = (long long)mx * my * mz * c;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % mx), y = (int)((v / mx) % my), z = (int)(v / ((long long)mx * my));
        float s = 0.f;
        for (int q = 0; q < 8; q++) {
            const long long u = (2 * x + (q & 1)) + (long long)nx * ((2 * y + ((q >> 1) & 1)) + (long long)ny * (2 * z + (q >> 2)));
            s += (float)in[u * c + ch];
        }
        out[i] = (__bf16)(0.125f * s);
    }
}

// d(in) of the pool: each fine voxel gets dout(coarse)/8 (fp32 -> fp32, added to dst when acc)
__global__ void k_pool2_bwd(const float *__restrict__ dout, float *__restrict__ din, int nx, int ny, int nz, int c,
                            int acc) {
    const long long n = (long long)nx * ny * nz * c;
    const int mx = nx / 2, my = ny / 2;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const long long u = (x / 2) + (long long)mx * ((y / 2) + (long long)my * (z / 2));
        const float g = 0.125f * dout[u * c + ch];
        din[i] = acc ? din[i] + g : g;
    }
}

// The same in pieces of 4 channels (c % 4 == 0)
__global__ void k_pool2_bwd_v4(const float *__restrict__ dout, float *__restrict__ din, int nx, int ny, int nz, int c,
                               int acc) {
    const int c4 = c / 4;
    const long long n = (long long)nx * ny * nz * c4;
    const int mx = nx / 2, my = ny / 2;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int g4 = (int)(i % c4);
        const long long v = i / c4;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const long long u = (x / 2) + (long long)mx * ((y / 2) + (long long)my * (z / 2));
        const float4 d = *reinterpret_cast<const float4 *>(dout + u * c + 4 * g4);
        float4 g = make_float4(0.125f * d.x, 0.125f * d.y, 0.125f * d.z, 0.125f * d.w);
        float4 *o = reinterpret_cast<float4 *>(din) + i;
        if (acc) {
            const float4 p = *o;
            g = make_float4(p.x + g.x, p.y + g.y, p.z + g.z, p.w + g.w);
        }
        *o = g;
    }
}

// out[fine][c0 + ch] = in[coarse][ch] (nearest-neighbour upsampling into a
// channel slice of a concatenation buffer with ctot channels)
__global__ void k_up2(const __bf16 *__restrict__ in, __bf16 *__restrict__ out, int nx, int ny, int nz, int c, int ctot,
                      int c0) {
    const long long n = (long long)nx * ny * nz * c;
    const int mx = nx / 2, my = ny / 2;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const long long u = (x / 2) + (long long)mx * ((y / 2) + (long long)my * (z / 2));
        out[v * ctot + c0 + ch] = in[u * c + ch];
    }
}

// The same copy in 16-byte pieces of 8 channels (c, ctot, c0 multiples of 8)
__global__ void k_up2_v8(const __bf16 *__restrict__ in, __bf16 *__restrict__ out, int nx, int ny, int nz, int c,
                         int ctot, int c0) {
    const int c8 = c / 8;
    const long long n = (long long)nx * ny * nz * c8;
    const int mx = nx / 2, my = ny / 2;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int g = (int)(i % c8);
        const long long v = i / c8;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        const long long u = (x / 2) + (long long)mx * ((y / 2) + (long long)my * (z / 2));
        *reinterpret_cast<bf16x8 *>(out + v * ctot + c0 + 8 * g) = *reinterpret_cast<const bf16x8 *>(in + u * c + 8 * g);
    }
}

// d(coarse) = sum over the 8 fine children of d(fine)[c0 + ch]
__global__ void k_up2_bwd(const float *__restrict__ dout, float *__restrict__ din, int nx, int ny, int nz, int c,
                          int ctot, int c0) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    const long long n = (long long)mx * my * mz * c;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        const int x = (int)(v % mx), y = (int)((v / mx) % my), z = (int)(v / ((long long)mx * my));
        float s = 0.f;
        for (int q = 0; q < 8; q++) {
            const long long u = (2 * x + (q & 1)) + (long long)nx * ((2 * y + ((q >> 1) & 1)) + (long long)ny * (2 * z + (q >> 2)));
            s += dout[u * ctot + c0 + ch];
        }
        din[i] = s;
    }
}

// The same in pieces of 4 channels (c, ctot, c0 multiples of 4)
__global__ void k_up2_bwd_v4(const float *__restrict__ dout, float *__restrict__ din, int nx, int ny, int nz, int c,
                             int ctot, int c0) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2, c4 = c / 4;
    const long long n = (long long)mx * my * mz * c4;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int g4 = (int)(i % c4);
        const long long v = i / c4;
        const int x = (int)(v % mx), y = (int)((v / mx) % my), z = (int)(v / ((long long)mx * my));
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = 0; q < 8; q++) {
            const long long u = (2 * x + (q & 1)) + (long long)nx * ((2 * y + ((q >> 1) & 1)) + (long long)ny * (2 * z + (q >> 2)));
            const float4 d = *reinterpret_cast<const float4 *>(dout + u * ctot + c0 + 4 * g4);
            s.x += d.x;
            s.y += d.y;
            s.z += d.z;
            s.w += d.w;
        }
        reinterpret_cast<float4 *>(din)[i] = s;
    }
}

// copy a channel slice: dst[v][d0 + ch] = src[v][s0 + ch], ch < c (bf16 or fp32)
// the same copy in 16-byte pieces (c, cs, s0, cd, d0 multiples of 16 / sizeof(T))
template <typename T>
__global__ void k_chan_copy_v16(const T *__restrict__ src, T *__restrict__ dst, long long nv, int c, int cs, int s0,
                                int cd, int d0) {
    constexpr int E = 16 / sizeof(T);
    const int ce = c / E;
    const long long n = nv * ce;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int g = (int)(i % ce);
        const long long v = i / ce;
        *reinterpret_cast<uint4 *>(dst + v * cd + d0 + g * E) = *reinterpret_cast<const uint4 *>(src + v * cs + s0 + g * E);
    }
}

template <typename T>
__global__ void k_chan_copy(const T *__restrict__ src, T *__restrict__ dst, long long nv, int c, int cs, int s0, int cd,
                            int d0) {
    const long long n = nv * c;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int ch = (int)(i % c);
        const long long v = i / c;
        dst[v * cd + d0 + ch] = src[v * cs + s0 + ch];
    }
}

// y[v] = sum_c w[c] x[v][c] (1x1 output layer, x bf16)
// y[v] = sum_ch w[ch] x[v][ch] for C channels: one thread per voxel reading its
// C bf16 values as 16-byte loads (C a multiple of 8), channels summed in order
template <int C>
__global__ void __launch_bounds__(256) k_conv1_fwd_t(const __bf16 *__restrict__ x, const float *__restrict__ w,
                                                     float *__restrict__ y, long long nv) {
    float wr[C];
#pragma unroll
    for (int ch = 0; ch < C; ch++) wr[ch] = w[ch];
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < nv; v += (long long)gridDim.x * 256) {
        float s = 0.f;
#pragma unroll
        for (int c8 = 0; c8 < C; c8 += 8) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(x + v * C + c8);
#pragma unroll
            for (int q = 0; q < 8; q++) s += wr[c8 + q] * (float)b[q];
        }
        y[v] = s;
    }
}

__global__ void k_conv1_fwd(const __bf16 *__restrict__ x, const float *__restrict__ w, float *__restrict__ y,
                            long long nv, int c) {
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < nv; v += (long long)gridDim.x * 256) {
        float s = 0.f;
        for (int ch = 0; ch < c; ch++) s += w[ch] * (float)x[v * c + ch];
        y[v] = s;
    }
}

// dx[v][c] = w[c] dy[v]; dW partials per block: part[b][c] = sum_v dy[v] x[v][c]
__global__ void __launch_bounds__(256) k_conv1_bwd(const __bf16 *__restrict__ x, const float *__restrict__ w,
                                                   const float *__restrict__ dy, float *__restrict__ dx,
                                                   float *__restrict__ part, long long nv, int c) {
    __shared__ float sm[4][64];
    float accw[64];
    for (int ch = 0; ch < c; ch++) accw[ch] = 0.f;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < nv; v += (long long)gridDim.x * 256) {
        const float g = dy[v];
        for (int ch = 0; ch < c; ch++) {
            dx[v * c + ch] = w[ch] * g;
            accw[ch] += g * (float)x[v * c + ch];
        }
    }
    for (int ch = 0; ch < c; ch++) {
        float s = accw[ch];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6][ch] = s;
    }
    __syncthreads();
    if (threadIdx.x < c) part[blockIdx.x * c + threadIdx.x] = sm[0][threadIdx.x] + sm[1][threadIdx.x] + sm[2][threadIdx.x] + sm[3][threadIdx.x];
}

// Adam on FP32 master weights; writes the BF16 copy
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       __bf16 *__restrict__ wb, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2, float scale) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float gi = g[i] * scale;
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = (__bf16)wi;
    }
}

int blocks_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (int)(b < 1 ? 1 : b);
}

template <typename T>
int chan_copy(const T *src, T *dst, long long nv, int c, int cs, int s0, int cd, int d0, void *stream) {
    constexpr int E = 16 / sizeof(T);
    if ((c | cs | s0 | cd | d0) % E == 0 && ((uintptr_t)src | (uintptr_t)dst) % 16 == 0)
        hipLaunchKernelGGL(k_chan_copy_v16<T>, dim3(blocks_for(nv * c / E)), dim3(256), 0, (hipStream_t)stream, src,
                           dst, nv, c, cs, s0, cd, d0);
    else
        hipLaunchKernelGGL(k_chan_copy<T>, dim3(blocks_for(nv * c)), dim3(256), 0, (hipStream_t)stream, src, dst, nv,
                           c, cs, s0, cd, d0);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

extern "C" {

int ngapnet_conv3_fwd(const void *in, const void *wt, float *out32, void *outb, int nx, int ny, int nz, int cin,
                      int cout, int relu, void *stream) {
    if (cin % CK || cout % CK || nx % TXV || ny % TYV || nz % TZV || cin <= 0 || cout <= 0) return -1;
    dim3 grid((nx / TXV) * (ny / TYV) * (nz / TZV), cout / CK);
    hipLaunchKernelGGL(k_conv3_fwd, grid, dim3(256), 0, (hipStream_t)stream, (const __bf16 *)in, (const __bf16 *)wt,
                       out32, (__bf16 *)outb, nx, ny, nz, cin, cout, relu);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_wgrad_parts(int nx, int ny, int nz, int cin, int cout) {
    const int ntiles = (nx / TXV) * (ny / TYV) * (nz / TZV);
    int G = ntiles < 256 ? ntiles : 256;
    (void)cin;
    (void)cout;
    return G;
}

int ngapnet_conv3_wgrad(const void *in, const void *dout, float *dw, float *part, int nx, int ny, int nz, int cin,
                        int cout, void *stream) {
    if (cin % CK || cout % CK || nx % TXV || ny % TYV || nz % TZV) return -1;
    const int G = ngapnet_wgrad_parts(nx, ny, nz, cin, cout);
    dim3 grid(G, cout / CK, cin / CK);
    hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(256), 0, (hipStream_t)stream, (const __bf16 *)in,
                       (const __bf16 *)dout, part, nx, ny, nz, cin, cout, G);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(27LL * cin * cout)), dim3(256), 0, (hipStream_t)stream,
                       (const float *)part, dw, cin, cout, G);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_flip_transpose(const void *wt, void *wd, int cin, int cout, void *stream) {
    hipLaunchKernelGGL(k_flip_transpose, dim3(blocks_for(27LL * cin * cout)), dim3(256), 0, (hipStream_t)stream,
                       (const __bf16 *)wt, (__bf16 *)wd, cin, cout);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_conv3_small_fwd(const float *in, const float *w, float *out32, void *outb, int nx, int ny, int nz,
                            int cin, int cout, int relu, void *stream) {
    const long long nv = (long long)nx * ny * nz;
    if (cin == 1 && cout == 32 && nv < (1LL << 31)) {
        hipLaunchKernelGGL(k_conv3_c1_fwd<32>, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, (hipStream_t)stream, in,
                           w, out32, (__bf16 *)outb, nx, ny, nz, relu);
        return hipGetLastError() == hipSuccess ? 0 : 1;
    }
    const long long n = nv * cout;
    hipLaunchKernelGGL(k_conv3_small_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, in, w,
                       out32, (__bf16 *)outb, nx, ny, nz, cin, cout, relu);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_conv3_small_wgrad(const float *in, const float *dout, float *dw, float *part, int nx, int ny, int nz,
                              int cin, int cout, void *stream) {
    if (part && cin == 1 && cout == 32 && (uintptr_t)dout % 16 == 0) {
        const long long bricks = (long long)((nx + SB_X - 1) / SB_X) * ((ny + SB_Y - 1) / SB_Y) * ((nz + SB_Z - 1) / SB_Z);
        const int grid = (int)(bricks < SW_BLOCKS ? bricks : SW_BLOCKS);
        hipLaunchKernelGGL(k_small_wgrad_brick<32>, dim3(grid), dim3(256), 0, (hipStream_t)stream, in, dout, part, nx,
                           ny, nz);
        hipLaunchKernelGGL(k_colsum, dim3((27 * cout + 15) / 16), dim3(256), 0, (hipStream_t)stream,
                           (const float *)part, dw, grid, 27 * cout);
    } else {
        hipLaunchKernelGGL(k_conv3_small_wgrad, dim3(27 * cin * cout), dim3(256), 0, (hipStream_t)stream, in, dout, dw,
                           nx, ny, nz, cin, cout);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

long long ngapnet_small_wgrad_part_size(int cout) { return (long long)SW_BLOCKS * 27 * cout; }

int ngapnet_relu_bwd(const float *dy, const void *y, float *dx, void *dxb, long long n, void *stream) {
    const bool v4 = n % 4 == 0 && ((uintptr_t)dy | (uintptr_t)dx) % 16 == 0 && ((uintptr_t)y | (uintptr_t)dxb) % 8 == 0;
    if (v4)
        hipLaunchKernelGGL(k_relu_bwd_v4, dim3(blocks_for(n / 4)), dim3(256), 0, (hipStream_t)stream, dy,
                           (const __bf16 *)y, dx, (__bf16 *)dxb, n / 4);
    else
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dy, (const __bf16 *)y,
                           dx, (__bf16 *)dxb, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_to_bf16(const float *s, void *d, long long n, void *stream) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, s, (__bf16 *)d, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_pool2(const void *in, void *out, int nx, int ny, int nz, int c, void *stream) {
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for((long long)nx * ny * nz * c / 8)), dim3(256), 0, (hipStream_t)stream,
                       (const __bf16 *)in, (__bf16 *)out, nx, ny, nz, c);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_pool2_bwd(const float *dout, float *din, int nx, int ny, int nz, int c, int acc, void *stream) {
    if (c % 4 == 0 && ((uintptr_t)dout | (uintptr_t)din) % 16 == 0)
        hipLaunchKernelGGL(k_pool2_bwd_v4, dim3(blocks_for((long long)nx * ny * nz * c / 4)), dim3(256), 0,
                           (hipStream_t)stream, dout, din, nx, ny, nz, c, acc);
    else
        hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for((long long)nx * ny * nz * c)), dim3(256), 0,
                           (hipStream_t)stream, dout, din, nx, ny, nz, c, acc);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_up2(const void *in, void *out, int nx, int ny, int nz, int c, int ctot, int c0, void *stream) {
    if (c % 8 == 0 && ctot % 8 == 0 && c0 % 8 == 0 && ((uintptr_t)in | (uintptr_t)out) % 16 == 0)
        hipLaunchKernelGGL(k_up2_v8, dim3(blocks_for((long long)nx * ny * nz * c / 8)), dim3(256), 0,
                           (hipStream_t)stream, (const __bf16 *)in, (__bf16 *)out, nx, ny, nz, c, ctot, c0);
    else
        hipLaunchKernelGGL(k_up2, dim3(blocks_for((long long)nx * ny * nz * c)), dim3(256), 0, (hipStream_t)stream,
                           (const __bf16 *)in, (__bf16 *)out, nx, ny, nz, c, ctot, c0);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_up2_bwd(const float *dout, float *din, int nx, int ny, int nz, int c, int ctot, int c0, void *stream) {
    if (c % 4 == 0 && ctot % 4 == 0 && c0 % 4 == 0 && ((uintptr_t)dout | (uintptr_t)din) % 16 == 0)
        hipLaunchKernelGGL(k_up2_bwd_v4, dim3(blocks_for((long long)nx * ny * nz * c / 32)), dim3(256), 0,
                           (hipStream_t)stream, dout, din, nx, ny, nz, c, ctot, c0);
    else
        hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for((long long)nx * ny * nz * c / 8)), dim3(256), 0,
                           (hipStream_t)stream, dout, din, nx, ny, nz, c, ctot, c0);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_chan_copy_bf16(const void *src, void *dst, long long nv, int c, int cs, int s0, int cd, int d0,
                           void *stream) {
    return chan_copy((const __bf16 *)src, (__bf16 *)dst, nv, c, cs, s0, cd, d0, stream);
}

int ngapnet_chan_copy_f32(const float *src, float *dst, long long nv, int c, int cs, int s0, int cd, int d0,
                          void *stream) {
    return chan_copy(src, dst, nv, c, cs, s0, cd, d0, stream);
}

int ngapnet_conv1_fwd(const void *x, const float *w, float *y, long long nv, int c, void *stream) {
    if (c == 32)
        hipLaunchKernelGGL(k_conv1_fwd_t<32>, dim3(blocks_for(nv)), dim3(256), 0, (hipStream_t)stream,
                           (const __bf16 *)x, w, y, nv);
    else
        hipLaunchKernelGGL(k_conv1_fwd, dim3(blocks_for(nv)), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)x, w,
                           y, nv, c);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_conv1_bwd(const void *x, const float *w, const float *dy, float *dx, float *dw, float *part, long long nv,
                      int c, void *stream) {
    if (c > 64) return -1;
    const int nb = 1024;
    if (c == 32)
        hipLaunchKernelGGL(k_conv1_bwd_t<32>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)x, w, dy, dx,
                           part, nv);
    else
        hipLaunchKernelGGL(k_conv1_bwd, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const __bf16 *)x, w, dy, dx,
                           part, nv, c);
    hipLaunchKernelGGL(k_colsum, dim3((c + 15) / 16), dim3(256), 0, (hipStream_t)stream, (const float *)part, dw, nb,
                       c);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_adam(float *w, const float *g, float *m, float *v, void *wb, long long n, float lr, float b1, float b2,
                 float eps, int step, float scale, void *stream) {
    float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, (__bf16 *)wb, n, lr,
                       b1, b2, eps, c1, c2, scale);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ngapnet_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 1; }

}  // extern "C"
