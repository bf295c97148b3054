// The U-Net as one native object: weights, Adam state, activation and
// gradient buffers on the device, and forward, backward, Adam and the
// scale-free correction stage that the pressure solver calls between PCG
// iterations. Every convolution and transfer is a kernel of ngap_net.hip;
// this file only sequences them and adds the loss and scaling kernels.
//
// Architecture (level l has dims (nx, ny, nz) >> l):
//   L0  c0 = relu(conv(x, 1->32))                 d0 = relu(conv([up d1, c0], 64->32))   y = w_out . d0
//   L1  c1 = relu(conv(pool c0, 32->32))          d1 = relu(conv([up d2, c1], 96->32))
//   L2  c2 = relu(conv(relu(conv(pool c1, 32->64)), 64->64))   d2 = relu(conv([up b, c2], 128->64))
//   L3  b  = relu(conv(relu(conv(pool c2, 64->64)), 64->64))
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ngapnet.h"

namespace {

enum Kind { SMALL = 0, MFMA = 1, ONE = 2 };
struct Spec {
    int lvl, cin, cout, kind;
};
// order of the flat weight vector
constexpr Spec kSpec[NGAPNET_NLAYERS] = {
    {0, 1, 32, SMALL}, {1, 32, 32, MFMA},  {2, 32, 64, MFMA},  {2, 64, 64, MFMA}, {3, 64, 64, MFMA},
    {3, 64, 64, MFMA}, {2, 128, 64, MFMA}, {1, 96, 32, MFMA}, {0, 64, 32, MFMA}, {0, 32, 1, ONE},
};
enum { C0 = 0, C1, C2A, C2, B1, B, D2, D1, D0, OUT };

long long layer_size(int i) {
    const Spec &s = kSpec[i];
    return (s.kind == ONE ? 1LL : 27LL) * s.cin * s.cout;
}

constexpr int RED_BLOCKS = 1024;

// sum of squares of (y - t) per block, and dy = 2 (y - t) / n
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ y, const float *__restrict__ t,
                                             float *__restrict__ dy, double *__restrict__ part, long long n) {
    __shared__ double sm[4];
    double acc = 0.0;
    const float sc = 2.0f / (float)n;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float d = y[i] - t[i];
        dy[i] = sc * d;
        acc += (double)d * d;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

// sum of squares of d r per block (d == nullptr: d = 1)
__global__ void __launch_bounds__(256) k_sumsq64(const double *__restrict__ r, const double *__restrict__ d,
                                                 double *__restrict__ part, long long n) {
    __shared__ double sm[4];
    double acc = 0.0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double v = d ? d[i] * r[i] : r[i];
        acc += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

// out[0] = sum(part) / n (fixed order), out[1] = sqrt(out[0])
__global__ void __launch_bounds__(256) k_fin_mean(const double *__restrict__ part, int nb, long long n,
                                                  double *__restrict__ out) {
    __shared__ double sm[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) a += part[i];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = sm[0] / (double)n;
        out[0] = m;
        out[1] = sqrt(m);
    }
}

// x = d r / s (FP32), s = rms[1]; zero when s is 0 or not finite (d == nullptr: d = 1)
__global__ void k_scale_in(const double *__restrict__ r, const double *__restrict__ d,
                           const double *__restrict__ rms, float *__restrict__ x, long long n) {
    const double s = rms[1];
    const bool ok = s > 0.0 && isfinite(s);
    const double inv = ok ? 1.0 / s : 0.0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        x[i] = (float)((d ? d[i] * r[i] : r[i]) * inv);
}

// e = s kappa y (FP64)
__global__ void k_scale_out(const float *__restrict__ y, const double *__restrict__ rms, double kappa,
                            double *__restrict__ e, long long n) {
    const double s = rms[1];
    const double f = (s > 0.0 && isfinite(s)) ? s * kappa : 0.0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        e[i] = f * (double)y[i];
}

int blocks(long long n) {
    long long b = (n + 255) / 256;
    return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

}  // namespace

struct ngapnet {
    int shp[4][3];
    long long V[4];
    long long off[NGAPNET_NLAYERS];
    long long nw;
    float *w32 = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    __bf16 *wb = nullptr, *wflip = nullptr;
    float *x = nullptr, *y = nullptr;
    __bf16 *c0, *p1, *c1, *p2, *c2a, *c2, *p3, *b1, *b, *cat2, *d2, *cat1, *d1, *cat0, *d0;
    bool grads = false;
    float *fa, *fb, *fc, *sk0, *sk1, *sk2, *dy, *c1x1, *wparts;
    __bf16 *gbf;
    double *red = nullptr;  // RED_BLOCKS partials + 2 results
    double kappa = 1.0;
    const double *dscale = nullptr;  // stage input scale (caller's device array), or none
    int step = 0;
    std::vector<void *> allocs;

    template <typename T>
    int get(T **p, long long n) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (size_t)(n > 0 ? n : 1)) != hipSuccess) return 1;
        allocs.push_back(q);
        *p = (T *)q;
        return 0;
    }
    ~ngapnet() {
        for (void *q : allocs) (void)hipFree(q);
    }
};

namespace {

#define NCK(x)                      \
    do {                            \
        int rc_ = (x);              \
        if (rc_) return rc_;        \
    } while (0)

int alloc_grads(ngapnet *M) {
    if (M->grads) return 0;
    const long long *V = M->V;
    long long big = V[0] * 64;
    if (V[1] * 96 > big) big = V[1] * 96;
    if (V[2] * 128 > big) big = V[2] * 128;
    long long parts = ngapnet_small_wgrad_part_size(32);
    for (int i = 0; i < NGAPNET_NLAYERS; i++)
        if (kSpec[i].kind == MFMA) {
            const int *s = M->shp[kSpec[i].lvl];
            long long G = ngapnet_wgrad_parts(s[0], s[1], s[2], kSpec[i].cin, kSpec[i].cout);
            if (G * layer_size(i) > parts) parts = G * layer_size(i);
        }
    if (M->get(&M->fa, big) || M->get(&M->fb, big) || M->get(&M->fc, big) || M->get(&M->sk0, V[0] * 32) ||
        M->get(&M->sk1, V[1] * 32) || M->get(&M->sk2, V[2] * 64) || M->get(&M->dy, V[0]) ||
        M->get(&M->c1x1, 1024LL * 64) || M->get(&M->wparts, parts) || M->get(&M->gbf, big))
        return NGAPNET_ERR_ALLOC;
    M->grads = true;
    return 0;
}

int conv(ngapnet *M, int li, const __bf16 *in, __bf16 *outb, hipStream_t s) {
    const Spec &L = kSpec[li];
    const int *d = M->shp[L.lvl];
    return ngapnet_conv3_fwd(in, M->wb + M->off[li], nullptr, outb, d[0], d[1], d[2], L.cin, L.cout, 1, s);
}

int forward(ngapnet *M, hipStream_t s) {
    const long long *V = M->V;
    int(*S)[3] = M->shp;
    NCK(ngapnet_conv3_small_fwd(M->x, M->w32 + M->off[C0], nullptr, M->c0, S[0][0], S[0][1], S[0][2], 1, 32, 1, s));
    NCK(ngapnet_pool2(M->c0, M->p1, S[0][0], S[0][1], S[0][2], 32, s));
    NCK(conv(M, C1, M->p1, M->c1, s));
    NCK(ngapnet_pool2(M->c1, M->p2, S[1][0], S[1][1], S[1][2], 32, s));
    NCK(conv(M, C2A, M->p2, M->c2a, s));
    NCK(conv(M, C2, M->c2a, M->c2, s));
    NCK(ngapnet_pool2(M->c2, M->p3, S[2][0], S[2][1], S[2][2], 64, s));
    NCK(conv(M, B1, M->p3, M->b1, s));
    NCK(conv(M, B, M->b1, M->b, s));
    NCK(ngapnet_up2(M->b, M->cat2, S[2][0], S[2][1], S[2][2], 64, 128, 0, s));
    NCK(ngapnet_chan_copy_bf16(M->c2, M->cat2, V[2], 64, 64, 0, 128, 64, s));
    NCK(conv(M, D2, M->cat2, M->d2, s));
    NCK(ngapnet_up2(M->d2, M->cat1, S[1][0], S[1][1], S[1][2], 64, 96, 0, s));
    NCK(ngapnet_chan_copy_bf16(M->c1, M->cat1, V[1], 32, 32, 0, 96, 64, s));
    NCK(conv(M, D1, M->cat1, M->d1, s));
    NCK(ngapnet_up2(M->d1, M->cat0, S[0][0], S[0][1], S[0][2], 32, 64, 0, s));
    NCK(ngapnet_chan_copy_bf16(M->c0, M->cat0, V[0], 32, 32, 0, 64, 32, s));
    NCK(conv(M, D0, M->cat0, M->d0, s));
    NCK(ngapnet_conv1_fwd(M->d0, M->w32 + M->off[OUT], M->y, V[0], 32, s));
    return 0;
}

// weight gradient of layer li into g; data gradient (FP32) into dx unless null
int conv_bwd(ngapnet *M, int li, const __bf16 *x, const __bf16 *gy, float *dx, hipStream_t s) {
    const Spec &L = kSpec[li];
    const int *d = M->shp[L.lvl];
    NCK(ngapnet_conv3_wgrad(x, gy, M->g + M->off[li], M->wparts, d[0], d[1], d[2], L.cin, L.cout, s));
    if (dx) {
        NCK(ngapnet_flip_transpose(M->wb + M->off[li], M->wflip + M->off[li], L.cin, L.cout, s));
        NCK(ngapnet_conv3_fwd(gy, M->wflip + M->off[li], dx, nullptr, d[0], d[1], d[2], L.cout, L.cin, 0, s));
    }
    return 0;
}

int backward(ngapnet *M, const float *dy, hipStream_t s) {
    const long long *V = M->V;
    int(*S)[3] = M->shp;
    float *fa = M->fa, *fb = M->fb, *fc = M->fc;
    __bf16 *gb = M->gbf;
    NCK(ngapnet_conv1_bwd(M->d0, M->w32 + M->off[OUT], dy, fa, M->g + M->off[OUT], M->c1x1, V[0], 32, s));
    NCK(ngapnet_relu_bwd(fa, M->d0, nullptr, gb, V[0] * 32, s));
    NCK(conv_bwd(M, D0, M->cat0, gb, fb, s));  // fb = d(cat0) [V0][64]
    NCK(ngapnet_chan_copy_f32(fb, M->sk0, V[0], 32, 64, 32, 32, 0, s));
    NCK(ngapnet_up2_bwd(fb, fa, S[0][0], S[0][1], S[0][2], 32, 64, 0, s));  // fa = d(d1)
    NCK(ngapnet_relu_bwd(fa, M->d1, nullptr, gb, V[1] * 32, s));
    NCK(conv_bwd(M, D1, M->cat1, gb, fb, s));  // fb = d(cat1) [V1][96]
    NCK(ngapnet_chan_copy_f32(fb, M->sk1, V[1], 32, 96, 64, 32, 0, s));
    NCK(ngapnet_up2_bwd(fb, fa, S[1][0], S[1][1], S[1][2], 64, 96, 0, s));  // fa = d(d2)
    NCK(ngapnet_relu_bwd(fa, M->d2, nullptr, gb, V[2] * 64, s));
    NCK(conv_bwd(M, D2, M->cat2, gb, fb, s));  // fb = d(cat2) [V2][128]
    NCK(ngapnet_chan_copy_f32(fb, M->sk2, V[2], 64, 128, 64, 64, 0, s));
    NCK(ngapnet_up2_bwd(fb, fa, S[2][0], S[2][1], S[2][2], 64, 128, 0, s));  // fa = d(b)
    NCK(ngapnet_relu_bwd(fa, M->b, nullptr, gb, V[3] * 64, s));
    NCK(conv_bwd(M, B, M->b1, gb, fb, s));  // fb = d(b1)
    NCK(ngapnet_relu_bwd(fb, M->b1, nullptr, gb, V[3] * 64, s));
    NCK(conv_bwd(M, B1, M->p3, gb, fa, s));  // fa = d(p3)
    NCK(ngapnet_pool2_bwd(fa, M->sk2, S[2][0], S[2][1], S[2][2], 64, 1, s));
    NCK(ngapnet_relu_bwd(M->sk2, M->c2, nullptr, gb, V[2] * 64, s));
    NCK(conv_bwd(M, C2, M->c2a, gb, fb, s));  // fb = d(c2a)
    NCK(ngapnet_relu_bwd(fb, M->c2a, nullptr, gb, V[2] * 64, s));
    NCK(conv_bwd(M, C2A, M->p2, gb, fa, s));  // fa = d(p2)
    NCK(ngapnet_pool2_bwd(fa, M->sk1, S[1][0], S[1][1], S[1][2], 32, 1, s));
    NCK(ngapnet_relu_bwd(M->sk1, M->c1, nullptr, gb, V[1] * 32, s));
    NCK(conv_bwd(M, C1, M->p1, gb, fa, s));  // fa = d(p1)
    NCK(ngapnet_pool2_bwd(fa, M->sk0, S[0][0], S[0][1], S[0][2], 32, 1, s));
    NCK(ngapnet_relu_bwd(M->sk0, M->c0, fc, nullptr, V[0] * 32, s));
    NCK(ngapnet_conv3_small_wgrad(M->x, fc, M->g + M->off[C0], M->wparts, S[0][0], S[0][1], S[0][2], 1, 32, s));
    return 0;
}

int hip_ok(hipError_t e) { return e == hipSuccess ? 0 : NGAPNET_ERR_HIP; }

}  // namespace

extern "C" {

int ngapnet_layer_info(int i, int *lvl, int *cin, int *cout, int *kind, long long *offset) {
    if (i < 0 || i >= NGAPNET_NLAYERS) return NGAPNET_ERR_ARG;
    long long o = 0;
    for (int j = 0; j < i; j++) o += layer_size(j);
    if (lvl) *lvl = kSpec[i].lvl;
    if (cin) *cin = kSpec[i].cin;
    if (cout) *cout = kSpec[i].cout;
    if (kind) *kind = kSpec[i].kind;
    if (offset) *offset = o;
    return 0;
}

long long ngapnet_num_weights(void) {
    long long o = 0;
    for (int j = 0; j < NGAPNET_NLAYERS; j++) o += layer_size(j);
    return o;
}

int ngapnet_create(ngapnet **out, int nx, int ny, int nz) {
    if (!out) return NGAPNET_ERR_ARG;
    *out = nullptr;
    for (int l = 0; l < 4; l++) {
        const int a = nx >> l, b = ny >> l, c = nz >> l;
        if (a <= 0 || b <= 0 || c <= 0 || (a << l) != nx || (b << l) != ny || (c << l) != nz || a % 16 || b % 4 ||
            c % 4)
            return NGAPNET_ERR_SHAPE;
    }
    if ((long long)nx * ny * nz * 128 >= (1LL << 40)) return NGAPNET_ERR_SHAPE;
    ngapnet *M = new ngapnet();
    for (int l = 0; l < 4; l++) {
        M->shp[l][0] = nx >> l;
        M->shp[l][1] = ny >> l;
        M->shp[l][2] = nz >> l;
        M->V[l] = (long long)M->shp[l][0] * M->shp[l][1] * M->shp[l][2];
    }
    long long o = 0;
    for (int j = 0; j < NGAPNET_NLAYERS; j++) {
        M->off[j] = o;
        o += layer_size(j);
    }
    M->nw = o;
    const long long *V = M->V;
    int bad = M->get(&M->w32, o) || M->get(&M->m, o) || M->get(&M->v, o) || M->get(&M->g, o) || M->get(&M->wb, o) ||
              M->get(&M->wflip, o) || M->get(&M->x, V[0]) || M->get(&M->y, V[0]) || M->get(&M->c0, V[0] * 32) ||
              M->get(&M->p1, V[1] * 32) || M->get(&M->c1, V[1] * 32) || M->get(&M->p2, V[2] * 32) ||
              M->get(&M->c2a, V[2] * 64) || M->get(&M->c2, V[2] * 64) || M->get(&M->p3, V[3] * 64) ||
              M->get(&M->b1, V[3] * 64) || M->get(&M->b, V[3] * 64) || M->get(&M->cat2, V[2] * 128) ||
              M->get(&M->d2, V[2] * 64) || M->get(&M->cat1, V[1] * 96) || M->get(&M->d1, V[1] * 32) ||
              M->get(&M->cat0, V[0] * 64) || M->get(&M->d0, V[0] * 32) || M->get(&M->red, RED_BLOCKS + 2);
    if (bad) {
        delete M;
        return NGAPNET_ERR_ALLOC;
    }
    if (hipMemset(M->w32, 0, sizeof(float) * o) != hipSuccess || hipMemset(M->m, 0, sizeof(float) * o) != hipSuccess ||
        hipMemset(M->v, 0, sizeof(float) * o) != hipSuccess || hipMemset(M->g, 0, sizeof(float) * o) != hipSuccess ||
        hipMemset(M->wb, 0, sizeof(__bf16) * o) != hipSuccess) {
        delete M;
        return NGAPNET_ERR_HIP;
    }
    *out = M;
    return 0;
}

int ngapnet_destroy(ngapnet *M) {
    if (!M) return 0;
    (void)hipDeviceSynchronize();
    delete M;
    return 0;
}

int ngapnet_set_weights(ngapnet *M, const float *host_w) {
    if (!M || !host_w) return NGAPNET_ERR_ARG;
    const size_t b = sizeof(float) * (size_t)M->nw;
    NCK(hip_ok(hipMemcpy(M->w32, host_w, b, hipMemcpyHostToDevice)));
    NCK(hip_ok(hipMemset(M->m, 0, b)));
    NCK(hip_ok(hipMemset(M->v, 0, b)));
    NCK(ngapnet_to_bf16(M->w32, M->wb, M->nw, nullptr));
    NCK(hip_ok(hipDeviceSynchronize()));
    M->step = 0;
    return 0;
}

int ngapnet_get_weights(ngapnet *M, float *host_w) {
    if (!M || !host_w) return NGAPNET_ERR_ARG;
    NCK(hip_ok(hipDeviceSynchronize()));
    return hip_ok(hipMemcpy(host_w, M->w32, sizeof(float) * (size_t)M->nw, hipMemcpyDeviceToHost));
}

int ngapnet_get_grads(ngapnet *M, float *host_g) {
    if (!M || !host_g) return NGAPNET_ERR_ARG;
    NCK(hip_ok(hipDeviceSynchronize()));
    return hip_ok(hipMemcpy(host_g, M->g, sizeof(float) * (size_t)M->nw, hipMemcpyDeviceToHost));
}

int ngapnet_set_kappa(ngapnet *M, double kappa) {
    if (!M || !(kappa > 0.0) || !std::isfinite(kappa)) return NGAPNET_ERR_ARG;
    M->kappa = kappa;
    return 0;
}

int ngapnet_forward(ngapnet *M, const float *x, float *y, void *stream) {
    if (!M || !x || !y) return NGAPNET_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    NCK(hip_ok(hipMemcpyAsync(M->x, x, sizeof(float) * (size_t)M->V[0], hipMemcpyDeviceToDevice, s)));
    NCK(forward(M, s));
    NCK(hip_ok(hipMemcpyAsync(y, M->y, sizeof(float) * (size_t)M->V[0], hipMemcpyDeviceToDevice, s)));
    return hip_ok(hipGetLastError());
}

int ngapnet_backward(ngapnet *M, const float *dy, void *stream) {
    if (!M || !dy) return NGAPNET_ERR_ARG;
    NCK(alloc_grads(M));
    hipStream_t s = (hipStream_t)stream;
    NCK(backward(M, dy, s));
    return hip_ok(hipGetLastError());
}

int ngapnet_model_adam(ngapnet *M, float lr, void *stream) {
    if (!M) return NGAPNET_ERR_ARG;
    M->step++;
    return ngapnet_adam(M->w32, M->g, M->m, M->v, M->wb, M->nw, lr, 0.9f, 0.999f, 1e-8f, M->step, 1.0f, stream);
}

int ngapnet_train_step(ngapnet *M, const float *x, const float *target, float lr, double *loss, void *stream) {
    if (!M || !x || !target) return NGAPNET_ERR_ARG;
    NCK(alloc_grads(M));
    hipStream_t s = (hipStream_t)stream;
    const long long n = M->V[0];
    NCK(hip_ok(hipMemcpyAsync(M->x, x, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s)));
    NCK(forward(M, s));
    hipLaunchKernelGGL(k_mse, dim3(RED_BLOCKS), dim3(256), 0, s, (const float *)M->y, target, M->dy, M->red, n);
    hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(256), 0, s, (const double *)M->red, RED_BLOCKS, n,
                       M->red + RED_BLOCKS);
    NCK(backward(M, M->dy, s));
    NCK(ngapnet_model_adam(M, lr, stream));
    if (loss) {
        NCK(hip_ok(hipMemcpyAsync(loss, M->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, s)));
        NCK(hip_ok(hipStreamSynchronize(s)));
    }
    return hip_ok(hipGetLastError());
}

int ngapnet_set_input_scale(ngapnet *M, const double *d, long long n) {
    if (!M) return NGAPNET_ERR_ARG;
    if (d && n != M->V[0]) return NGAPNET_ERR_SHAPE;
    M->dscale = d;
    return 0;
}

int ngapnet_stage(void *ctx, const double *r, double *e, long long n, void *stream) {
    ngapnet *M = (ngapnet *)ctx;
    if (!M || !r || !e) return NGAPNET_ERR_ARG;
    if (n != M->V[0]) return NGAPNET_ERR_SHAPE;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sumsq64, dim3(RED_BLOCKS), dim3(256), 0, s, r, M->dscale, M->red, n);
    hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(256), 0, s, (const double *)M->red, RED_BLOCKS, n,
                       M->red + RED_BLOCKS);
    hipLaunchKernelGGL(k_scale_in, dim3(blocks(n)), dim3(256), 0, s, r, M->dscale,
                       (const double *)(M->red + RED_BLOCKS), M->x, n);
    NCK(forward(M, s));
    hipLaunchKernelGGL(k_scale_out, dim3(blocks(n)), dim3(256), 0, s, (const float *)M->y,
                       (const double *)(M->red + RED_BLOCKS), M->kappa, e, n);
    return hip_ok(hipGetLastError());
}

}  // extern "C"
