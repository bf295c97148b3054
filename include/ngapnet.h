// C ABI of the U-Net kernels (libngapnet.so). No counterpart in NGA2: the
// reference's pressure solve (src/solver/hypre_str_class.f90) has no learned
// stage. The network supplies initial guesses and a flexible-PCG stage.
// All pointers are device pointers; activations are channels-last [voxel][channel]; `stream` is a
// hipStream_t of this library's HIP runtime (NULL: default stream). A caller
// whose buffers come from another runtime instance (PyTorch wheels bundle
// their own) finishes its own work on them before the call and calls
// ngapnet_sync() before reading the results. Returns 0 on success, -1 for
// shapes the kernels do not take (MFMA layers: channels multiple of 32, grid
// multiple of 16 x 4 x 4), 1 for a launch failure.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
int ngapnet_conv3_fwd(const void *in_bf16, const void *wt_bf16, float *out32, void *outb_bf16, int nx, int ny,
                      int nz, int cin, int cout, int relu, void *stream);
int ngapnet_wgrad_parts(int nx, int ny, int nz, int cin, int cout);
int ngapnet_conv3_wgrad(const void *in_bf16, const void *dout_bf16, float *dw, float *part, int nx, int ny, int nz,
                        int cin, int cout, void *stream);
int ngapnet_flip_transpose(const void *wt_bf16, void *wd_bf16, int cin, int cout, void *stream);
int ngapnet_conv3_small_fwd(const float *in, const float *w, float *out32, void *outb_bf16, int nx, int ny, int nz,
                            int cin, int cout, int relu, void *stream);
/* part: NULL, or ngapnet_small_wgrad_part_size(cout) floats of scratch for the
 * one-pass kernel (cin == 1, cout == 32); other shapes use one workgroup per weight */
int ngapnet_conv3_small_wgrad(const float *in, const float *dout, float *dw, float *part, int nx, int ny, int nz,
                              int cin, int cout, void *stream);
long long ngapnet_small_wgrad_part_size(int cout);
int ngapnet_relu_bwd(const float *dy, const void *y_bf16, float *dx, void *dxb_bf16, long long n, void *stream);
int ngapnet_to_bf16(const float *s, void *d_bf16, long long n, void *stream);
int ngapnet_pool2(const void *in_bf16, void *out_bf16, int nx, int ny, int nz, int c, void *stream);
int ngapnet_pool2_bwd(const float *dout, float *din, int nx, int ny, int nz, int c, int acc, void *stream);
int ngapnet_up2(const void *in_bf16, void *out_bf16, int nx, int ny, int nz, int c, int ctot, int c0, void *stream);
int ngapnet_up2_bwd(const float *dout, float *din, int nx, int ny, int nz, int c, int ctot, int c0, void *stream);
int ngapnet_chan_copy_bf16(const void *src, void *dst, long long nv, int c, int cs, int s0, int cd, int d0,
                           void *stream);
int ngapnet_chan_copy_f32(const float *src, float *dst, long long nv, int c, int cs, int s0, int cd, int d0,
                          void *stream);
int ngapnet_conv1_fwd(const void *x_bf16, const float *w, float *y, long long nv, int c, void *stream);
int ngapnet_conv1_bwd(const void *x_bf16, const float *w, const float *dy, float *dx, float *dw, float *part,
                      long long nv, int c, void *stream);
int ngapnet_adam(float *w, const float *g, float *m, float *v, void *wb_bf16, long long n, float lr, float b1,
                 float b2, float eps, int step, float scale, void *stream);
int ngapnet_sync(void);

/* ---- the network as one object (ngap_netmodel.hip) ----
 * Weights are one flat FP32 vector, layer after layer in the order
 * c0 c1 c2a c2 b1 b d2 d1 d0 out. A 3x3x3 layer is [27][cout][cin] with
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1); the 1x1 output layer is [cin].
 * Grid (nx, ny, nz) needs nx % 128 == 0 and ny % 32 == 0 and nz % 32 == 0,
 * so that every level is tiled by 16 x 4 x 4. */
#define NGAPNET_NLAYERS 10
#define NGAPNET_ERR_SHAPE (-1)
#define NGAPNET_ERR_ARG (-2)
#define NGAPNET_ERR_HIP 1
#define NGAPNET_ERR_ALLOC 2
typedef struct ngapnet ngapnet;
/* layer i: level, channels, kind (0 FP32 3x3x3, 1 BF16 MFMA 3x3x3, 2 1x1), offset in the flat vector */
int ngapnet_layer_info(int i, int *lvl, int *cin, int *cout, int *kind, long long *offset);
long long ngapnet_num_weights(void);
int ngapnet_create(ngapnet **m, int nx, int ny, int nz);
int ngapnet_destroy(ngapnet *m);
/* host FP32 weights in and out; set_weights also resets the Adam state */
int ngapnet_set_weights(ngapnet *m, const float *host_w);
int ngapnet_get_weights(ngapnet *m, float *host_w);
/* gradients of the last backward, flat like the weights */
int ngapnet_get_grads(ngapnet *m, float *host_g);
/* y = net(x), x and y device FP32 vectors of nx*ny*nz cells (x fastest) */
int ngapnet_forward(ngapnet *m, const float *x, float *y, void *stream);
/* gradients of sum(dy * y) for the last forward */
int ngapnet_backward(ngapnet *m, const float *dy, void *stream);
/* Adam (beta 0.9, 0.999, eps 1e-8) on the FP32 masters; refreshes the BF16 copies */
int ngapnet_model_adam(ngapnet *m, float lr, void *stream);
/* forward, MSE loss against target, backward, Adam; *loss (host) = mean((y - target)^2) */
int ngapnet_train_step(ngapnet *m, const float *x, const float *target, float lr, double *loss, void *stream);
/* output scale of the correction stage */
int ngapnet_set_kappa(ngapnet *m, double kappa);
/* Correction stage with the signature of ngap_stage_fn (ngapsolve.h):
 * e = s * kappa * net(d r / s), s = rms(d r), r and e device FP64 vectors of n cells,
 * d the input scale of ngapnet_set_input_scale (1 when none is set).
 * ctx is the ngapnet object. */
int ngapnet_stage(void *ctx, const double *r, double *e, long long n, void *stream);
/* Per-cell input scale d of the stage (device FP64, n cells, x fastest), for
 * example 1 / diag(A): the network then sees the Jacobi-scaled residual, which
 * carries the local coefficient (the density of a two-phase operator). The
 * array is the caller's and must stay valid while the stage is attached;
 * d == NULL removes the scale. No counterpart in the reference. */
int ngapnet_set_input_scale(ngapnet *m, const double *d, long long n);
#ifdef __cplusplus
}
#endif
