// Internal declarations shared by the solver's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ngap_kernels.h"
#include "ngapsolve.h"

namespace ngap {

int grid_for(long long n);
int grid_parts(long long n);

template <typename T>
void launch_apply(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *x, T *y,
                  double *part);
// y = A x with the FP64 operator and x stored in FP32 (the mixed solve's
// search directions); part as in launch_apply
void launch_apply_x32(hipStream_t s, const Geo &g, const double *c, const double *a0, const double *a1,
                      const double *a2, const float *x, double *y, double *part);
// one symmetric-order Gauss-Seidel sweep, out of place (vin == nullptr: zero
// initial iterate); c1 = colour updated first (0 red, 1 black). Levels of rows
// <= 128 cells run whole-row tiles (k_sweep_row) unless tiles is set.
template <typename T>
void launch_sweep(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                  const T *vin, T *vout, int c1, bool tiles = false);
// two fused sweeps of the same colour order (vin != vout)
template <typename T>
void launch_sweep2(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                   const T *vin, T *vout, int c1);
// first symmetric sweep from a zero iterate as a stencil (no LDS)
template <typename T>
void launch_smooth0(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f, T *v,
                    int c1);
template <typename T>
void launch_resid(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f,
                  const T *x, T *r);
// fc = P^T r (vec = false: the one-cell-per-thread kernel on every level)
template <typename T>
void launch_gather(hipStream_t s, const Geo &F, const Geo &C, int d, const T *r, const T *pm, const T *pp, T *fc,
                   bool vec = true);
// vout = vin + P vc (vout may be vin; vin == nullptr: vout = P vc)
template <typename T>
void launch_prolong_to(hipStream_t s, const Geo &F, const Geo &C, int d, const T *pm, const T *pp, const T *vc,
                       const T *vin, T *vout, bool vec = true);
template <typename Tin, typename Tout>
void launch_dense(hipStream_t s, int n, const double *Ainv, const Tin *f, Tout *v);
int apply_blocks(const Geo &g, int &kc);

// The small levels of a V-cycle in one workgroup (see k_bottom).
constexpr int BOT_MAX_LEVELS = 16;
constexpr long long BOT_MAX_CELLS = 1024;
template <typename T>
struct BotLevel {
    Geo g;
    int cdir;   // coarsening direction to the next level
    int relax;  // 0: skipped level (no smoothing, see set_relax)
    const T *c, *a0, *a1, *a2;
    const T *pm, *pp;  // transfer weights from the finer level (coarse-indexed)
    T *f, *v, *w;
};
template <typename T>
struct BotArgs {
    int nlev;  // levels lb..lb+nlev-1, the last one solved by the dense inverse
    int npre, npost;
    const double *Ainv;
    BotLevel<T> L[BOT_MAX_LEVELS];
};
template <typename T>
void launch_bottom(hipStream_t s, const BotArgs<T> &a);
void launch_strength(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2, double *part,
                     int *nparts);
// counts of the cells with two or more directions and of the anisotropic ones
// (k_aniso_count) as block partials: part[0..nparts), part[MAX_PARTS..)
void launch_aniso_count(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2,
                        double *part, int *nparts);
// The coarsest level's dense (pseudo-)inverse in one workgroup (levels of at
// most COARSE_LDS_MAX cells): out[0] = 1 when singular, out[1] = 1 when the
// Cholesky factorisation failed.
constexpr int COARSE_LDS_MAX = 128;
void launch_coarse_inverse(hipStream_t s, const Geo &g, const double *c, const double *a0, const double *a1,
                           const double *a2, double *Ainv, double *out);
void launch_to_f32(hipStream_t s, long long n, const double *src, float *dst);
void launch_to_f64(hipStream_t s, long long n, const float *src, double *dst);

// One multigrid level on the device.
template <typename T>
struct LevelData {
    T *c = nullptr, *a0 = nullptr, *a1 = nullptr, *a2 = nullptr;
    T *pm = nullptr, *pp = nullptr; // weights of the transfer from the finer level (coarse-indexed)
    T *f = nullptr, *v = nullptr;   // right-hand side and final iterate (levels >= 1)
    T *w = nullptr;                 // second iterate buffer (out-of-place sweeps), all levels
};

// The coarse level's weights (pm, pp), couplings and diagonal from the fine
// level along direction d; with c32, their FP32 copies in the same pass.
void launch_coarse_build(hipStream_t s, const Geo &F, const Geo &C, int d, const double *c, const double *a0,
                         const double *a1, const double *a2, double *pm, double *pp, double *cc, double *ca0,
                         double *ca1, double *ca2, double *sc, const LevelData<float> *c32);

struct Level {
    Geo g{};
    int cdir = -1;
    bool relax = true;              // smoothed in the V-cycle (skipped relaxation)
    double str[3] = {0.0, 0.0, 0.0}; // sum |a_d| per direction, from the coarsening choice
    LevelData<double> d;
    LevelData<float> s;
};

}  // namespace ngap
