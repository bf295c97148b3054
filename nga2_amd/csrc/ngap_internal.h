// Internal declarations shared by the solver's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ngap_kernels.h"
#include "ngapsolve.h"

namespace ngap {

int grid_for(long long n);

template <typename T>
void launch_apply(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *x, T *y,
                  double *part);
template <typename T>
void launch_rbgs(hipStream_t s, const Geo &g, const T *c, const T *a0, const T *a1, const T *a2, const T *f, T *v,
                 int col, int zero_guess);
template <typename T>
void launch_resid_restrict(hipStream_t s, const Geo &F, const Geo &C, int d, const T *c, const T *a0, const T *a1,
                           const T *a2, const T *f, const T *v, const T *pm, const T *pp, T *fc);
template <typename T>
void launch_prolong(hipStream_t s, const Geo &F, const Geo &C, int d, const T *pm, const T *pp, const T *vc, T *v);
template <typename Tin, typename Tout>
void launch_dense(hipStream_t s, int n, const double *Ainv, const Tin *f, Tout *v);
void launch_coarse_build(hipStream_t s, const Geo &F, const Geo &C, int d, const double *c, const double *a0,
                         const double *a1, const double *a2, double *pm, double *pp, double *cc, double *ca0,
                         double *ca1, double *ca2, double *sc);
void launch_strength(hipStream_t s, const Geo &g, const double *a0, const double *a1, const double *a2, double *part,
                     int *nparts);
void launch_to_f32(hipStream_t s, long long n, const double *src, float *dst);
void launch_to_f64(hipStream_t s, long long n, const float *src, double *dst);

// One multigrid level on the device.
template <typename T>
struct LevelData {
    T *c = nullptr, *a0 = nullptr, *a1 = nullptr, *a2 = nullptr;
    T *pm = nullptr, *pp = nullptr; // weights of the transfer from the finer level (coarse-indexed)
    T *f = nullptr, *v = nullptr;   // work vectors (levels >= 1)
};

struct Level {
    Geo g{};
    int cdir = -1;
    LevelData<double> d;
    LevelData<float> s;
};

}  // namespace ngap
