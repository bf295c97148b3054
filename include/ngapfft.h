/* fourier3d on the GPU (libngapfft.so): the FFT direct solve of NGA2's
 * fourier3d linsol (fourier3d_class.f90) for a stencil operator that is the
 * same in every cell of a periodic, uniform grid.
 *
 * Method (fourier3d_class.f90:1103-1202):
 * - setup: the operator must be circulant, every interior cell's stencil
 *   within circ_tol of the first interior cell's (:1114-1126). The stencil is
 *   placed at its offsets modulo the grid (:1128-1137) and transformed
 *   (:1140); the zero mode is set to 1 (:1143-1144); any mode with
 *   |F| < 1000 eps is an error (:1146-1152); F is inverted (:1154-1160).
 * - solve: rhs is transformed, its zero mode set to 0 (:1039-1040), multiplied
 *   by 1/F, transformed back and divided by nx ny nz (:1047-1099); the real
 *   part is the interior of sol, whose ghosts are then filled (:1196).
 * The transforms are rocFFT real-to-complex/complex-to-real plans over the
 * directions with more than one cell; for a real stencil F(-k) = conj F(k), so
 * the half spectrum holds the whole of the reference's complex transform.
 *
 * Layout: opr(nst, imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_) and the
 * ghosted rhs/sol, as for ngapsolve.h. Return codes are those of ngapsolve.h.
 * The constructor's other checks (a uniform mesh, fourier3d_class.f90:146-148)
 * concern the caller's grid spacing, which this interface does not see; the
 * caller makes them. */
#ifndef NGAPFFT_H
#define NGAPFFT_H

#include "ngapsolve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ngap_fft ngap_fft;

/* fourier3d_from_args + fourier3d_init (fourier3d_class.f90:109-216):
 * stc(nst,3) column-major. NGAP_ERR_UNSUPPORTED when a direction with more
 * than one cell is not periodic (:146-148). */
int ngap_fft_create(ngap_fft **f, const ngap_grid *g, int nst, const int *stc);

/* fourier3d_setup (:1103-1166). opr: host array; circ_tol: the circulance
 * tolerance, which NGA2 takes as 6 eps / min_meshsize^4 (:1124).
 * NGAP_ERR_NOT_CIRCULANT or NGAP_ERR_SINGULAR where the reference dies. */
int ngap_fft_setup(ngap_fft *f, const double *opr, double circ_tol);
int ngap_fft_setup_device(ngap_fft *f, const double *d_opr, double circ_tol);

/* fourier3d_solve (:1170-1202): sol = A^-1 rhs with zero mean, ghosts filled.
 * rhs/sol: host arrays (ghosted). The ghosts of a non-periodic direction (one
 * cell thick grids only) are left as the caller had them. */
int ngap_fft_solve(ngap_fft *f, const double *rhs, double *sol);
int ngap_fft_solve_device(ngap_fft *f, const double *d_rhs, double *d_sol);

/* fourier3d_destroy (:1206-1221) and the object itself. */
int ngap_fft_destroy(ngap_fft *f);

/* Time of the last solve's kernels and transforms on the handle's stream
 * (HIP events), in ms; 0 before the first solve. */
int ngap_fft_last_ms(const ngap_fft *f, double *ms);

#ifdef __cplusplus
}
#endif

#endif
