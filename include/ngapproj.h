/* Device-side projection around the pressure solve (part of libngapsolve.so).
 *
 * These are the callers' side of the solve, restated as HIP kernels over the
 * caller's own ghosted arrays in device memory:
 * - operator assembly: incomp_class.f90:766-787, and tpns%update_laplacian at
 *   tpns_class.f90:1565-1593;
 * - divergence: incomp_class.f90:1144-1159, tpns_class.f90:1597-1612;
 * - pressure gradient: incomp_class.f90:1163-1184, and with the two-phase
 *   jump tpns_class.f90:1695-1716;
 * - right-hand side scaling and velocity correction of the drivers:
 *   examples/channel/src/simulation.f90:318-328,
 *   examples/falling_drop/src/simulation.f90:302-311;
 * - flow-rate correction: incomp_class.f90:1577-1716, tpns_class.f90:2131-2267;
 * - surface-tension jump: tpns_class.f90:1616-1691;
 * - mean removal: incomp_class.f90:1720-1740, with config_class.f90:196-214;
 * - ghost fill: pgrid_class.f90:439-518, single rank.
 * With them, a time step keeps U, V, W, P, rho_U/V/W and opr on the GPU: the
 * host copies of rhs, sol and opr around each solve go away.
 *
 * Layout: every 3-D array is the Fortran array (imino_:imaxo_,
 * jmino_:jmaxo_, kmino_:kmaxo_), i fastest, as a flat device array of
 * nxo*nyo*nzo doubles. The metric arrays are NGA2's masked metrics:
 * - divp_x(0:1,i,j,k) and divu_x(-1:0,i,j,k) are pairs, the first index
 *   fastest (2*nxo*nyo*nzo doubles each), likewise for y and z;
 * - opr(7,i,j,k) in stencil order centre,+x,-x,+y,-y,+z,-z.
 *
 * Return codes are those of ngapsolve.h. */
#ifndef NGAPPROJ_H
#define NGAPPROJ_H

#include "ngapsolve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ngap_proj ngap_proj;

/* Upload the masked metrics, cell volumes and VF (host arrays, layout above)
 * and compute the fluid volume (config_class.f90:176-192). */
int ngap_proj_create(ngap_proj **p, const ngap_grid *g, const double *divp_x, const double *divp_y,
                     const double *divp_z, const double *divu_x, const double *divu_y, const double *divu_z,
                     const double *vol, const double *vf);
int ngap_proj_destroy(ngap_proj *p);

/* opr = -vol div((1/rho_face) grad .) over interior cells, ghost entries 0.
 * rho_u/v/w: device face densities, or all NULL for constant density
 * (incomp_class.f90:766-787; tpns_class.f90:1565-1593). */
int ngap_proj_laplacian(ngap_proj *p, const double *rho_u, const double *rho_v, const double *rho_w, double *opr);

/* div = divp_x U(i:i+1) + divp_y V(j:j+1) + divp_z W(k:k+1) over interior
 * cells, ghosts filled as cfg%sync does (incomp_class.f90:1144-1159). */
int ngap_proj_div(ngap_proj *p, const double *u, const double *v, const double *w, double *div);

/* (Px, Py, Pz) = divu_x P(i-1:i) etc. over interior cells, ghosts filled
 * (incomp_class.f90:1163-1184). */
int ngap_proj_pgrad(ngap_proj *p, const double *pres, double *px, double *py, double *pz);

/* P -= sum(vol VF P) / fluid_vol on interior cells with VF > 0, then the
 * ghosts are filled (incomp_class.f90:1720-1740). *mean (host, may be NULL)
 * gets the removed value. */
int ngap_proj_shift_p(ngap_proj *p, double *pres, double *mean);

/* Ghost fill of a device array: periodic wrap in x, then y, then z;
 * non-periodic ghosts untouched (pgrid_class.f90:459-516 on one rank). */
int ngap_proj_sync(ngap_proj *p, double *a);

/* Right-hand side of the pressure solve over every cell of the ghosted
 * arrays: rhs = -(((vol * div) * rho) / dt). Replaces the drivers'
 * `psolv%rhs=-cfg%vol*div*rho/dt` (examples/channel/src/simulation.f90:318,
 * constant density rho) and `psolv%rhs=-cfg%vol*div/dt`
 * (examples/falling_drop/src/simulation.f90:302: pass rho = 1, which is
 * exact). dt must be finite and nonzero. */
int ngap_proj_rhs(ngap_proj *p, const double *div, double rho, double dt, double *rhs);

/* Velocity correction after the solve, over every cell:
 *   P += sol                      (when pres and sol are both non-NULL)
 *   U -= (dt * Px) / rho_U        (likewise V, W)
 * rho_u/v/w: device face densities (falling_drop/src/simulation.f90:307-311),
 * or all NULL for the constant density rho (channel/src/simulation.f90:
 * 324-328). */
int ngap_proj_correct(ngap_proj *p, double dt, double rho, const double *rho_u, const double *rho_v,
                      const double *rho_w, const double *px, const double *py, const double *pz, double *u, double *v,
                      double *w, double *pres, const double *sol);

/* Two-phase pressure gradient: (Px, Py, Pz) = divu P - dPj over interior
 * cells, ghosts filled (tpns_class.f90:1695-1716). dpjx/y/z are the jump
 * increments written by ngap_proj_csf_jump. */
int ngap_proj_pgrad_jump(ngap_proj *p, const double *pres, const double *dpjx, const double *dpjy,
                         const double *dpjz, double *px, double *py, double *pz);

/* Surface-tension jump by CSF (tpns%add_surface_tension_jump,
 * tpns_class.f90:1616-1691, without a contact model). On the faces of cells
 * imin_..imax_+1 (likewise j, k) the jump is
 *   Pjx = sigma * curv_face * (divu_x(-1) VF(i-1) + divu_x(0) VF(i)),
 * curv_face the SD*vol-weighted mean curvature of the two cells (0 where
 * their SD*vol sum is not positive). pjx/y/z hold the previous jump on entry
 * and the new one on exit; dpjx/y/z = new - previous over every cell; then
 * div += dt * (sum(divp_x DPjx/rho_U) + ...) on interior cells. sd, curv and
 * vf are the VOF surface density, curvature and volume fraction (ghosted);
 * rho_u/v/w the face densities. */
int ngap_proj_csf_jump(ngap_proj *p, double sigma, double dt, const double *sd, const double *curv, const double *vf,
                       const double *rho_u, const double *rho_v, const double *rho_w, double *pjx, double *pjy,
                       double *pjz, double *dpjx, double *dpjy, double *dpjz, double *div);

/* One boundary condition for correct_mfr, as NGA2's bcond type holds it
 * (face, rdir, canCorrect, and the iterator's cell map): face 0/1/2 for
 * 'x'/'y'/'z', rdir +1 or -1, ijk the Fortran (i,j,k) of n_all cells of
 * which the first n_in are the iterator's interior cells (itr%n_) and the
 * rest its overlap cells (itr%no_). */
typedef struct ngap_bcond {
    int face;
    int rdir;
    int can_correct;
    int n_in;
    int n_all;
    const int *ijk;
} ngap_bcond;

/* Upload the boundary conditions (copied; replaces any previous set) with
 * the cell sizes dx(imino_:imaxo_), dy(jmino_:jmaxo_), dz(kmino_:kmaxo_) of
 * the mesh, and compute each one's area and the correctable area
 * (incomp_class.f90:1577-1665). nbc = 0 removes them. */
int ngap_proj_set_bconds(ngap_proj *p, int nbc, const ngap_bcond *bc, const double *dx, const double *dy,
                         const double *dz);

/* correct_mfr (incomp_class.f90:1669-1716 with the constant density rho;
 * tpns_class.f90:2219-2267 with rho = 1): the flow rate through each
 * boundary condition, mfr_error = their sum, and unless |mfr_error| or the
 * correctable area is below 10 eps, U += rdir * (-mfr_error / (rho * area))
 * on every cell of each correctable boundary condition (likewise V, W by
 * face). Everything stays on the device; *mfr_error (host, may be NULL)
 * gets the mismatch. Without boundary conditions it does nothing. */
int ngap_proj_correct_mfr(ngap_proj *p, double rho, double *u, double *v, double *w, double *mfr_error);

/* ngap_proj_shift_p without a handle and without allocating: no counterpart
 * in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight time-step loops, harnesses that forbid
 * hipMalloc). vol and vf are the device arrays cfg%vol and cfg%VF, pres the
 * device pressure (all ghosted, layout above). It removes
 * sum(vol VF P) / sum(vol VF) over interior cells from the interior cells
 * with VF > 0 and fills the ghosts, with the same sums and the same result
 * bits as ngap_proj_create + ngap_proj_shift_p (incomp_class.f90:1720-1740,
 * config_class.f90:176-214). d_mean (device, may be NULL) gets the removed
 * value; a fluid volume that is not positive leaves pres unchanged and the
 * mean NaN. The partial sums live in a fixed 16.4 KB device array of the
 * library (one per device, held until the process ends), whatever the grid
 * size; grids of 2^40/7 cells or more are refused (NGAP_ERR_ARG). stream:
 * a hipStream_t of the library's HIP runtime on which the work is enqueued
 * (the call then returns without waiting), or NULL for the default stream,
 * which the call synchronizes before returning. Calls on one device share
 * the array: calls that overlap in time on one device race. */
int ngap_proj_shift_p_noalloc(const ngap_grid *g, const double *vol, const double *vf, double *pres, double *d_mean,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif
