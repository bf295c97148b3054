/* Device-side projection around the pressure solve (part of libngapsolve.so).
 *
 * These are the callers' side of the solve, restated as HIP kernels over the
 * caller's own ghosted arrays in device memory:
 * - operator assembly: incomp_class.f90:766-787, and tpns%update_laplacian at
 *   tpns_class.f90:1565-1593;
 * - divergence: incomp_class.f90:1144-1159, tpns_class.f90:1597-1612;
 * - pressure gradient: incomp_class.f90:1163-1184;
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

#ifdef __cplusplus
}
#endif

#endif
