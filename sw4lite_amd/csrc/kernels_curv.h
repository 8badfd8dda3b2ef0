// kernels_curv.h -- launchers of the curvilinear-grid kernels (kernels_curv.hip).
#pragma once
#include "kernels_common.h"

namespace sw4amd {

// met (4 components, layout co), jac = h^2 z_r from the grid depth z over the
// box; bad[0] is set to 1 where z_r <= 0 (the caller clears it first).
int launch_metric(int co, const Box& b, double h, const double* z, double* met, double* jac, int* bad,
                  hipStream_t s);
// lu = L(u) on the curvilinear grid: rows 1..6 with the SBP closure when
// onesided[4] == 1, then k up to ke-2.
int launch_rhs4sgcurv(int co, const Box& b, const int* onesided, const SbpTables* d_sbp, const double* u,
                      const double* mu, const double* la, const double* met, const double* jac, const double* strx,
                      const double* stry, double* lu, hipStream_t s);
// up -= x/y supergrid damping with rho -> rho jac, k over [kb+2, ke-2].
int launch_addsgd4c(int co, const Box& b, double* up, const double* u, const double* um, const double* rho,
                    const double* jac, const SgCoef& sg, double beta, hipStream_t s);
// Ghost points of side 4 (k = 0 from k = 1..4) or 5 (k = nz+1 from nz..nz-3)
// for zero traction minus forcing (3 interleaved values per (i,j) of the box
// plane, or NULL for none).
// Interface of a curvilinear grid stacked on a Cartesian one
// (EW::enforceCartTopo, EW.C:3503-3531), over the common (i, j) range
// [ib, ie] x [jb, je]: the Cartesian planes kbc .. kbc+g-1 (its ghosts above
// the interface) take the curvilinear planes kec-2g .. kec-g-1, then the
// curvilinear planes kec-g .. kec (the common plane and the ghosts below it)
// take the Cartesian planes kbc+2g .. kbc+g, with g ghost points.  cart and
// curv are the two grids' boxes (same i, j bounds).
int launch_enforce_cart_topo(int co, const Box& cart, double* ucart, const Box& curv, double* ucurv, int g,
                             hipStream_t s);
int launch_freesurf_curv(int co, const Box& b, int nz, int side, double* u, const double* mu, const double* la,
                         const double* met, const SbpTables* d_sbp, const double* forcing, const double* strx,
                         const double* stry, hipStream_t s);

// max over the box of -lambda_min(A)/(4 rho), A the principal-part matrix
// of the curvilinear operator (EW.C:5063-5136); npartial blocks, result in
// out[0].
int launch_maxeig_curv(int co, const Box& b, const double* met, const double* jac, const double* mu, const double* la,
                       const double* rho, double* partial, int npartial, double* out, hipStream_t s);

}  // namespace sw4amd
