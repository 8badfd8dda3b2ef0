// kernels_adj.h -- transposed (adjoint) kernels of the FP64 time step and the
// derivatives of its linear maps with respect to the material
// (kernels_adj.hip).  Component-blocked layout (corder = 1) only.
//
// Notation: L(u; mu, la) is rhs4sg (rhs4sg_rev.C:43-864) on its rows (every
// interior point: the interior formula on k1..k2, the SBP formula on the six
// rows next to a free surface).  For an adjoint field z given on the rows,
//   L^T z          the transpose, on every point of the box (ghosts included),
//   G_mu(z, u)     d/dmu [z . L(u; mu, la)] on every point (L is linear in
//   G_la(z, u)     mu and la, so these are exact bilinear forms).
#pragma once
#include "kernels_common.h"

namespace sw4amd {

// Rows of L: interior formula on k1..k2, SBP rows at the top (k = 1..6) and
// bottom (k = nz-5..nz) when one-sided there.
struct RowMap {
   int nx, ny, nz;
   int k1, k2;
   int top, bot;
};
RowMap make_rowmap(const Box& b, int nz, const int* onesided);

// Copies of the material and stretching on the box grown by two points per
// side (mu = lambda = 0, stretching 1 in the pad) and a padded work field, so
// that the transposed stencils read without bounds checks.
struct AdjPad {
   Box pb;      // the padded box
   double* w;   // 3 components on pb
   double* mu;  // on pb
   double* la;
   double* sx;  // pb.ie - pb.ib + 1 values, and so on
   double* sy;
   double* sz;
   double* wb;  // W on the box (3 components), 2W, and S = strx stry strz:
   double* w2b; // the inputs of the LDS predictor sweep that evaluates
   double* sb;  // cof L(W)/S = 2W - 2W + cof L(W)/S with rho := S, dt = 1
   Box pb4;     // the box grown by four points per side
   double* u4;  // u on pb4 (3 components, 0 outside the box)
   double* z4;  // z on the interior-formula rows (0 elsewhere), on pb4
   double* gy;  // damping transpose on pb: beta y / rho on the interior (3 components),
   double* dd;  // u - um in the box (3 components),
   double* rh;  // rho in the box (0 in the pad)
};
Box pad_box(const Box& b);
long adj_pad_doubles(const Box& b);                 // doubles AdjPad needs
AdjPad make_adj_pad(const Box& b, double* mem);     // carve AdjPad out of mem
int launch_adj_pad_material(const Box& b, const double* mu, const double* la, const double* strx, const double* stry,
                            const double* strz, const AdjPad& pad, hipStream_t s);

// out = L^T z (overwritten on every point); z is read on the rows only.
// pad: filled by launch_adj_pad_material for the same material.
int launch_rhs4sg_T(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* mu,
                    const double* la, double h, const double* strx, const double* stry, const double* strz,
                    double* out, const AdjPad& pad, hipStream_t s);
// gmu += G_mu(z, u), gla += G_la(z, u) on every point (pad as above).
int launch_rhs4sg_matgrad(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* u,
                          const double* mu, const double* la, double h, const double* strx, const double* stry,
                          const double* strz, double* gmu, double* gla, const AdjPad& pad, hipStream_t s);
// Supergrid damping up -= D(u - um; rho) at the interior points
// (addsgd4fort_indrev, ew-cfromfort.C:985-1041).  For the adjoint y of up:
// dbar = -D^T y (overwritten on every point), grho += d/drho [-y . D].
// inv_rho = false leaves out the derivative through the 1/rho factor at the
// damped point (y . D / rho), for callers that form it from state differences.
int launch_addsgd4_T(const Box& b, const double* y, const double* u, const double* um, const double* rho,
                     const SgCoef& sg, double beta, double* dbar, double* grho, hipStream_t s, bool inv_rho = true);
// As launch_addsgd4_T with unchecked reads of padded fields (pad as above).
int launch_addsgd4_T_pad(const Box& b, const double* y, const double* u, const double* um, const double* rho,
                         const SgCoef& sg, double beta, double* dbar, double* grho, const AdjPad& pad, hipStream_t s,
                         bool inv_rho = true);
// Transpose of bcfortsg with zero boundary data on the adjoint y of its
// output (sides in reverse order): Dirichlet/supergrid ghosts are dropped, a
// free-surface ghost plane is pushed into the points its closure reads.  The
// closure's mu/lambda derivatives, evaluated on the field fwd the BC acted on,
// go to gmu, gla.  plane: scratch of (nx+4)(ny+4)*3 doubles.
int launch_bcfortsg_T(const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd, double* y,
                      const double* fwd, double h, const SbpTables* d_sbp, const double* mu, const double* la,
                      const double* strx, const double* stry, double* gmu, double* gla, double* plane,
                      hipStream_t s);

// pointwise helpers over n values
int launch_axpy(long n, double a, const double* x, double* y, hipStream_t s);              // y += a x
int launch_rows_scale(const Box& b, double coef, const double* rho, const double* y, double* z,
                      hipStream_t s);                                                        // z = coef/rho y on the interior, 0 elsewhere
int launch_rho_grad(const Box& b, double coef, const double* y, const double* lu, const double* f,
                    const double* rho, double* grho, hipStream_t s);                        // grho -= coef y.(lu+f)/rho^2
// grho -= coef y.(a - b)/rho on the interior (b may be null)
int launch_rho_diff(const Box& b, double coef, const double* y, const double* a, const double* bsub, const double* rho,
                    double* grho, hipStream_t s);
int launch_scatter_add(int n, const int* idx, const double* val, double* y, hipStream_t s);  // y[idx] += val

}  // namespace sw4amd
