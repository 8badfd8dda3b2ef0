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
#include <vector>

#include "kernels_common.h"
#include "kernels_fast.h"

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
   double* o;   // the transposed sweep's output cof L(W)/S (3 components on
                // pb), S = sx sy sz, on every point of the box
};
Box pad_box(const Box& b);
long adj_pad_doubles(const Box& b);                 // doubles AdjPad needs
AdjPad make_adj_pad(const Box& b, double* mem);     // carve AdjPad out of mem
int launch_adj_pad_material(const Box& b, const double* mu, const double* la, const double* strx, const double* stry,
                            const double* strz, const AdjPad& pad, hipStream_t s);

// out = L^T z (overwritten on every point); z is read on the rows only.
// pad: filled by launch_adj_pad_material for the same material.  weighted:
// pad.w already holds W = S z on the interior-formula rows
// (launch_vjp_corr_pw / launch_vjp_pred_pw).  out null: the result stays in
// pad.o on the padded box.
int launch_rhs4sg_T(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* mu,
                    const double* la, double h, const double* strx, const double* stry, const double* strz,
                    double* out, const AdjPad& pad, hipStream_t s, bool weighted = false);
// gmu += G_mu(z, u), gla += G_la(z, u) on every point.
int launch_rhs4sg_matgrad(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* u,
                          const double* mu, const double* la, double h, const double* strx, const double* stry,
                          const double* strz, double* gmu, double* gla, hipStream_t s);
// Supergrid damping up -= D(u - um; rho) at the interior points
// (addsgd4fort_indrev, ew-cfromfort.C:985-1041).  For the adjoint y of up:
// dbar = -D^T y (overwritten on every point), grho += d/drho [-y . D].
// inv_rho = false leaves out the derivative through the 1/rho factor at the
// damped point (y . D / rho), for callers that form it from state differences.
// LDS k-march over column tiles.
// tiles (damp_T_tiles of the same box and profiles): only those tiles run,
// dbar keeps its values elsewhere (where the transposed damping vanishes).
int launch_addsgd4_T_march(const Box& b, const double* y, const double* u, const double* um, const double* rho,
                           const SgCoef& sg, double beta, double* dbar, double* grho, hipStream_t s,
                           bool inv_rho = true, const DampTile* tiles = nullptr, int ntiles = 0);
// The column tiles and planes of the damping transpose whose outputs can be
// nonzero (the damping profiles at a point or a neighbour nonzero), from the
// host profiles (box-sized, from the box start).
std::vector<DampTile> damp_T_tiles(const Box& b, const double* dcx, const double* dcy, const double* dcz);
// Transpose of bcfortsg with zero boundary data on the adjoint y of its
// output (sides in reverse order): Dirichlet/supergrid ghosts are dropped, a
// free-surface ghost plane is pushed into the points its closure reads.  The
// closure's mu/lambda derivatives, evaluated on the field fwd the BC acted on,
// go to gmu, gla.  plane: scratch of (nx+4)(ny+4)*3 doubles.
int launch_bcfortsg_T(const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd, double* y,
                      const double* fwd, double h, const SbpTables* d_sbp, const double* mu, const double* la,
                      const double* strx, const double* stry, double* gmu, double* gla, double* plane,
                      hipStream_t s);

// Argument checks shared by sw4hip_bcfortsg_adjoint and its allocation-free
// variant (abi.cpp).
int check_bcfortsg_adjoint(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny, int nz,
                           const double* y, const double* fwd, const int* bccnd, const double* sbop, const double* mu,
                           const double* lambda, const double* strx, const double* stry, const double* gmu,
                           const double* glambda);

// pointwise helper over n values
int launch_axpy(long n, double a, const double* x, double* y, hipStream_t s);              // y += a x

// Fused pointwise stages of the reverse step (every point of the box unless
// said).  corrector: z = c4/rho y on the interior (0 elsewhere), the
// transposed operator's input W = S z on the interior-formula rows into
// pad.w (its pad must hold zeros), grho -= y.(u1 - up)/rho on the interior;
int launch_vjp_corr_pw(const Box& b, const RowMap& m, const double* strx, const double* stry, const double* strz,
                       double c4, const double* rho, const double* y, const double* u1, const double* up, double* z,
                       double* grho, const AdjPad& pad, hipStream_t s);
// predictor: z = dt2/rho yp, W likewise, grho -= dt2 yp.acc/rho;
int launch_vjp_pred_pw(const Box& b, const RowMap& m, const double* strx, const double* stry, const double* strz,
                       double dt2, const double* rho, const double* yp, const double* acc, double* z, double* grho,
                       const AdjPad& pad, hipStream_t s);
// y += pad.o/dt2 on the box (pad.o: the transposed operator's result);
int launch_vjp_axpy_pad(const Box& b, double dt2i, const AdjPad& pad, double* y, hipStream_t s);
// adjoints of U^n and U^{n-1} through one step, with a = p1.o/dt2 and o2 =
// p2.o (the two transposed operators' results, on the padded box):
// lb += db - 2a + 2yp + o2, lc = a - db - yp (lc is overwritten).
int launch_vjp_lam(const Box& b, double dt2i, const double* db, const AdjPad& p1, const double* yp,
                   const AdjPad& p2, double* lb, double* lc, hipStream_t s);
// Receiver traces g[s][r][c] (s = 0..nsteps) as adjoint sources, grouped by
// the grid value they act on: group q adds to lam[uidx[q]] the entries
// ent[off[q] .. off[q+1]-1] (e = 3r + c) in that order, so receivers that
// share a grid point sum deterministically.
struct ReceiverAdj {
   int nu;            // groups
   int nrec;
   const int* uidx;   // [nu] component offsets into the field
   const int* off;    // [nu + 1]
   const int* ent;    // [3 nrec]
   const int* mode;   // [nrec] 0 displacement, 1 velocity
   const double* g;   // [(nsteps + 1) 3 nrec]
};
// Adjoint source of state s: displacement receivers add g[s], velocity
// receivers (U^s - U^{s-2}) / (2 dt) add g[s]/(2 dt) - g[s+2]/(2 dt)
// (g[s+2] only for s + 2 <= nsteps); nothing for s outside 1..nsteps.
int launch_receiver_sources_T(const ReceiverAdj& R, int s, int nsteps, double twodt, double* lam, hipStream_t st);

}  // namespace sw4amd
