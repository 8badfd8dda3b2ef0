// sbp_row_body.h -- one SBP boundary row (k = 1..6 below a free surface, or
// the six rows above one at the bottom) of h^2 L(u), written over a row-space
// accessor so that the FP64 and the FP32 slab kernels share the arithmetic.
//
// Accessor S (R = double or float): u(c, di, dj, q), mu(di, dj, q),
// la(di, dj, q) at row-space plane q (plane q is grid plane
// kbase + sgn (q - 1), rhs4sg_rev.C:353-629; row kr lives on q = kr), and the
// stretching sx(di), sy(dj).
#pragma once
#include "kernels_common.h"

namespace sw4amd {

// SBP row kr of h^2 L(u): the same terms as rhs_sbp_point (kernels_common.h),
// evaluated through the row-space accessor.
template <class R, class S>
__device__ __forceinline__ void sbp_row_body(const S& s, int kr, R sgn, const SbpTablesT<R>* __restrict__ sbp, R r[3])
{
   const R i12 = R(1.0 / 12), i144 = R(1.0 / 144);
   const int off[4] = {-2, -1, 1, 2};
   const R sx = s.sx(0), sy = s.sy(0);
   // in-plane second derivatives (x and y), 1/6 included
   {
      R am[5], al[5], mx[4], lx[4], my[4], ly[4], f[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
         am[t] = s.mu(t - 2, 0, kr) * s.sx(t - 2);
         al[t] = s.la(t - 2, 0, kr) * s.sx(t - 2);
      }
      varcoef4(am, mx);
      varcoef4(al, lx);
#pragma unroll
      for (int t = 0; t < 5; t++) {
         am[t] = s.mu(0, t - 2, kr) * s.sy(t - 2);
         al[t] = s.la(0, t - 2, kr) * s.sy(t - 2);
      }
      varcoef4(am, my);
      varcoef4(al, ly);
      R cx[4], cy[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
         cx[q] = 2 * mx[q] + lx[q];
         cy[q] = 2 * my[q] + ly[q];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
#pragma unroll
         for (int t = 0; t < 5; t++) f[t] = s.u(c, t - 2, 0, kr);
         R acc = sx * varsum4(c == 0 ? cx : mx, f);
#pragma unroll
         for (int t = 0; t < 5; t++) f[t] = s.u(c, 0, t - 2, kr);
         acc += sy * varsum4(c == 1 ? cy : my, f);
         r[c] = acc * R(1.0 / 6);
      }
   }
   // k-direction SBP second derivative and the ghost-point term
   R zz0 = 0, zz1 = 0, zz2 = 0;
#pragma unroll 1
   for (int q = 1; q <= 8; q++) {
      R cmu = 0, c2ml = 0;
#pragma unroll 2
      for (int m = 1; m <= 8; m++) {
         const R a = sbp->acof[(kr - 1) + 6 * (q - 1) + 48 * (m - 1)];
         cmu += a * s.mu(0, 0, m);
         c2ml += a * (2 * s.mu(0, 0, m) + s.la(0, 0, m));
      }
      zz0 += cmu * s.u(0, 0, 0, q);
      zz1 += cmu * s.u(1, 0, 0, q);
      zz2 += c2ml * s.u(2, 0, 0, q);
   }
   {
      const R gh = sbp->ghcof[kr - 1];
      const R m1 = s.mu(0, 0, 1), l1 = s.la(0, 0, 1);
      r[0] += zz0 + gh * m1 * s.u(0, 0, 0, 0);
      r[1] += zz1 + gh * m1 * s.u(1, 0, 0, 0);
      r[2] += zz2 + gh * (2 * m1 + l1) * s.u(2, 0, 0, 0);
   }
   // in-plane mixed derivatives (coef sampled along the outer direction)
   auto mixed_xy = [&](bool coef_la, int c, bool outer_x) {
      R g[4];
#pragma unroll
      for (int a = 0; a < 4; a++) {
         const int oi = outer_x ? off[a] : 0, oj = outer_x ? 0 : off[a];
         R fb[4];
#pragma unroll
         for (int t = 0; t < 4; t++) fb[t] = s.u(c, oi + (outer_x ? 0 : off[t]), oj + (outer_x ? off[t] : 0), kr);
         g[a] = (coef_la ? s.la(oi, oj, kr) : s.mu(oi, oj, kr)) * d4(fb[0], fb[1], fb[2], fb[3]);
      }
      return d4(g[0], g[1], g[2], g[3]);
   };
   // mixed(coef, u, c, da, db): outer derivative along da, inner along db
   r[0] += sx * sy * i144 * (mixed_xy(true, 1, true) + mixed_xy(false, 1, false));
   r[1] += sx * sy * i144 * (mixed_xy(false, 0, true) + mixed_xy(true, 0, false));
   // (la w_z)_x, (la w_z)_y, (mu u_z)_x, (mu v_z)_y with the SBP first derivative
   R w3x[4], w3y[4], w1x[4], w2y[4];
#pragma unroll
   for (int a = 0; a < 4; a++) {
      R a3x = 0, a3y = 0, a1x = 0, a2y = 0;
#pragma unroll 2
      for (int q = 1; q <= 8; q++) {
         const R bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
         a3x += bq * s.u(2, off[a], 0, q);
         a3y += bq * s.u(2, 0, off[a], q);
         a1x += bq * s.u(0, off[a], 0, q);
         a2y += bq * s.u(1, 0, off[a], q);
      }
      w3x[a] = s.la(off[a], 0, kr) * a3x;
      w3y[a] = s.la(0, off[a], kr) * a3y;
      w1x[a] = s.mu(off[a], 0, kr) * a1x;
      w2y[a] = s.mu(0, off[a], kr) * a2y;
   }
   // (mu w_x)_z, (mu w_y)_z, (la u_x)_z, (la v_y)_z with the SBP derivative
   R t3x = 0, t3y = 0, t1x = 0, t2y = 0;
#pragma unroll 1
   for (int q = 1; q <= 8; q++) {
      const R bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
      const R dx3 = d4(s.u(2, -2, 0, q), s.u(2, -1, 0, q), s.u(2, 1, 0, q), s.u(2, 2, 0, q));
      const R dy3 = d4(s.u(2, 0, -2, q), s.u(2, 0, -1, q), s.u(2, 0, 1, q), s.u(2, 0, 2, q));
      const R dx1 = d4(s.u(0, -2, 0, q), s.u(0, -1, 0, q), s.u(0, 1, 0, q), s.u(0, 2, 0, q));
      const R dy2 = d4(s.u(1, 0, -2, q), s.u(1, 0, -1, q), s.u(1, 0, 1, q), s.u(1, 0, 2, q));
      t3x += bq * (s.mu(0, 0, q) * i12 * dx3);
      t3y += bq * (s.mu(0, 0, q) * i12 * dy3);
      t1x += bq * (s.la(0, 0, q) * i12 * dx1);
      t2y += bq * (s.la(0, 0, q) * i12 * dy2);
   }
   r[0] += sx * (i12 * d4(w3x[0], w3x[1], w3x[2], w3x[3]));
   r[0] += sx * t3x;
   r[1] += sy * (i12 * d4(w3y[0], w3y[1], w3y[2], w3y[3]));
   r[1] += sy * t3y;
   r[2] += sx * (i12 * d4(w1x[0], w1x[1], w1x[2], w1x[3]));
   r[2] += sy * (i12 * d4(w2y[0], w2y[1], w2y[2], w2y[3]));
   r[2] += sx * t1x;
   r[2] += sy * t2y;
}

}  // namespace sw4amd
