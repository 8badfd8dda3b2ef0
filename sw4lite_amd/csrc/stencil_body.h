// stencil_body.h -- the interior 4th-order divergence-of-stress stencil of
// rhs4sg_rev.C:111-347 written once over an accessor, so that the kernels
// reading global memory and the kernels reading LDS-staged planes execute the
// same arithmetic.
//
// Accessor S (R = double for the FP64 path, float for the FP32 form):
//   R u(int c, int di, int dj, int dk)  displacement component c at offset
//   R mu(int di, int dj, int dk)        only axis-aligned offsets are used
//   R la(int di, int dj, int dk)
//   R sx(int di), sy(int dj), sz(int dk)  supergrid stretching
// Every offset is a compile-time constant after unrolling.
#pragma once
#include "sw4hip_internal.h"

namespace sw4amd {

template <class R>
__device__ __forceinline__ R d4c(R fm2, R fm1, R fp1, R fp2)
{
   return fm2 - fp2 + 8 * (fp1 - fm1);
}

template <class R>
__device__ __forceinline__ void varcoef4c(const R a[5], R m[4])
{
   m[0] = a[1] - R(0.75) * (a[2] + a[0]);
   m[1] = a[0] + a[3] + 3 * (a[2] + a[1]);
   m[2] = a[1] + a[4] + 3 * (a[3] + a[2]);
   m[3] = a[3] - R(0.75) * (a[2] + a[4]);
}

template <class R>
__device__ __forceinline__ R varsum4c(const R m[4], const R f[5])
{
   return m[0] * (f[0] - f[2]) + m[1] * (f[1] - f[2]) + m[2] * (f[3] - f[2]) + m[3] * (f[4] - f[2]);
}

// (coef * D_DB u_C)_DA, centred 4th order in both directions, times 144h^2;
// coef = la when LA else mu, sampled along DA.
template <int DA, int DB, int C, bool LA, class R, class S>
__device__ __forceinline__ R mixed_term(const S& s)
{
   R g[4];
#pragma unroll
   for (int p = 0; p < 4; p++) {
      const int op = p < 2 ? p - 2 : p - 1;
      const int ai = DA == 0 ? op : 0, aj = DA == 1 ? op : 0, ak = DA == 2 ? op : 0;
      R fb[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int ot = t < 2 ? t - 2 : t - 1;
         fb[t] = s.u(C, ai + (DB == 0 ? ot : 0), aj + (DB == 1 ? ot : 0), ak + (DB == 2 ? ot : 0));
      }
      g[p] = (LA ? s.la(ai, aj, ak) : s.mu(ai, aj, ak)) * d4c(fb[0], fb[1], fb[2], fb[3]);
   }
   return d4c(g[0], g[1], g[2], g[3]);
}

// r = h^2 L(u) at one interior point (no 1/h^2 factor).
template <class R, class S>
__device__ __forceinline__ void rhs4_interior_body(const S& s, R r[3])
{
   R am[5], al[5], mx[4], lx[4], my[4], ly[4], mz[4], lz[4];
#pragma unroll
   for (int q = 0; q < 5; q++) {
      am[q] = s.mu(q - 2, 0, 0) * s.sx(q - 2);
      al[q] = s.la(q - 2, 0, 0) * s.sx(q - 2);
   }
   varcoef4c(am, mx);
   varcoef4c(al, lx);
#pragma unroll
   for (int q = 0; q < 5; q++) {
      am[q] = s.mu(0, q - 2, 0) * s.sy(q - 2);
      al[q] = s.la(0, q - 2, 0) * s.sy(q - 2);
   }
   varcoef4c(am, my);
   varcoef4c(al, ly);
#pragma unroll
   for (int q = 0; q < 5; q++) {
      am[q] = s.mu(0, 0, q - 2) * s.sz(q - 2);
      al[q] = s.la(0, 0, q - 2) * s.sz(q - 2);
   }
   varcoef4c(am, mz);
   varcoef4c(al, lz);
   R cx[4], cy[4], cz[4];
#pragma unroll
   for (int q = 0; q < 4; q++) {
      cx[q] = 2 * mx[q] + lx[q];
      cy[q] = 2 * my[q] + ly[q];
      cz[q] = 2 * mz[q] + lz[q];
   }
   const R sx = s.sx(0), sy = s.sy(0), sz = s.sz(0);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      R f[5], acc;
#pragma unroll
      for (int q = 0; q < 5; q++) f[q] = s.u(c, q - 2, 0, 0);
      acc = sx * varsum4c(c == 0 ? cx : mx, f);
#pragma unroll
      for (int q = 0; q < 5; q++) f[q] = s.u(c, 0, q - 2, 0);
      acc += sy * varsum4c(c == 1 ? cy : my, f);
#pragma unroll
      for (int q = 0; q < 5; q++) f[q] = s.u(c, 0, 0, q - 2);
      acc += sz * varsum4c(c == 2 ? cz : mz, f);
      r[c] = acc * R(1.0 / 6);
   }
   const R i144 = R(1.0 / 144);
   const R sxy = sx * sy * i144, sxz = sx * sz * i144, syz = sy * sz * i144;
   r[0] += sxy * mixed_term<0, 1, 1, true, R>(s) + sxz * mixed_term<0, 2, 2, true, R>(s) +
           sxy * mixed_term<1, 0, 1, false, R>(s) + sxz * mixed_term<2, 0, 2, false, R>(s);
   r[1] += sxy * mixed_term<0, 1, 0, false, R>(s) + sxy * mixed_term<1, 0, 0, true, R>(s) +
           syz * mixed_term<1, 2, 2, true, R>(s) + syz * mixed_term<2, 1, 2, false, R>(s);
   r[2] += sxz * mixed_term<0, 2, 0, false, R>(s) + syz * mixed_term<1, 2, 1, false, R>(s) +
           sxz * mixed_term<2, 0, 0, true, R>(s) + syz * mixed_term<2, 1, 1, true, R>(s);
}

}  // namespace sw4amd
