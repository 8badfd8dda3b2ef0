// kernels_common.h -- device-side building blocks shared by the one-thread-
// per-point kernels and the fused step kernels.  Each routine restates the
// reference arithmetic of one stencil (file:line cited) for a single grid
// point; layouts are selected at compile time by CO.
#pragma once
#include "stencil_body.h"
#include "sw4hip_internal.h"

namespace sw4amd {

// 1-D supergrid arrays (dc, str, corner) of the three axes, each indexed
// from the first index of its axis (EW.C:4671-4801).
struct SgCoef {
   const double *dcx, *dcy, *dcz;
   const double *strx, *stry, *strz;
   const double *cox, *coy, *coz;
};

// Sorted grid-point sources (EW.C:6370-6433): sources first[r]..first[r+1]-1
// act on the same grid point, so one thread owns one unique point.
struct SourceTable {
   int nunique;
   const int* first;     // nunique+1
   const int* ijk;       // 3 per source
   const double* force;  // 3 per source
   const double* freq;
   const double* t0;
   const int* tfkind;    // 0 Ricker, 1 Gaussian, 2 C6SmoothBump
};


// Read accessors for the displacement-like field a stencil is applied to.
// Plain: a stored array.  Accel: (up - 2u + um)/dt^2 formed on the fly
// (dpdmtfort, ew-cfromfort.C:116-141), so the corrector never stores Uacc.
struct Plain {
   const double* __restrict__ p;
   __device__ __forceinline__ double operator[](long q) const { return p[q]; }
};
struct Accel {
   const double* __restrict__ up;
   const double* __restrict__ u;
   const double* __restrict__ um;
   double dt2i;
   __device__ __forceinline__ double operator[](long q) const { return dt2i * (up[q] - 2 * u[q] + um[q]); }
};
// A few k-planes ks..ks+npl-1 of a component-blocked array over the full
// (i,j) extent, addressed with the full box's component-blocked indices.
struct Slab {
   const double* __restrict__ p;
   long npts;   // points of the full box
   long shift;  // nij * (ks - kb)
   long nslab;  // nij * npl
   __device__ __forceinline__ double operator[](long q) const
   {
      const int c = (q >= npts) + (q >= 2 * npts);
      return p[q - c * npts - shift + c * nslab];
   }
};

// 4th-order centred first difference times 12h: f(-2)-f(+2)+8(f(+1)-f(-1))
template <class R>
__device__ __forceinline__ R d4(R fm2, R fm1, R fp1, R fp2)
{
   return fm2 - fp2 + 8 * (fp1 - fm1);
}

// Coefficients of the variable-coefficient second difference (a u_x)_x on
// a 5-point window of a = coefficient*stretching (rhs4sg_rev.C:120-145).
template <class R>
__device__ __forceinline__ void varcoef4(const R a[5], R m[4])
{
   m[0] = a[1] - R(0.75) * (a[2] + a[0]);
   m[1] = a[0] + a[3] + 3 * (a[2] + a[1]);
   m[2] = a[1] + a[4] + 3 * (a[3] + a[2]);
   m[3] = a[3] - R(0.75) * (a[2] + a[4]);
}

template <class R>
__device__ __forceinline__ R varsum4(const R m[4], const R f[5])
{
   return m[0] * (f[0] - f[2]) + m[1] * (f[1] - f[2]) + m[2] * (f[3] - f[2]) + m[3] * (f[4] - f[2]);
}

// x/y (and z when WITHZ) second-derivative parts of the three components,
// including the 1/6 factor (rhs4sg_rev.C:120-221).
template <int CO, bool WITHZ, class UA>
__device__ __forceinline__ void second_derivs(const Box& b, int i, int j, int k, const UA& u,
                                              const double* __restrict__ mu,
                                              const double* __restrict__ la,
                                              const double* __restrict__ strx,
                                              const double* __restrict__ stry,
                                              const double* __restrict__ strz, double r[3])
{
   double am[5], al[5], mx[4], lx[4], my[4], ly[4], mz[4], lz[4];
   double f[5];
#pragma unroll
   for (int s = 0; s < 5; s++) {
      const double sx = strx[i + s - 2 - b.ib];
      am[s] = mu[sidx(b, i + s - 2, j, k)] * sx;
      al[s] = la[sidx(b, i + s - 2, j, k)] * sx;
   }
   varcoef4(am, mx);
   varcoef4(al, lx);
#pragma unroll
   for (int s = 0; s < 5; s++) {
      const double sy = stry[j + s - 2 - b.jb];
      am[s] = mu[sidx(b, i, j + s - 2, k)] * sy;
      al[s] = la[sidx(b, i, j + s - 2, k)] * sy;
   }
   varcoef4(am, my);
   varcoef4(al, ly);
   if (WITHZ) {
#pragma unroll
      for (int s = 0; s < 5; s++) {
         const double sz = strz[k + s - 2 - b.kb];
         am[s] = mu[sidx(b, i, j, k + s - 2)] * sz;
         al[s] = la[sidx(b, i, j, k + s - 2)] * sz;
      }
      varcoef4(am, mz);
      varcoef4(al, lz);
   }
   double cx[4], cy[4], cz[4];
#pragma unroll
   for (int q = 0; q < 4; q++) {
      cx[q] = 2 * mx[q] + lx[q];
      cy[q] = 2 * my[q] + ly[q];
      cz[q] = WITHZ ? 2 * mz[q] + lz[q] : 0.0;
   }
   const double sx = strx[i - b.ib], sy = stry[j - b.jb];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double acc;
#pragma unroll
      for (int s = 0; s < 5; s++) f[s] = u[vidx<CO>(b, c, i + s - 2, j, k)];
      acc = sx * varsum4(c == 0 ? cx : mx, f);
#pragma unroll
      for (int s = 0; s < 5; s++) f[s] = u[vidx<CO>(b, c, i, j + s - 2, k)];
      acc += sy * varsum4(c == 1 ? cy : my, f);
      if (WITHZ) {
#pragma unroll
         for (int s = 0; s < 5; s++) f[s] = u[vidx<CO>(b, c, i, j, k + s - 2)];
         acc += strz[k - b.kb] * varsum4(c == 2 ? cz : mz, f);
      }
      r[c] = acc * (1.0 / 6);
   }
}

// (coef * D_db u_c)_da, centred 4th order in both directions, times 144h^2.
template <int CO, class UA>
__device__ __forceinline__ double mixed(const Box& b, int i, int j, int k,
                                        const double* __restrict__ coef, const UA& u, int c,
                                        int da, int db)
{
   double g[4];
   const int off[4] = {-2, -1, 1, 2};
#pragma unroll
   for (int s = 0; s < 4; s++) {
      const int ia = i + (da == 0) * off[s], ja = j + (da == 1) * off[s], ka = k + (da == 2) * off[s];
      double fb[4];
#pragma unroll
      for (int t = 0; t < 4; t++)
         fb[t] = u[vidx<CO>(b, c, ia + (db == 0) * off[t], ja + (db == 1) * off[t], ka + (db == 2) * off[t])];
      g[s] = coef[sidx(b, ia, ja, ka)] * d4(fb[0], fb[1], fb[2], fb[3]);
   }
   return d4(g[0], g[1], g[2], g[3]);
}

// Global-memory accessor of rhs4_interior_body (stencil_body.h).
template <int CO, class UA>
struct GlobalSten {
   const Box& b;
   int i, j, k;
   const UA& uu;
   const double* __restrict__ mu_;
   const double* __restrict__ la_;
   const double* __restrict__ strx_;
   const double* __restrict__ stry_;
   const double* __restrict__ strz_;
   __device__ __forceinline__ double u(int c, int di, int dj, int dk) const
   {
      return uu[vidx<CO>(b, c, i + di, j + dj, k + dk)];
   }
   __device__ __forceinline__ double mu(int di, int dj, int dk) const { return mu_[sidx(b, i + di, j + dj, k + dk)]; }
   __device__ __forceinline__ double la(int di, int dj, int dk) const { return la_[sidx(b, i + di, j + dj, k + dk)]; }
   __device__ __forceinline__ double sx(int d) const { return strx_[i + d - b.ib]; }
   __device__ __forceinline__ double sy(int d) const { return stry_[j + d - b.jb]; }
   __device__ __forceinline__ double sz(int d) const { return strz_[k + d - b.kb]; }
};

// Interior point of L(u) without the 1/h^2 factor (rhs4sg_rev.C:111-347).
template <int CO, class UA>
__device__ __forceinline__ void rhs_interior_point(const Box& b, int i, int j, int k, const UA& u,
                                                   const double* __restrict__ mu,
                                                   const double* __restrict__ la,
                                                   const double* __restrict__ strx,
                                                   const double* __restrict__ stry,
                                                   const double* __restrict__ strz, double r[3])
{
   GlobalSten<CO, UA> s{b, i, j, k, u, mu, la, strx, stry, strz};
   rhs4_interior_body(s, r);
}

// SBP boundary row kr (1..6) at grid plane k; kbase/sgn map boundary-relative
// index m to grid plane kbase + sgn (m-1): top rhs4sg_rev.C:349-601,
// bottom rhs4sg_rev.C:602-863.  No 1/h^2 factor.
template <int CO, class UA>
__device__ __forceinline__ void rhs_sbp_point(const Box& b, int i, int j, int k, int kr, int kbase,
                                              int sgn, const SbpTables* __restrict__ sbp, const UA& u,
                                              const double* __restrict__ mu,
                                              const double* __restrict__ la,
                                              const double* __restrict__ strx,
                                              const double* __restrict__ stry, double r[3])
{
#define KK(m) (kbase + sgn * ((m)-1))
   const double i12 = 1.0 / 12, i144 = 1.0 / 144;
   const double sx = strx[i - b.ib], sy = stry[j - b.jb];
   const int off[4] = {-2, -1, 1, 2};
   second_derivs<CO, false, UA>(b, i, j, k, u, mu, la, strx, stry, nullptr, r);

   double zz0 = 0, zz1 = 0, zz2 = 0;
#pragma unroll 1
   for (int q = 1; q <= 8; q++) {
      double cmu = 0, c2ml = 0;
#pragma unroll
      for (int m = 1; m <= 8; m++) {
         const double a = sbp->acof[(kr - 1) + 6 * (q - 1) + 48 * (m - 1)];
         const long pm = sidx(b, i, j, KK(m));
         cmu += a * mu[pm];
         c2ml += a * (2 * mu[pm] + la[pm]);
      }
      zz0 += cmu * u[vidx<CO>(b, 0, i, j, KK(q))];
      zz1 += cmu * u[vidx<CO>(b, 1, i, j, KK(q))];
      zz2 += c2ml * u[vidx<CO>(b, 2, i, j, KK(q))];
   }
   {
      const long p1 = sidx(b, i, j, KK(1));
      const double gh = sbp->ghcof[kr - 1];
      r[0] += zz0 + gh * mu[p1] * u[vidx<CO>(b, 0, i, j, KK(0))];
      r[1] += zz1 + gh * mu[p1] * u[vidx<CO>(b, 1, i, j, KK(0))];
      r[2] += zz2 + gh * (2 * mu[p1] + la[p1]) * u[vidx<CO>(b, 2, i, j, KK(0))];
   }
   r[0] += sx * sy * i144 * (mixed<CO, UA>(b, i, j, k, la, u, 1, 0, 1) + mixed<CO, UA>(b, i, j, k, mu, u, 1, 1, 0));
   r[1] += sx * sy * i144 * (mixed<CO, UA>(b, i, j, k, mu, u, 0, 0, 1) + mixed<CO, UA>(b, i, j, k, la, u, 0, 1, 0));

   double w3x[4], w3y[4], w1x[4], w2y[4];
#pragma unroll
   for (int s = 0; s < 4; s++) {
      double a3x = 0, a3y = 0, a1x = 0, a2y = 0;
#pragma unroll 1
      for (int q = 1; q <= 8; q++) {
         const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
         a3x += bq * u[vidx<CO>(b, 2, i + off[s], j, KK(q))];
         a3y += bq * u[vidx<CO>(b, 2, i, j + off[s], KK(q))];
         a1x += bq * u[vidx<CO>(b, 0, i + off[s], j, KK(q))];
         a2y += bq * u[vidx<CO>(b, 1, i, j + off[s], KK(q))];
      }
      w3x[s] = la[sidx(b, i + off[s], j, k)] * a3x;
      w3y[s] = la[sidx(b, i, j + off[s], k)] * a3y;
      w1x[s] = mu[sidx(b, i + off[s], j, k)] * a1x;
      w2y[s] = mu[sidx(b, i, j + off[s], k)] * a2y;
   }
   double t3x = 0, t3y = 0, t1x = 0, t2y = 0;
#pragma unroll 1
   for (int q = 1; q <= 8; q++) {
      const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
      const int kq = KK(q);
      const long pq = sidx(b, i, j, kq);
      const double dx3 = d4(u[vidx<CO>(b, 2, i - 2, j, kq)], u[vidx<CO>(b, 2, i - 1, j, kq)],
                            u[vidx<CO>(b, 2, i + 1, j, kq)], u[vidx<CO>(b, 2, i + 2, j, kq)]);
      const double dy3 = d4(u[vidx<CO>(b, 2, i, j - 2, kq)], u[vidx<CO>(b, 2, i, j - 1, kq)],
                            u[vidx<CO>(b, 2, i, j + 1, kq)], u[vidx<CO>(b, 2, i, j + 2, kq)]);
      const double dx1 = d4(u[vidx<CO>(b, 0, i - 2, j, kq)], u[vidx<CO>(b, 0, i - 1, j, kq)],
                            u[vidx<CO>(b, 0, i + 1, j, kq)], u[vidx<CO>(b, 0, i + 2, j, kq)]);
      const double dy2 = d4(u[vidx<CO>(b, 1, i, j - 2, kq)], u[vidx<CO>(b, 1, i, j - 1, kq)],
                            u[vidx<CO>(b, 1, i, j + 1, kq)], u[vidx<CO>(b, 1, i, j + 2, kq)]);
      t3x += bq * (mu[pq] * i12 * dx3);
      t3y += bq * (mu[pq] * i12 * dy3);
      t1x += bq * (la[pq] * i12 * dx1);
      t2y += bq * (la[pq] * i12 * dy2);
   }
   r[0] += sx * (i12 * d4(w3x[0], w3x[1], w3x[2], w3x[3]));
   r[0] += sx * t3x;
   r[1] += sy * (i12 * d4(w3y[0], w3y[1], w3y[2], w3y[3]));
   r[1] += sy * t3y;
   r[2] += sx * (i12 * d4(w1x[0], w1x[1], w1x[2], w1x[3]));
   r[2] += sy * (i12 * d4(w2y[0], w2y[1], w2y[2], w2y[3]));
   r[2] += sx * t1x;
   r[2] += sy * t2y;
#undef KK
}

// Point q of the ghost shell, the 2-point frame outside [1,nx]x[1,ny]x[1,nz]
// (the k planes kb, kb+1, ke-1, ke first, then the j and i frames of the
// remaining planes); false past the last shell point.
__device__ __forceinline__ bool shell_point(const Box& b, long q, int& i, int& j, int& k)
{
   const long ni = b.ni, nj = b.je - b.jb + 1;
   const long nxi = ni - 4, nyi = nj - 4, nzi = (b.ke - b.kb + 1) - 4;
   const long nk_planes = 4 * ni * nj;      // k in {kb,kb+1,ke-1,ke}
   const long nj_planes = 4 * ni * nzi;     // j in {jb,jb+1,je-1,je}, k interior
   const long ni_planes = 4 * nyi * nzi;    // i in {ib,ib+1,ie-1,ie}, j,k interior
   if (q < nk_planes) {
      const long pl = q / (ni * nj), r = q % (ni * nj);
      k = pl < 2 ? b.kb + (int)pl : b.ke - 3 + (int)pl;
      j = b.jb + (int)(r / ni);
      i = b.ib + (int)(r % ni);
      return true;
   }
   q -= nk_planes;
   if (q < nj_planes) {
      const long pl = q / (ni * nzi), r = q % (ni * nzi);
      j = pl < 2 ? b.jb + (int)pl : b.je - 3 + (int)pl;
      k = b.kb + 2 + (int)(r / ni);
      i = b.ib + (int)(r % ni);
      return true;
   }
   q -= nj_planes;
   if (q < ni_planes) {
      // four consecutive threads take the four ghost columns of one row, so a
      // wave touches 16 rows (two cache lines each) rather than 64
      const long s = q % 4, r = q / 4;
      i = s < 2 ? b.ib + (int)s : b.ie - 3 + (int)s;
      k = b.kb + 2 + (int)(r / nyi);
      j = b.jb + 2 + (int)(r % nyi);
      return true;
   }
   (void)nxi;
   return false;
}

// Damping increment (to be subtracted from up) of the 4th-order supergrid
// term at one point, all three components (ew-cfromfort.C:985-1041).
template <int CO>
__device__ __forceinline__ void addsgd4_point(const Box& b, int i, int j, int k,
                                              const double* __restrict__ u,
                                              const double* __restrict__ um,
                                              const double* __restrict__ rho, const SgCoef& sg,
                                              double beta, double out[3])
{
   const int ii = i - b.ib, jj = j - b.jb, kk = k - b.kb;
   const double r0 = rho[sidx(b, i, j, k)];
   const double birho = beta / r0;
   const double rxp = rho[sidx(b, i + 1, j, k)] * sg.dcx[ii + 1], rx0 = r0 * sg.dcx[ii],
                rxm = rho[sidx(b, i - 1, j, k)] * sg.dcx[ii - 1];
   const double ryp = rho[sidx(b, i, j + 1, k)] * sg.dcy[jj + 1], ry0 = r0 * sg.dcy[jj],
                rym = rho[sidx(b, i, j - 1, k)] * sg.dcy[jj - 1];
   const double rzp = rho[sidx(b, i, j, k + 1)] * sg.dcz[kk + 1], rz0 = r0 * sg.dcz[kk],
                rzm = rho[sidx(b, i, j, k - 1)] * sg.dcz[kk - 1];
   const double wx = sg.strx[ii] * sg.coy[jj] * sg.coz[kk];
   const double wy = sg.stry[jj] * sg.cox[ii] * sg.coz[kk];
   const double wz = sg.strz[kk] * sg.cox[ii] * sg.coy[jj];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double dx[5], dy[5], dz[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
         long q = vidx<CO>(b, c, i + t - 2, j, k);
         dx[t] = u[q] - um[q];
         q = vidx<CO>(b, c, i, j + t - 2, k);
         dy[t] = u[q] - um[q];
         q = vidx<CO>(b, c, i, j, k + t - 2);
         dz[t] = u[q] - um[q];
      }
      const double ax = rxp * (dx[4] - 2 * dx[3] + dx[2]) - 2 * rx0 * (dx[3] - 2 * dx[2] + dx[1]) +
                        rxm * (dx[2] - 2 * dx[1] + dx[0]);
      const double ay = ryp * (dy[4] - 2 * dy[3] + dy[2]) - 2 * ry0 * (dy[3] - 2 * dy[2] + dy[1]) +
                        rym * (dy[2] - 2 * dy[1] + dy[0]);
      const double az = rzp * (dz[4] - 2 * dz[3] + dz[2]) - 2 * rz0 * (dz[3] - 2 * dz[2] + dz[1]) +
                        rzm * (dz[2] - 2 * dz[1] + dz[0]);
      out[c] = birho * (wx * ax + wy * ay + wz * az);
   }
}

// Ghost value below (kl=1, top) or above (kl=-1, bottom) the free surface at
// plane k, from zero-traction minus the boundary forcing t[3]
// (ew-cfromfort.C:683-742).  u is an array or a read accessor (operator[]
// on the layout's offsets returning double).
template <int CO, class UA>
__device__ __forceinline__ void freesurf_point(const Box& b, int i, int j, int k, int kl, double h,
                                               const double* sbop, const UA& u,
                                               const double* __restrict__ mu,
                                               const double* __restrict__ la, const double t[3],
                                               const double* __restrict__ strx,
                                               const double* __restrict__ stry, double g[3])
{
   const double d4a = 2.0 / 3.0, d4b = -1.0 / 12.0;
   const long p = sidx(b, i, j, k);
   const double sx = strx[i - b.ib], sy = stry[j - b.jb];
#define UU(c, di, dj, dk) u[vidx<CO>(b, c, i + (di), j + (dj), k + (dk))]
   const double wx = sx * (d4a * (UU(2, 1, 0, 0) - UU(2, -1, 0, 0)) + d4b * (UU(2, 2, 0, 0) - UU(2, -2, 0, 0)));
   const double ux = sx * (d4a * (UU(0, 1, 0, 0) - UU(0, -1, 0, 0)) + d4b * (UU(0, 2, 0, 0) - UU(0, -2, 0, 0)));
   const double wy = sy * (d4a * (UU(2, 0, 1, 0) - UU(2, 0, -1, 0)) + d4b * (UU(2, 0, 2, 0) - UU(2, 0, -2, 0)));
   const double vy = sy * (d4a * (UU(1, 0, 1, 0) - UU(1, 0, -1, 0)) + d4b * (UU(1, 0, 2, 0) - UU(1, 0, -2, 0)));
   double dz0 = 0, dz1 = 0, dz2 = 0;
#pragma unroll
   for (int m = 1; m <= 4; m++) {
      dz0 += sbop[m] * UU(0, 0, 0, kl * (m - 1));
      dz1 += sbop[m] * UU(1, 0, 0, kl * (m - 1));
      dz2 += sbop[m] * UU(2, 0, 0, kl * (m - 1));
   }
#undef UU
   const double m0 = mu[p], l0 = la[p];
   g[0] = (-dz0 - kl * wx + kl * h * t[0] / m0) / sbop[0];
   g[1] = (-dz1 - kl * wy + kl * h * t[1] / m0) / sbop[0];
   g[2] = (-dz2 + (-kl * l0 * (ux + vy) + kl * h * t[2]) / (2 * m0 + l0)) / sbop[0];
}

// Source time functions and second derivatives (time_functions.C): Ricker
// :175-182/:202-209, Gaussian :284-290/:311-317, C6SmoothBump
// :1637-1648/:1676-1687.  Exponent cut-off -700 (Source.C:2183-2187).
__device__ __forceinline__ double time_function(int kind, int deriv2, double freq, double t)
{
   const double cutoff = -700.0;
   if (kind == 0) {
      const double x = M_PI * freq * t, f = x * x;
      if (!(-f > cutoff)) return 0.0;
      if (!deriv2) return (2 * f - 1) * exp(-f);
      return M_PI * M_PI * freq * freq * (6 - 24 * f + 8 * f * f) * exp(-f);
   } else if (kind == 1) {
      const double tf = t * freq, f = tf * tf / 2;
      if (!(-f > cutoff)) return 0.0;
      if (!deriv2) return freq / sqrt(2 * M_PI) * exp(-f);
      return freq / sqrt(2 * M_PI) * freq * freq * (2 * f - 1) * exp(-f);
   } else if (kind == 2) {
      const double s = t * freq;
      if (s < 0 || s > 1) return 0.0;
      const double p = s * (1 - s);
      const double p2 = p * p, p4 = p2 * p2, p5 = p4 * p, p6 = p4 * p2;
      if (!deriv2) return 51480 * (p6 * p);
      const double q = 1 - 2 * s;
      return 51480 * freq * freq * 7 * (6 * q * q * p5 - 2 * p6);
   }
   return 0.0;
}

// ---------------------------------------------------------------------------
// launchers (kernels_basic.hip)
// ---------------------------------------------------------------------------
int launch_rhs4sg(int co, const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                  double* lu, const double* u, const double* mu, const double* la, double h,
                  const double* strx, const double* stry, const double* strz, hipStream_t s);
int launch_pred(int co, const Box& b, double* up, const double* u, const double* um,
                const double* lu, const double* fo, const double* rho, double dt2, hipStream_t s);
int launch_corr(int co, const Box& b, double* up, const double* lu, const double* fo,
                const double* rho, double dt4, hipStream_t s);
int launch_dpdmt(const Box& b, const double* up, const double* u, const double* um, double* uacc,
                 double dt2i, hipStream_t s);
int launch_addsgd4(int co, const Box& b, double* up, const double* u, const double* um,
                   const double* rho, const SgCoef& sg, double beta, hipStream_t s);
int launch_bcfortsg(int co, const Box& b, const int* wind, int nx, int ny, int nz, double* u,
                    double h, const int* bccnd, const SbpTables* d_sbp, const double* mu,
                    const double* la, const double* const* bforce, const double* strx,
                    const double* stry, hipStream_t s);
int launch_force(int co, const Box& b, double* F, const SourceTable& st, double t, int deriv2,
                 bool zero_first, hipStream_t s);
int launch_zero(double* a, long n, hipStream_t s);

}  // namespace sw4amd
