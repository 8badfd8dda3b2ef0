// kernels_adj.hip -- transposed kernels of the FP64 step (declarations and
// notation in kernels_adj.h).  Thread per point, bounds-checked reads.
//
// Interior rows: the interior formula of rhs4sg is, for the stretching
// weights S = strx(i) stry(j) strz(k), S-symmetric: its second-difference
// coefficients between two points depend only on the material around their
// midpoint (varcoef4), and each mixed term (coef D_b u)_a is the transpose of
// its partner (coef D_a u)_b.  Hence L^T z = S^{-1} L(S z) evaluated with the
// same stencil on every point of the box, reading 0 outside the rows.  The
// in-plane part of the SBP rows is the same with S = strx stry.  The k
// direction of the SBP rows (acof, ghcof, bope terms, rhs4sg_rev.C:353-629)
// is transposed term by term.
#include <algorithm>

#include "kernels_adj.h"
#include "kernels_fast.h"
#include "kernels_step.h"
#include "stencil_body.h"

namespace sw4amd {

namespace {
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}  // namespace

RowMap make_rowmap(const Box& b, int nz, const int* onesided)
{
   RowMap m;
   m.nx = b.ie - b.ib - 3;
   m.ny = b.je - b.jb - 3;
   m.nz = nz;
   m.top = onesided[4] == 1;
   m.bot = onesided[5] == 1;
   m.k1 = m.top ? 7 : 1;
   m.k2 = m.bot ? nz - 6 : nz;
   return m;
}

// coefficient of f(a) in d4 = f(-2) - f(2) + 8 (f(1) - f(-1)), a in {-2,-1,1,2}
__device__ __forceinline__ double cd4(int a) { return a == -2 ? 1.0 : a == -1 ? -8.0 : a == 1 ? 8.0 : -1.0; }

__device__ __forceinline__ bool inbox(const Box& b, int i, int j, int k)
{
   return i >= b.ib && i <= b.ie && j >= b.jb && j <= b.je && k >= b.kb && k <= b.ke;
}
__device__ __forceinline__ bool int_row(const RowMap& m, int i, int j, int k)
{
   return i >= 1 && i <= m.nx && j >= 1 && j <= m.ny && k >= m.k1 && k <= m.k2;
}
__device__ __forceinline__ bool sbp_row(const RowMap& m, int i, int j, int k)
{
   return i >= 1 && i <= m.nx && j >= 1 && j <= m.ny &&
          ((m.top && k >= 1 && k <= 6) || (m.bot && k >= m.nz - 5 && k <= m.nz));
}
// plane k holds rows (of either kind) on its columns 1..nx x 1..ny
__device__ __forceinline__ bool plane_rows(const RowMap& m, int k)
{
   return (k >= m.k1 && k <= m.k2) || (m.top && k >= 1 && k <= 6) || (m.bot && k >= m.nz - 5 && k <= m.nz);
}

struct AdjCtx {
   Box b;
   RowMap m;
   const SbpTables* sbp;
   const double* mu;
   const double* la;
   const double* strx;
   const double* stry;
   const double* strz;
   __device__ __forceinline__ double vmu(int i, int j, int k) const { return inbox(b, i, j, k) ? mu[sidx(b, i, j, k)] : 0.0; }
   __device__ __forceinline__ double vla(int i, int j, int k) const { return inbox(b, i, j, k) ? la[sidx(b, i, j, k)] : 0.0; }
   __device__ __forceinline__ double sx(int i) const { return i >= b.ib && i <= b.ie ? strx[i - b.ib] : 1.0; }
   __device__ __forceinline__ double sy(int j) const { return j >= b.jb && j <= b.je ? stry[j - b.jb] : 1.0; }
   __device__ __forceinline__ double sz(int k) const { return k >= b.kb && k <= b.ke ? strz[k - b.kb] : 1.0; }
   // a component-blocked field, 0 outside the box
   __device__ __forceinline__ double f(const double* p, int c, int i, int j, int k) const
   {
      return inbox(b, i, j, k) ? p[sidx(b, i, j, k) + b.npts * c] : 0.0;
   }
   // z on the interior-formula rows / on the SBP rows, 0 elsewhere
   __device__ __forceinline__ double zi(const double* z, int c, int i, int j, int k) const
   {
      return int_row(m, i, j, k) ? z[sidx(b, i, j, k) + b.npts * c] : 0.0;
   }
   __device__ __forceinline__ double zb(const double* z, int c, int i, int j, int k) const
   {
      return sbp_row(m, i, j, k) ? z[sidx(b, i, j, k) + b.npts * c] : 0.0;
   }
};

// x and y parts of the interior formula with the xy mixed terms: the
// in-plane part of an SBP row (rhs4sg_rev.C:353-629 before the k terms).
template <class S>
__device__ __forceinline__ void rhs4_xy_body(const S& s, double r[3])
{
   double am[5], al[5], mx[4], lx[4], my[4], ly[4];
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
   double cx[4], cy[4];
#pragma unroll
   for (int q = 0; q < 4; q++) {
      cx[q] = 2 * mx[q] + lx[q];
      cy[q] = 2 * my[q] + ly[q];
   }
   const double sx = s.sx(0), sy = s.sy(0);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double f[5];
#pragma unroll
      for (int q = 0; q < 5; q++) f[q] = s.u(c, q - 2, 0, 0);
      double acc = sx * varsum4c(c == 0 ? cx : mx, f);
#pragma unroll
      for (int q = 0; q < 5; q++) f[q] = s.u(c, 0, q - 2, 0);
      acc += sy * varsum4c(c == 1 ? cy : my, f);
      r[c] = acc * (1.0 / 6);
   }
   const double sxy = sx * sy * (1.0 / 144);
   r[0] += sxy * (mixed_term<0, 1, 1, true, double>(s) + mixed_term<1, 0, 1, false, double>(s));
   r[1] += sxy * (mixed_term<0, 1, 0, false, double>(s) + mixed_term<1, 0, 0, true, double>(s));
}

// k-direction terms of the SBP rows of one side, transposed, at the point
// (i,j,k): row kr lives on grid plane kbase + sgn (kr-1), plane q (0..8) of
// the row space on kbase + sgn (q-1).
__device__ void sbp_col_T(const AdjCtx& A, const double* z, int i, int j, int k, int kbase, int sgn, double r[3])
{
   const int q = sgn * (k - kbase) + 1;
   if (q < 0 || q > 8) return;
   const SbpTables* sbp = A.sbp;
   const double i12 = 1.0 / 12;
   const bool col = i >= 1 && i <= A.m.nx && j >= 1 && j <= A.m.ny;
   // mu, lambda on the side's planes m = 1..8 of this column
   double muc[8], lac[8];
#pragma unroll
   for (int mm = 0; mm < 8; mm++) {
      const int kmg = kbase + sgn * mm;
      muc[mm] = A.vmu(i, j, kmg);
      lac[mm] = A.vla(i, j, kmg);
   }
   // mu, lambda around this point in its plane (x and y taps)
   double mxa[4], mya[4], lxa[4], lya[4];
   const double sxa[4] = {A.sx(i + 2), A.sx(i + 1), A.sx(i - 1), A.sx(i - 2)};
   const double sya[4] = {A.sy(j + 2), A.sy(j + 1), A.sy(j - 1), A.sy(j - 2)};
#pragma unroll
   for (int t = 0; t < 4; t++) {
      const int a = t < 2 ? t - 2 : t - 1;
      mxa[t] = A.vmu(i - a, j, k);
      mya[t] = A.vmu(i, j - a, k);
      lxa[t] = A.vla(i - a, j, k);
      lya[t] = A.vla(i, j - a, k);
   }
#pragma unroll
   for (int kr = 1; kr <= 6; kr++) {
      const int krg = kbase + sgn * (kr - 1);
      const double z0 = A.zb(z, 0, i, j, krg), z1 = A.zb(z, 1, i, j, krg), z2 = A.zb(z, 2, i, j, krg);
      // acof / ghcof: u_c(q) enters row kr of the same column
      double amu = 0, a2ml = 0;
      if (q == 0) {
         const double gh = sbp->ghcof[kr - 1];
         amu = gh * muc[0];
         a2ml = gh * (2 * muc[0] + lac[0]);
      } else {
#pragma unroll
         for (int mm = 0; mm < 8; mm++) {
            const double a = sbp->acof[(kr - 1) + 6 * (q - 1) + 48 * mm];
            amu += a * muc[mm];
            a2ml += a * (2 * muc[mm] + lac[mm]);
         }
      }
      if (col) {
         r[0] += z0 * amu;
         r[1] += z1 * amu;
         r[2] += z2 * a2ml;
      }
      if (q < 1) continue;
      const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
      double s2 = 0, s0 = 0, s1 = 0;
      const double lar = lac[kr - 1], mur = muc[kr - 1];
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int a = t < 2 ? t - 2 : t - 1;
         const double ca = cd4(a);
         const double wx0 = sxa[t] * A.zb(z, 0, i - a, j, krg);
         const double wy1 = sya[t] * A.zb(z, 1, i, j - a, krg);
         const double wx2 = sxa[t] * A.zb(z, 2, i - a, j, krg);
         const double wy2 = sya[t] * A.zb(z, 2, i, j - a, krg);
         // (la w_z)_x, (la w_z)_y: coefficient la at this column, row plane
         s2 += ca * lar * (wx0 + wy1);
         // (mu w_x)_z, (mu w_y)_z: coefficient mu at the row's column, plane q
         s2 += ca * (mxa[t] * wx0 + mya[t] * wy1);
         // (mu u_z)_x and (la u_x)_z
         s0 += ca * (mur * wx2 + lxa[t] * wx2);
         // (mu v_z)_y and (la v_y)_z
         s1 += ca * (mur * wy2 + lya[t] * wy2);
      }
      r[0] += bq * i12 * s0;
      r[1] += bq * i12 * s1;
      r[2] += bq * i12 * s2;
   }
}

// Planes per workgroup for an LDS k-march of `tiles` column tiles over `npl`
// planes when `per_cu` workgroups fit on a CU: the k-block count near npl/64
// that minimises (rounds of resident workgroups) x (planes per workgroup +
// the 4 planes each one primes).  Workgroups run in rounds, so a count just
// above a multiple of the resident slots costs a whole extra round.
int march_kchunk(long tiles, int npl, int per_cu)
{
   static int cus = 0;
   if (!cus) {
      int dev = 0, v = 0;
      if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) {
         (void)hipGetLastError();
         v = 256;
      }
      cus = v;
   }
   const long slots = (long)cus * per_cu;
   const int nc0 = std::max(1, (npl + 32) / 64);
   int best = nc0;
   long bcost = -1;
   for (int nc = std::max(1, nc0 / 2); nc <= 2 * nc0 + 1; nc++) {
      const long rounds = (tiles * nc + slots - 1) / slots;
      const long cost = rounds * ((npl + nc - 1) / nc + 4);
      if (bcost < 0 || cost < bcost) {
         bcost = cost;
         best = nc;
      }
   }
   return (npl + best - 1) / best;
}

Box pad_box(const Box& b) { return make_box(b.ib - 2, b.ie + 2, b.jb - 2, b.je + 2, b.kb - 2, b.ke + 2); }

long adj_pad_doubles(const Box& b)
{
   const Box pb = pad_box(b);
   return 8 * pb.npts + pb.ni + (pb.je - pb.jb + 1) + (pb.ke - pb.kb + 1);
}

AdjPad make_adj_pad(const Box& b, double* mem)
{
   AdjPad p;
   p.pb = pad_box(b);
   p.w = mem;
   p.mu = mem + 3 * p.pb.npts;
   p.la = p.mu + p.pb.npts;
   p.sx = p.la + p.pb.npts;
   p.sy = p.sx + p.pb.ni;
   p.sz = p.sy + (p.pb.je - p.pb.jb + 1);
   p.o = p.sz + (p.pb.ke - p.pb.kb + 1);

   return p;
}

__global__ void k_adj_pad_material(Box b, Box pb, const double* __restrict__ mu, const double* __restrict__ la,
                                   const double* __restrict__ strx, const double* __restrict__ stry,
                                   const double* __restrict__ strz, double* __restrict__ mup, double* __restrict__ lap,
                                   double* __restrict__ sxp, double* __restrict__ syp, double* __restrict__ szp)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q < pb.npts) {
      const int i = pb.ib + (int)(q % pb.ni);
      const int j = pb.jb + (int)((q / pb.ni) % (pb.je - pb.jb + 1));
      const int k = pb.kb + (int)(q / pb.nij);
      const bool in = inbox(b, i, j, k);
      mup[q] = in ? mu[sidx(b, i, j, k)] : 0.0;
      lap[q] = in ? la[sidx(b, i, j, k)] : 0.0;
   }
   if (q < pb.ni) sxp[q] = (pb.ib + q >= b.ib && pb.ib + q <= b.ie) ? strx[pb.ib + q - b.ib] : 1.0;
   if (q < pb.je - pb.jb + 1) syp[q] = (pb.jb + q >= b.jb && pb.jb + q <= b.je) ? stry[pb.jb + q - b.jb] : 1.0;
   if (q < pb.ke - pb.kb + 1) szp[q] = (pb.kb + q >= b.kb && pb.kb + q <= b.ke) ? strz[pb.kb + q - b.kb] : 1.0;
}

// padded W: INPLANE = false: S zI (interior-formula rows, S = sx sy sz);
// true: sx sy zB (SBP rows)
template <bool INPLANE>
__global__ void k_adj_weight(AdjCtx A, AdjPad P, const double* __restrict__ z)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const Box& pb = P.pb;
   if (q >= pb.npts) return;
   const int i = pb.ib + (int)(q % pb.ni);
   const int j = pb.jb + (int)((q / pb.ni) % (pb.je - pb.jb + 1));
   const int k = pb.kb + (int)(q / pb.nij);
   const bool row = INPLANE ? sbp_row(A.m, i, j, k) : int_row(A.m, i, j, k);
   const double s = row ? A.sx(i) * A.sy(j) * (INPLANE ? 1.0 : A.sz(k)) : 0.0;
   const long p = row ? sidx(A.b, i, j, k) : 0;
   for (int c = 0; c < 3; c++) {
      const double v = row ? s * z[p + A.b.npts * c] : 0.0;
      P.w[q + pb.npts * c] = v;
   }
}

// out = o on the box (o lives on the padded box)
__global__ void k_adj_crop(Box b, Box pb, const double* __restrict__ o, double* __restrict__ out)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const long q = sidx(pb, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) out[p + b.npts * c] = o[q + pb.npts * c];
}

// Accessor of rhs4_xy_body over the padded arrays (no checks).
struct PadSten {
   const AdjPad& P;
   int i, j, k;
   __device__ __forceinline__ double u(int c, int di, int dj, int dk) const
   {
      return P.w[sidx(P.pb, i + di, j + dj, k + dk) + P.pb.npts * c];
   }
   __device__ __forceinline__ double mu(int di, int dj, int dk) const { return P.mu[sidx(P.pb, i + di, j + dj, k + dk)]; }
   __device__ __forceinline__ double la(int di, int dj, int dk) const { return P.la[sidx(P.pb, i + di, j + dj, k + dk)]; }
   __device__ __forceinline__ double sx(int d) const { return P.sx[i + d - P.pb.ib]; }
   __device__ __forceinline__ double sy(int d) const { return P.sy[j + d - P.pb.jb]; }
   __device__ __forceinline__ double sz(int d) const { return P.sz[k + d - P.pb.kb]; }
};

// SBP rows, in-plane part on the row planes of one side: out += cof
// (sx sy)^{-1} Lxy(W), W = sx sy zB
__global__ void __launch_bounds__(256) k_rhs4sg_T_xy(AdjCtx A, AdjPad P, int kbase, int sgn, Box ob,
                                                     double* __restrict__ out, double cof)
{
   const Box& b = A.b;
   const int i = b.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = kbase + sgn * (int)blockIdx.z;  // rows 1..6
   if (i > b.ie || j > b.je) return;
   const PadSten s{P, i, j, k};
   double r[3];
   rhs4_xy_body(s, r);
   const double sc = cof / (A.sx(i) * A.sy(j));
   const long p = sidx(ob, i, j, k);
   for (int c = 0; c < 3; c++) out[p + ob.npts * c] += sc * r[c];
}

// SBP rows, k-direction terms on the planes 0..8 of one side
__global__ void __launch_bounds__(256) k_rhs4sg_T_col(AdjCtx A, int kbase, int sgn, const double* __restrict__ z,
                                                      Box ob, double* __restrict__ out, double cof)
{
   const Box& b = A.b;
   const int i = b.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = kbase + sgn * ((int)blockIdx.z - 1);  // planes q = 0..8
   if (i > b.ie || j > b.je || k < b.kb || k > b.ke) return;
   double r[3] = {0, 0, 0};
   sbp_col_T(A, z, i, j, k, kbase, sgn, r);
   const long p = sidx(ob, i, j, k);
   for (int c = 0; c < 3; c++) out[p + ob.npts * c] += cof * r[c];
}

// ---------------------------------------------------------------------------
// Material derivatives
// ---------------------------------------------------------------------------
// acof / ghcof / bope terms of the SBP rows of one side at p
__device__ void sbp_col_grad(const AdjCtx& A, const double* z, const double* u, int i, int j, int k, int kbase,
                             int sgn, double& gm, double& gl)
{
   const int q = sgn * (k - kbase) + 1;
   if (q < 1 || q > 8) return;
   const SbpTables* sbp = A.sbp;
   const double i12 = 1.0 / 12;
   const bool col = i >= 1 && i <= A.m.nx && j >= 1 && j <= A.m.ny;
   // u on the side's planes 1..8 of this column, and on its ghost plane
   double uk[8][3], ug[3];
#pragma unroll
   for (int qq = 0; qq < 8; qq++) {
      const int kq = kbase + sgn * qq;
#pragma unroll
      for (int c = 0; c < 3; c++) uk[qq][c] = A.f(u, c, i, j, kq);
   }
#pragma unroll
   for (int c = 0; c < 3; c++) ug[c] = A.f(u, c, i, j, kbase - sgn);
   // x and y differences of u in this point's plane
   double dx2 = 0, dy2 = 0, dx0 = 0, dy1 = 0;
#pragma unroll
   for (int t = 0; t < 4; t++) {
      const int a = t < 2 ? t - 2 : t - 1;
      dx2 += cd4(a) * A.f(u, 2, i + a, j, k);
      dy2 += cd4(a) * A.f(u, 2, i, j + a, k);
      dx0 += cd4(a) * A.f(u, 0, i + a, j, k);
      dy1 += cd4(a) * A.f(u, 1, i, j + a, k);
   }
   const double sxi = A.sx(i), syj = A.sy(j);
#pragma unroll
   for (int kr = 1; kr <= 6; kr++) {
      const int krg = kbase + sgn * (kr - 1);
      const double z0 = A.zb(z, 0, i, j, krg), z1 = A.zb(z, 1, i, j, krg), z2 = A.zb(z, 2, i, j, krg);
      if (col) {
         // mu, la at plane m = q enter row kr through acof(kr, q', q) u(q')
         double sm = 0, sl = 0;
#pragma unroll
         for (int qq = 0; qq < 8; qq++) {
            const double a = sbp->acof[(kr - 1) + 6 * qq + 48 * (q - 1)];
            sm += a * (z0 * uk[qq][0] + z1 * uk[qq][1] + 2 * z2 * uk[qq][2]);
            sl += a * z2 * uk[qq][2];
         }
         if (q == 1) {  // ghost term gh mu(1) u(0)
            const double gh = sbp->ghcof[kr - 1];
            sm += gh * (z0 * ug[0] + z1 * ug[1] + 2 * z2 * ug[2]);
            sl += gh * z2 * ug[2];
         }
         gm += sm;
         gl += sl;
      }
      // (mu w_x)_z, (mu w_y)_z, (la u_x)_z, (la v_y)_z: coefficient at (i,j) plane q, row (i,j,kr)
      const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
      gm += bq * i12 * (sxi * dx2 * z0 + syj * dy2 * z1);
      gl += bq * i12 * (sxi * dx0 + syj * dy1) * z2;
   }
   // (la w_z)_x, (la w_z)_y, (mu u_z)_x, (mu v_z)_y: coefficient at p on row plane kr = q
   if (q <= 6) {
      double bu0 = 0, bu1 = 0, bu2 = 0;
#pragma unroll
      for (int qq = 0; qq < 8; qq++) {
         const double bq = sgn * sbp->bope[(q - 1) + 6 * qq];
         bu0 += bq * uk[qq][0];
         bu1 += bq * uk[qq][1];
         bu2 += bq * uk[qq][2];
      }
      double wx0 = 0, wy1 = 0, wx2 = 0, wy2 = 0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int a = t < 2 ? t - 2 : t - 1;
         wx0 += cd4(a) * A.sx(i - a) * A.zb(z, 0, i - a, j, k);
         wy1 += cd4(a) * A.sy(j - a) * A.zb(z, 1, i, j - a, k);
         wx2 += cd4(a) * A.sx(i - a) * A.zb(z, 2, i - a, j, k);
         wy2 += cd4(a) * A.sy(j - a) * A.zb(z, 2, i, j - a, k);
      }
      gl += i12 * bu2 * (wx0 + wy1);
      gm += i12 * (bu0 * wx2 + bu1 * wy2);
   }
}

// Interior-row terms of G_mu, G_la at every point p of the box (ghosts
// included), marching along k.  A workgroup owns a 32 x 8 tile of (i,j)
// columns.  Plane k of u (0 off the box) and of z (0 off the interior-formula
// rows) sit in LDS with a halo of 2, double-buffered so that one barrier per
// plane suffices; the next plane is fetched into registers before the current
// one is evaluated.  The thread's own columns of u and z (k-2..k+2) rotate in
// registers.
//
// Along axis d, for component c, with w(e) = u_c(p + e e_d) and
// y(e) = s_d(p + e e_d) z_c(p + e e_d), e = -2..2: the five rows p + e e_d
// whose varcoef4 coefficients contain the material at p contribute
// s_d(p)/6 sum_e w(e) K_e(y) (second_form below: the psi factors of varcoef4
// summed over the rows, which only reach u within 2 points of p).  The mixed
// terms (coef D_a u_b)_d are D[a][b] = d4 of u_b along a at p times
// V[r][d] = d4 of s_d z_r along d, taken with its sign reversed, at p.
namespace {
constexpr int GTX = 32, GTY = 8, GNT = GTX * GTY;
constexpr int GHX = GTX + 4, GHY = GTY + 4, GPL = GHX * GHY;  // plane with a halo of 2
constexpr int GSL = (GPL + GNT - 1) / GNT;                      // staging slots per thread
}  // namespace

// sum_e w(e) K_e(y), arrays indexed e + 2
__device__ __forceinline__ double second_form(const double w[5], const double y[5])
{
   const double k0 = y[1] - 0.75 * y[2] - 0.25 * y[0];
   const double k1 = y[3] + 3 * y[2] - 5 * y[1] + y[0];
   const double k2 = 3 * (y[3] + y[1]) - 0.75 * (y[4] + y[0]) - 4.5 * y[2];
   const double k3 = y[4] - 5 * y[3] + 3 * y[2] + y[1];
   const double k4 = y[3] - 0.75 * y[2] - 0.25 * y[4];
   return w[0] * k0 + w[1] * k1 + w[2] * k2 + w[3] * k3 + w[4] * k4;
}
// d4 at e = 0: sum_a cd4(a) f(a), and the same of f(-a)
__device__ __forceinline__ double d4w(const double f[5]) { return f[0] - f[4] + 8 * (f[3] - f[1]); }
__device__ __forceinline__ double d4r(const double f[5]) { return f[4] - f[0] + 8 * (f[1] - f[3]); }

__global__ void __launch_bounds__(GNT, 2) k_rhs4sg_matgrad_march(AdjCtx A, int ntx, int nty, int kc,
                                                                 const double* __restrict__ z,
                                                                 const double* __restrict__ u,
                                                                 double* __restrict__ gmu, double* __restrict__ gla,
                                                                 double cof)
{
   __shared__ double su[2][3 * GPL];
   __shared__ double sz[2][3 * GPL];
   const Box& b = A.b;
   const RowMap& m = A.m;
   const int t = blockIdx.x;
   const int tix = t % ntx, tr = t / ntx, tjy = tr % nty, tkz = tr / nty;
   const int tid = threadIdx.x, tx = tid % GTX, ty = tid / GTX;
   const int ia = b.ib + tix * GTX, ja = b.jb + tjy * GTY;
   const int i = ia + tx, j = ja + ty;
   const bool active = i <= b.ie && j <= b.je;
   const int k0 = b.kb + tkz * kc, k1 = min(b.ke, k0 + kc - 1);
   const long n = b.npts;
   // staging slots: in-plane offset of the point (-1 off the box) and whether
   // its column carries rows (interior-formula or SBP rows: same columns)
   int soff[GSL];
   bool srow[GSL];
#pragma unroll
   for (int q = 0; q < GSL; q++) {
      const int e = tid + GNT * q, yy = e / GHX, xx = e - yy * GHX;
      const int gi = ia - 2 + xx, gj = ja - 2 + yy;
      const bool in = e < GPL && gi >= b.ib && gi <= b.ie && gj >= b.jb && gj <= b.je;
      soff[q] = in ? (gj - b.jb) * b.ni + (gi - b.ib) : -1;
      srow[q] = in && gi >= 1 && gi <= m.nx && gj >= 1 && gj <= m.ny;
   }
   // reads go to a point of the box and are masked afterwards (no branches)
   double pu[GSL][3], pz[GSL][3];
   auto fetch = [&](int kk) {  // plane kk (inside the box) into pu, pz
      const long pk = (long)(kk - b.kb) * b.nij;
      const bool krow = plane_rows(m, kk);
#pragma unroll
      for (int q = 0; q < GSL; q++) {
         const long a = pk + max(soff[q], 0);
         const bool zin = krow && srow[q];
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const double uv = u[a + n * c], zv = z[a + n * c];
            pu[q][c] = soff[q] >= 0 ? uv : 0.0;
            pz[q][c] = zin ? zv : 0.0;
         }
      }
   };
   auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < GSL; q++) {
         const int e = tid + GNT * q;
         if (e < GPL) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
               su[buf][c * GPL + e] = pu[q][c];
               sz[buf][c * GPL + e] = pz[q][c];
            }
         }
      }
   };
   // own column (inactive threads keep zeros)
   const int ic = active ? i : b.ib, jc = active ? j : b.jb;
   const long col = (long)(jc - b.jb) * b.ni + (ic - b.ib);
   const bool crow = active && i >= 1 && i <= m.nx && j >= 1 && j <= m.ny;
   double uc[5][3], zc[5][3], szr[5];
   auto column = [&](int kk, double* uq, double* zq, double& sq) {
      const long pk = (long)(min(max(kk, b.kb), b.ke) - b.kb) * b.nij + col;
      const bool uin = active && kk >= b.kb && kk <= b.ke, zin = crow && kk >= m.k1 && kk <= m.k2;
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const double uv = u[pk + n * c], zv = z[pk + n * c];
         uq[c] = uin ? uv : 0.0;
         zq[c] = zin ? zv : 0.0;
      }
      sq = A.sz(kk);
   };
#pragma unroll
   for (int q = 0; q < 4; q++) column(k0 - 2 + q, uc[q], zc[q], szr[q]);
   // one plane ahead: the column values at k+2 and gmu, gla at this point,
   // so that their latency overlaps the plane before
   double nu[3], nz[3], nsz;
   column(k0 + 2, nu, nz, nsz);
   double ngm = gmu[(long)(k0 - b.kb) * b.nij + col], ngl = gla[(long)(k0 - b.kb) * b.nij + col];
   double sxo[5], syo[5];
#pragma unroll
   for (int e = 0; e < 5; e++) {
      sxo[e] = A.sx(ic + e - 2);
      syo[e] = A.sy(jc + e - 2);
   }
   fetch(k0);
   store(0);
   __syncthreads();
   const int ctr = (ty + 2) * GHX + tx + 2;
   for (int k = k0; k <= k1; k++) {
      const int cur = (k - k0) & 1;
#pragma unroll
      for (int c = 0; c < 3; c++) {
         uc[4][c] = nu[c];
         zc[4][c] = nz[c];
      }
      szr[4] = nsz;
      const double gmc = ngm, glc = ngl;
      if (k < k1) {
         fetch(k + 1);
         column(k + 3, nu, nz, nsz);
         const long pn = (long)(k + 1 - b.kb) * b.nij + col;  // col: a point of the box also where inactive
         ngm = gmu[pn];
         ngl = gla[pn];
      }
      if (active) {
         const double* U = su[cur];
         const double* Zs = sz[cur];
         double D[3][3], V[3][3], gs = 0, gd = 0;
#pragma unroll
         for (int c = 0; c < 3; c++) {
            double w[5], y[5], S;
#pragma unroll
            for (int e = 0; e < 5; e++) {
               w[e] = U[c * GPL + ctr + e - 2];
               y[e] = sxo[e] * Zs[c * GPL + ctr + e - 2];
            }
            S = sxo[2] * second_form(w, y);
            D[0][c] = d4w(w);
            V[c][0] = d4r(y);
            gs += S;
            if (c == 0) gd += S;
#pragma unroll
            for (int e = 0; e < 5; e++) {
               w[e] = U[c * GPL + ctr + (e - 2) * GHX];
               y[e] = syo[e] * Zs[c * GPL + ctr + (e - 2) * GHX];
            }
            S = syo[2] * second_form(w, y);
            D[1][c] = d4w(w);
            V[c][1] = d4r(y);
            gs += S;
            if (c == 1) gd += S;
#pragma unroll
            for (int e = 0; e < 5; e++) {
               w[e] = uc[e][c];
               y[e] = szr[e] * zc[e][c];
            }
            S = szr[2] * second_form(w, y);
            D[2][c] = d4w(w);
            V[c][2] = d4r(y);
            gs += S;
            if (c == 2) gd += S;
         }
         // on an SBP row plane only the in-plane terms belong to the rows
         // (their k terms are the acof/bope stencils of k_rhs4sg_matgrad_sbp)
         const double fi = k >= m.k1 && k <= m.k2 ? 1.0 : 0.0;
         const double s0 = sxo[2], s1 = syo[2], s2 = szr[2], s2i = fi * s2;
         const double ml = V[0][0] * (s1 * D[1][1] + s2i * D[2][2]) + V[1][1] * (s0 * D[0][0] + s2i * D[2][2]) +
                           V[2][2] * (s0 * D[0][0] + s1 * D[1][1]);
         const double mm = s0 * (D[0][1] * V[0][1] + D[0][2] * V[0][2]) + s1 * (D[1][0] * V[1][0] + D[1][2] * V[1][2]) +
                           s2i * (D[2][0] * V[2][0] + D[2][1] * V[2][1]);
         const long p = (long)(k - b.kb) * b.nij + col;
         gmu[p] = gmc + cof * ((gs + gd) * (1.0 / 6) + mm * (1.0 / 144));
         gla[p] = glc + cof * (gd * (1.0 / 6) + ml * (1.0 / 144));
      }
      if (k < k1) store(cur ^ 1);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++) {
#pragma unroll
         for (int c = 0; c < 3; c++) {
            uc[q][c] = uc[q + 1][c];
            zc[q][c] = zc[q + 1][c];
         }
         szr[q] = szr[q + 1];
      }
   }
}

// SBP rows of one side on its planes 0..8: in-plane part (row planes) and
// the k-direction terms
__global__ void __launch_bounds__(256) k_rhs4sg_matgrad_sbp(AdjCtx A, int kbase, int sgn, const double* __restrict__ z,
                                                            const double* __restrict__ u, double* __restrict__ gmu,
                                                            double* __restrict__ gla, double cof)
{
   const Box& b = A.b;
   const int i = b.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = kbase + sgn * (int)blockIdx.z;  // planes 1..8 of the side
   if (i > b.ie || j > b.je || k < b.kb || k > b.ke) return;
   double gm = 0, gl = 0;
   sbp_col_grad(A, z, u, i, j, k, kbase, sgn, gm, gl);
   const long p = sidx(b, i, j, k);
   gmu[p] += cof * gm;
   gla[p] += cof * gl;
}

// ---------------------------------------------------------------------------
// Supergrid damping, transposed
// ---------------------------------------------------------------------------
struct DampCtx {
   Box b;
   int nx, ny, nz;
   SgCoef sg;
   const double* rho;
   double beta;
   __device__ __forceinline__ bool interior(int i, int j, int k) const
   {
      return i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
   }
};

// Damping transpose as an LDS k-march over 32 x 8 column tiles (the tile
// geometry of k_rhs4sg_matgrad_march).  Planes of g = (beta/rho) y (interior
// points), d = u - um and rho (box points), 0 elsewhere, with a halo of 2,
// double-buffered with one barrier per plane; the own columns of g
// (k-2..k+2), d and rho (k-1..k+1) rotate in registers.  Per point:
// dbar = -sum_ax w_ax D2(r D2(s_ax g)) and grho += -sum_ax dc_ax D2(d) w_ax
// D2(s_ax g) at the point (r = rho dc_ax; D2 the centred second difference;
// w_x = coy coz, w_y = cox coz, w_z = cox coy).
// tiles: when given, block b runs tiles[b] (column tile and planes) instead
// of the full tiling.
__global__ void __launch_bounds__(GNT, 2) k_addsgd4_T_march(DampCtx D, int ntx, int nty, int kc,
                                                            const double* __restrict__ y, const double* __restrict__ u,
                                                            const double* __restrict__ um, double* __restrict__ dbar,
                                                            double* __restrict__ grho,
                                                            const DampTile* __restrict__ tiles)
{
   __shared__ double sgl[2][3 * GPL];
   __shared__ double sdl[2][3 * GPL];
   __shared__ double srl[2][GPL];
   const Box& b = D.b;
   const SgCoef& sg = D.sg;
   const double* rho = D.rho;
   const double beta = D.beta;
   int tix, tjy, k0, k1;
   if (tiles) {
      const DampTile d = tiles[blockIdx.x];
      tix = d.tix;
      tjy = d.tjy;
      k0 = d.k0;
      k1 = d.k1;
   } else {
      const int t = blockIdx.x;
      tix = t % ntx;
      const int tr = t / ntx;
      tjy = tr % nty;
      k0 = b.kb + (tr / nty) * kc;
      k1 = min(b.ke, k0 + kc - 1);
   }
   const int tid = threadIdx.x, tx = tid % GTX, ty = tid / GTX;
   const int ia = b.ib + tix * GTX, ja = b.jb + tjy * GTY;
   const int i = ia + tx, j = ja + ty;
   const bool active = i <= b.ie && j <= b.je;
   const long n = b.npts;
   // staging slots: in-plane offset (-1 off the box), interior column
   int soff[GSL];
   bool sint[GSL];
#pragma unroll
   for (int q = 0; q < GSL; q++) {
      const int e = tid + GNT * q, yy = e / GHX, xx = e - yy * GHX;
      const int gi = ia - 2 + xx, gj = ja - 2 + yy;
      const bool in = e < GPL && gi >= b.ib && gi <= b.ie && gj >= b.jb && gj <= b.je;
      soff[q] = in ? (gj - b.jb) * b.ni + (gi - b.ib) : -1;
      sint[q] = in && gi >= 1 && gi <= D.nx && gj >= 1 && gj <= D.ny;
   }
   // reads go to a point of the box and are masked afterwards (no branches)
   double pg[GSL][3], pd[GSL][3], pr[GSL];
   auto fetch = [&](int kk) {  // plane kk (inside the box)
      const long pk = (long)(kk - b.kb) * b.nij;
      const bool kint = kk >= 1 && kk <= D.nz;
#pragma unroll
      for (int q = 0; q < GSL; q++) {
         const long a = pk + max(soff[q], 0);
         const bool in = soff[q] >= 0, gin = kint && sint[q];
         const double r = rho[a], br = beta / r;
         pr[q] = in ? r : 0.0;
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const double yv = y[a + n * c], uv = u[a + n * c], mv = um[a + n * c];
            pg[q][c] = gin ? br * yv : 0.0;
            pd[q][c] = in ? uv - mv : 0.0;
         }
      }
   };
   auto store = [&](int buf) {
#pragma unroll
      for (int q = 0; q < GSL; q++) {
         const int e = tid + GNT * q;
         if (e < GPL) {
            srl[buf][e] = pr[q];
#pragma unroll
            for (int c = 0; c < 3; c++) {
               sgl[buf][c * GPL + e] = pg[q][c];
               sdl[buf][c * GPL + e] = pd[q][c];
            }
         }
      }
   };
   // own column (inactive threads read column (ib, jb) and keep zeros)
   const int ic = active ? i : b.ib, jc = active ? j : b.jb;
   const long col = (long)(jc - b.jb) * b.ni + (ic - b.ib);
   const bool cint = active && i >= 1 && i <= D.nx && j >= 1 && j <= D.ny;
   double gc[5][3], dc[3][3], rc[3], szr[5], dczr[3];
   auto gcolumn = [&](int kk, double* g, double& sz) {
      const long a = (long)(min(max(kk, b.kb), b.ke) - b.kb) * b.nij + col;
      const bool gin = cint && kk >= 1 && kk <= D.nz;
      const double br = beta / rho[a];
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const double yv = y[a + n * c];
         g[c] = gin ? br * yv : 0.0;
      }
      sz = kk >= b.kb && kk <= b.ke ? sg.strz[kk - b.kb] : 1.0;
   };
   auto dcolumn = [&](int kk, double* d, double& r, double& dz) {
      const long a = (long)(min(max(kk, b.kb), b.ke) - b.kb) * b.nij + col;
      const bool in = active && kk >= b.kb && kk <= b.ke;
      const double rv = rho[a];
      r = in ? rv : 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const double uv = u[a + n * c], mv = um[a + n * c];
         d[c] = in ? uv - mv : 0.0;
      }
      dz = sg.dcz[min(max(kk, b.kb), b.ke) - b.kb];
   };
#pragma unroll
   for (int q = 0; q < 4; q++) gcolumn(k0 - 2 + q, gc[q], szr[q]);
#pragma unroll
   for (int q = 0; q < 2; q++) dcolumn(k0 - 1 + q, dc[q], rc[q], dczr[q]);
   // one plane ahead: the column values the next plane adds (g at k+2, d and
   // rho at k+1) and grho at this point, so that their latency overlaps
   // the plane before
   double ng[3], nsz, nd[3], nr, ndz;
   gcolumn(k0 + 2, ng, nsz);
   dcolumn(k0 + 1, nd, nr, ndz);
   double ngr = grho[(long)(k0 - b.kb) * b.nij + col];  // col: a point of the box also where inactive
   const int ii = ic - b.ib, jj = jc - b.jb, ni = b.ie - b.ib, nj = b.je - b.jb;
   double sxo[5], syo[5];
#pragma unroll
   for (int e = 0; e < 5; e++) {
      const int iq = ic + e - 2, jq = jc + e - 2;
      sxo[e] = iq >= b.ib && iq <= b.ie ? sg.strx[iq - b.ib] : 1.0;
      syo[e] = jq >= b.jb && jq <= b.je ? sg.stry[jq - b.jb] : 1.0;
   }
   const double dcx[3] = {sg.dcx[max(ii - 1, 0)], sg.dcx[ii], sg.dcx[min(ii + 1, ni)]};
   const double dcy[3] = {sg.dcy[max(jj - 1, 0)], sg.dcy[jj], sg.dcy[min(jj + 1, nj)]};
   const double cxi = sg.cox[ii], cyj = sg.coy[jj];
   fetch(k0);
   store(0);
   __syncthreads();
   const int ctr = (ty + 2) * GHX + tx + 2;
   for (int k = k0; k <= k1; k++) {
      const int cur = (k - k0) & 1;
#pragma unroll
      for (int c = 0; c < 3; c++) {
         gc[4][c] = ng[c];
         dc[2][c] = nd[c];
      }
      szr[4] = nsz;
      rc[2] = nr;
      dczr[2] = ndz;
      const double grc = ngr;
      if (k < k1) {
         fetch(k + 1);
         gcolumn(k + 3, ng, nsz);
         dcolumn(k + 2, nd, nr, ndz);
         ngr = grho[(long)(k + 1 - b.kb) * b.nij + col];
      }
      if (active) {
         const double* G = sgl[cur];
         const double* Dd = sdl[cur];
         const double* R = srl[cur];
         const int kk = k - b.kb;
         const double czk = sg.coz[kk];
         const double rx[3] = {R[ctr - 1] * dcx[0], R[ctr] * dcx[1], R[ctr + 1] * dcx[2]};
         const double ry[3] = {R[ctr - GHX] * dcy[0], R[ctr] * dcy[1], R[ctr + GHX] * dcy[2]};
         const double rz[3] = {rc[0] * dczr[0], rc[1] * dczr[1], rc[2] * dczr[2]};
         const double wx = cyj * czk, wy = cxi * czk, wz = cxi * cyj;
         double gr = 0;
         const long p = (long)kk * b.nij + col;
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const double* g = G + c * GPL + ctr;
            const double* d = Dd + c * GPL + ctr;
            double hx[5], hy[5], hz[5];
#pragma unroll
            for (int o = 0; o < 5; o++) {
               hx[o] = sxo[o] * g[o - 2];
               hy[o] = syo[o] * g[(o - 2) * GHX];
               hz[o] = szr[o] * gc[o][c];
            }
            const double ax[3] = {hx[2] - 2 * hx[1] + hx[0], hx[3] - 2 * hx[2] + hx[1], hx[4] - 2 * hx[3] + hx[2]};
            const double ay[3] = {hy[2] - 2 * hy[1] + hy[0], hy[3] - 2 * hy[2] + hy[1], hy[4] - 2 * hy[3] + hy[2]};
            const double az[3] = {hz[2] - 2 * hz[1] + hz[0], hz[3] - 2 * hz[2] + hz[1], hz[4] - 2 * hz[3] + hz[2]};
            const double tx_ = wx * (rx[2] * ax[2] - 2 * rx[1] * ax[1] + rx[0] * ax[0]);
            const double ty_ = wy * (ry[2] * ay[2] - 2 * ry[1] * ay[1] + ry[0] * ay[0]);
            const double tz_ = wz * (rz[2] * az[2] - 2 * rz[1] * az[1] + rz[0] * az[0]);
            dbar[p + n * c] = -(tx_ + ty_ + tz_);
            const double dxx = d[1] - 2 * d[0] + d[-1];
            const double dyy = d[GHX] - 2 * d[0] + d[-GHX];
            const double dzz = dc[2][c] - 2 * dc[1][c] + dc[0][c];
            gr -= dcx[1] * dxx * wx * ax[1] + dcy[1] * dyy * wy * ay[1] + dczr[1] * dzz * wz * az[1];
         }
         grho[p] = grc + gr;
      }
      if (k < k1) store(cur ^ 1);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++) {
#pragma unroll
         for (int c = 0; c < 3; c++) gc[q][c] = gc[q + 1][c];
         szr[q] = szr[q + 1];
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
#pragma unroll
         for (int c = 0; c < 3; c++) dc[q][c] = dc[q + 1][c];
         rc[q] = rc[q + 1];
         dczr[q] = dczr[q + 1];
      }
   }
}

__global__ void k_damp_invrho(DampCtx D, const double* __restrict__ y, const double* __restrict__ u,
                              const double* __restrict__ um, double* __restrict__ grho)
{
   const Box& b = D.b;
   const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = 1 + blockIdx.z;
   if (i > D.nx || j > D.ny) return;
   double dm[3];
   addsgd4_point<1>(b, i, j, k, u, um, D.rho, D.sg, D.beta, dm);
   const long p = sidx(b, i, j, k);
   grho[p] += (y[p] * dm[0] + y[p + b.npts] * dm[1] + y[p + 2 * b.npts] * dm[2]) / D.rho[p];
}

// ---------------------------------------------------------------------------
// Ghost-point boundary conditions, transposed
// ---------------------------------------------------------------------------
struct WinT {
   int i0, i1, j0, j1, k0, k1;
};

__global__ void k_bc_zero_T(Box b, WinT w, double* __restrict__ y)
{
   const long wi = w.i1 - w.i0 + 1, wj = w.j1 - w.j0 + 1, wk = w.k1 - w.k0 + 1;
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= wi * wj * wk) return;
   const int i = w.i0 + (int)(q % wi);
   const int j = w.j0 + (int)((q / wi) % wj);
   const int k = w.k0 + (int)(q / (wi * wj));
   for (int c = 0; c < 3; c++) y[vidx<1>(b, c, i, j, k)] = 0.0;
}

// phase 1: the adjoint of the ghost plane k - kl moves to the plane buffer
__global__ void k_freesurf_take(Box b, int k, int kl, double* __restrict__ y, double* __restrict__ plane)
{
   const int i = b.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie || j > b.je) return;
   const long q = (i - b.ib) + b.ni * (j - b.jb);
   const long nij = b.nij;
   const bool own = i >= b.ib + 2 && i <= b.ie - 2 && j >= b.jb + 2 && j <= b.je - 2;
   for (int c = 0; c < 3; c++) {
      const long g = vidx<1>(b, c, i, j, k - kl);
      plane[q + nij * c] = own ? y[g] : 0.0;
      if (own) y[g] = 0.0;
   }
}

// phase 2: gather into the points the closure reads (freesurf_point,
// ew-cfromfort.C:683-742) and the derivatives of g2 in mu, la at plane k.
__global__ void k_freesurf_push(Box b, int k, int kl, double h, const SbpTables* __restrict__ sbp,
                                const double* __restrict__ plane, const double* __restrict__ fwd,
                                const double* __restrict__ mu, const double* __restrict__ la,
                                const double* __restrict__ strx, const double* __restrict__ stry,
                                double* __restrict__ y, double* __restrict__ gmu, double* __restrict__ gla)
{
   (void)h;
   const int i = b.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie || j > b.je) return;
   const long nij = b.nij;
   const double* sbop = sbp->sbop;
   const double d4a = 2.0 / 3.0, d4b = -1.0 / 12.0;
   auto G = [&](int c, int ii, int jj) {  // ghost adjoint at column (ii,jj), 0 off the closure's columns
      if (ii < b.ib + 2 || ii > b.ie - 2 || jj < b.jb + 2 || jj > b.je - 2) return 0.0;
      return plane[(ii - b.ib) + b.ni * (jj - b.jb) + nij * c];
   };
   auto K = [&](int ii, int jj) {  // d g2 / d ux = -kl la / (2 mu + la) / sbop0
      if (ii < b.ib + 2 || ii > b.ie - 2 || jj < b.jb + 2 || jj > b.je - 2) return 0.0;
      const long pp = sidx(b, ii, jj, k);
      return -kl * la[pp] / (2 * mu[pp] + la[pp]) / sbop[0];
   };
   // column terms: g_c = -sum_m sbop[m] u_c(k + kl (m-1)) / sbop0 + ...
   for (int c = 0; c < 3; c++) {
      const double gc = G(c, i, j);
      if (gc != 0.0)
         for (int m = 1; m <= 4; m++) y[vidx<1>(b, c, i, j, k + kl * (m - 1))] -= sbop[m] / sbop[0] * gc;
   }
   // in-plane first differences at plane k
   double a0 = 0, a1 = 0, a2 = 0;
   const int off[4] = {-2, -1, 1, 2};
   const double cp[4] = {-d4b, -d4a, d4a, d4b};  // coefficient of u(i+off) in d4a(u+1 - u-1) + d4b(u+2 - u-2)
   for (int t = 0; t < 4; t++) {
      const int d = off[t];
      const int ir = i - d, jr = j - d;  // closure columns that read (i,j) at offset d
      const double sxr = ir >= b.ib && ir <= b.ie ? strx[ir - b.ib] : 1.0;
      const double syr = jr >= b.jb && jr <= b.je ? stry[jr - b.jb] : 1.0;
      a2 += -kl / sbop[0] * cp[t] * (G(0, ir, j) * sxr + G(1, i, jr) * syr);
      a0 += K(ir, j) * G(2, ir, j) * sxr * cp[t];
      a1 += K(i, jr) * G(2, i, jr) * syr * cp[t];
   }
   y[vidx<1>(b, 0, i, j, k)] += a0;
   y[vidx<1>(b, 1, i, j, k)] += a1;
   y[vidx<1>(b, 2, i, j, k)] += a2;
   // mu, la derivatives of g2 = (-dz2 - kl la (ux + vy) / (2 mu + la)) / sbop0
   const double g2 = G(2, i, j);
   if (g2 != 0.0) {
      auto U = [&](int c, int ii, int jj) { return fwd[vidx<1>(b, c, ii, jj, k)]; };
      const double sx = strx[i - b.ib], sy = stry[j - b.jb];
      const double ux = sx * (d4a * (U(0, i + 1, j) - U(0, i - 1, j)) + d4b * (U(0, i + 2, j) - U(0, i - 2, j)));
      const double vy = sy * (d4a * (U(1, i, j + 1) - U(1, i, j - 1)) + d4b * (U(1, i, j + 2) - U(1, i, j - 2)));
      const long pp = sidx(b, i, j, k);
      const double m0 = mu[pp], l0 = la[pp], den = 2 * m0 + l0, S = ux + vy;
      gmu[pp] += g2 * kl * l0 * S * 2 / (den * den) / sbop[0];
      gla[pp] += -g2 * kl * S * 2 * m0 / (den * den) / sbop[0];
   }
}

// ---------------------------------------------------------------------------
// pointwise helpers
// ---------------------------------------------------------------------------
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q < n) y[q] += a * x[q];
}

// z = coef/rho y on the interior (0 elsewhere) and, for the transposed
// operator, W = S z on the interior-formula rows (0 elsewhere in the box;
// S = strx stry strz) on the padded box; grho -= gs y.(a - bs)/rho on the
// interior (bs may be null).
__device__ __forceinline__ void vjp_scaled_rows(const AdjCtx& A, const AdjPad& P, long p, double coef, double gs,
                                                const double* __restrict__ rho, const double* __restrict__ y,
                                                const double* __restrict__ a, const double* __restrict__ bs,
                                                double* __restrict__ z, double* __restrict__ grho)
{
   const Box& b = A.b;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const bool in = i >= b.ib + 2 && i <= b.ie - 2 && j >= b.jb + 2 && j <= b.je - 2 && k >= b.kb + 2 && k <= b.ke - 2;
   const bool row = int_row(A.m, i, j, k);
   const double r = rho[p], sc = in ? coef / r : 0.0;
   const double sw = row ? A.sx(i) * A.sy(j) * A.sz(k) : 0.0;
   const long q = sidx(P.pb, i, j, k);
   double g = 0;
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      const double yv = y[v], zv = sc * yv;
      z[v] = zv;
      P.w[q + P.pb.npts * c] = row ? sw * zv : 0.0;
      g += yv * (a[v] - (bs ? bs[v] : 0.0));
   }
   if (in) grho[p] -= gs * g / r;
}

// corrector: z = c4/rho y, W; grho -= y.(u1 - up)/rho
__global__ void k_vjp_corr_pw(AdjCtx A, AdjPad P, double c4, const double* __restrict__ rho,
                              const double* __restrict__ y, const double* __restrict__ u1,
                              const double* __restrict__ up, double* __restrict__ z, double* __restrict__ grho)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= A.b.npts) return;
   vjp_scaled_rows(A, P, p, c4, 1.0, rho, y, u1, up, z, grho);
}

// predictor: z = dt2/rho yp, W; grho -= dt2 yp.acc/rho
__global__ void k_vjp_pred_pw(AdjCtx A, AdjPad P, double dt2, const double* __restrict__ rho,
                              const double* __restrict__ yp, const double* __restrict__ acc, double* __restrict__ z,
                              double* __restrict__ grho)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= A.b.npts) return;
   vjp_scaled_rows(A, P, p, dt2, dt2, rho, yp, acc, nullptr, z, grho);
}

// y += o1/dt2 on the box, o1 the transposed operator's result on the padded box
__global__ void k_vjp_axpy_pad(Box b, Box pb, double dt2i, const double* __restrict__ o1, double* __restrict__ y)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const long q = sidx(pb, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) y[p + b.npts * c] += dt2i * o1[q + pb.npts * c];
}

// with a = o1/dt2, o1 and o2 read on the padded box:
// lb += db - 2a + 2yp + o2, lc = a - db - yp
__global__ void k_vjp_lam(Box b, Box pb, double dt2i, const double* __restrict__ db, const double* __restrict__ o1,
                          const double* __restrict__ yp, const double* __restrict__ o2, double* __restrict__ lb,
                          double* __restrict__ lc)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const long q = sidx(pb, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      const double a = dt2i * o1[q + pb.npts * c], d = db[v], y = yp[v];
      lb[v] += d - 2 * a + 2 * y + o2[q + pb.npts * c];
      lc[v] = a - d - y;
   }
}

__global__ void k_receiver_sources_T(ReceiverAdj R, int s, int nsteps, double twodt, double* __restrict__ lam)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q >= R.nu) return;
   const long n = 3L * R.nrec;
   double v = 0;
   for (int t = R.off[q]; t < R.off[q + 1]; t++) {
      const int e = R.ent[t];
      double w = R.g[s * n + e];
      if (R.mode[e / 3] == 1) {
         w /= twodt;
         if (s + 2 <= nsteps) w -= R.g[(s + 2) * n + e] / twodt;
      }
      v += w;
   }
   lam[R.uidx[q]] += v;
}

// ===========================================================================
// launchers
// ===========================================================================
int launch_adj_pad_material(const Box& b, const double* mu, const double* la, const double* strx, const double* stry,
                            const double* strz, const AdjPad& P, hipStream_t s)
{
   k_adj_pad_material<<<cdiv(P.pb.npts, 256), 256, 0, s>>>(b, P.pb, mu, la, strx, stry, strz, P.mu, P.la, P.sx, P.sy,
                                                           P.sz);
   return check_hip(hipGetLastError(), "adjoint pad launch");
}

int launch_rhs4sg_T(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* mu,
                    const double* la, double h, const double* strx, const double* stry, const double* strz,
                    double* out, const AdjPad& P, hipStream_t s, bool weighted)
{
   const AdjCtx A{b, m, d_sbp, mu, la, strx, stry, strz};
   const double cof = 1.0 / (h * h);
   const dim3 blk(64, 4, 1);
   const unsigned gi = cdiv(b.ni, 64), gj = cdiv(b.je - b.jb + 1, 4);
   // interior rows: S^{-1} L(S zI) on every point of the box, by the LDS
   // predictor sweep over the padded box in its transposed mode (cof L(W)/S;
   // W, mu and lambda vanish in the pad, the stretching is 1 there)
   if (!weighted) k_adj_weight<false><<<cdiv(P.pb.npts, 256), 256, 0, s>>>(A, P, z);
   SgCoef sg{};
   sg.strx = P.sx;
   sg.stry = P.sy;
   sg.strz = P.sz;
   const Box& pb = P.pb;
   const int npl = pb.ke - pb.kb - 3;  // the sweep's 32 x 8 tiles over the box's columns
   const long tiles = cdiv(pb.ni - 4, 32) * cdiv(pb.je - pb.jb - 3, 8);
   int rc = launch_sweep_lds_T(pb, pb.kb + 2, pb.ke - 2, march_kchunk(tiles, npl, 2), P.w, P.mu, P.la, sg, P.o, h, s);
   if (rc) return rc;
   // the result on the box, or left on the padded box (out null)
   const Box ob = out ? b : pb;
   double* const dst = out ? out : P.o;
   if (out) k_adj_crop<<<cdiv(b.npts, 256), 256, 0, s>>>(b, pb, P.o, out);
   // SBP rows: in-plane part from (sx sy) zB, then the k-direction terms
   if (m.top || m.bot) {
      k_adj_weight<true><<<cdiv(P.pb.npts, 256), 256, 0, s>>>(A, P, z);
      if (m.top) k_rhs4sg_T_xy<<<dim3(gi, gj, 6), blk, 0, s>>>(A, P, 1, 1, ob, dst, cof);
      if (m.bot) k_rhs4sg_T_xy<<<dim3(gi, gj, 6), blk, 0, s>>>(A, P, m.nz, -1, ob, dst, cof);
      if (m.top) k_rhs4sg_T_col<<<dim3(gi, gj, 9), blk, 0, s>>>(A, 1, 1, z, ob, dst, cof);
      if (m.bot) k_rhs4sg_T_col<<<dim3(gi, gj, 9), blk, 0, s>>>(A, m.nz, -1, z, ob, dst, cof);
   }
   return check_hip(hipGetLastError(), "rhs4sg transpose launch");
}

int launch_rhs4sg_matgrad(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* u,
                          const double* mu, const double* la, double h, const double* strx, const double* stry,
                          const double* strz, double* gmu, double* gla, hipStream_t s)
{
   const AdjCtx A{b, m, d_sbp, mu, la, strx, stry, strz};
   const double cof = 1.0 / (h * h);
   const dim3 blk(64, 4, 1);
   const unsigned gi = cdiv(b.ni, 64), gj = cdiv(b.je - b.jb + 1, 4);
   {
      const int ntx = (int)cdiv(b.ni, GTX), nty = (int)cdiv(b.je - b.jb + 1, GTY);
      const int kc = march_kchunk((long)ntx * nty, b.ke - b.kb + 1, 2), nkc = (int)cdiv(b.ke - b.kb + 1, kc);
      k_rhs4sg_matgrad_march<<<ntx * nty * nkc, GNT, 0, s>>>(A, ntx, nty, kc, z, u, gmu, gla, cof);
   }
   if (m.top) k_rhs4sg_matgrad_sbp<<<dim3(gi, gj, 8), blk, 0, s>>>(A, 1, 1, z, u, gmu, gla, cof);
   if (m.bot) k_rhs4sg_matgrad_sbp<<<dim3(gi, gj, 8), blk, 0, s>>>(A, m.nz, -1, z, u, gmu, gla, cof);
   return check_hip(hipGetLastError(), "rhs4sg material-derivative launch");
}

std::vector<DampTile> damp_T_tiles(const Box& b, const double* dcx, const double* dcy, const double* dcz)
{
   // the kernel reads the profiles at the point and its two neighbours,
   // clamped to the box; its outputs at a point vanish where all three are 0
   auto damped = [](const double* dc, int n) {
      std::vector<char> on(n, 0);
      for (int q = 0; q < n; q++)
         on[q] = dc[std::max(q - 1, 0)] != 0.0 || dc[q] != 0.0 || dc[std::min(q + 1, n - 1)] != 0.0;
      return on;
   };
   const int ni = (int)b.ni, nj = b.je - b.jb + 1, nk = b.ke - b.kb + 1;
   const std::vector<char> ox = damped(dcx, ni), oy = damped(dcy, nj), oz = damped(dcz, nk);
   const int ntx = (int)cdiv(ni, GTX), nty = (int)cdiv(nj, GTY);
   const int kc = march_kchunk((long)ntx * nty, nk, 2);
   std::vector<DampTile> out;
   for (int kb = b.kb; kb <= b.ke; kb += kc) {
      const int ke = std::min(b.ke, kb + kc - 1);
      for (int tjy = 0; tjy < nty; tjy++)
         for (int tix = 0; tix < ntx; tix++) {
            bool side = false;
            for (int q = tix * GTX; q < std::min((tix + 1) * GTX, ni); q++) side = side || ox[q];
            for (int q = tjy * GTY; q < std::min((tjy + 1) * GTY, nj); q++) side = side || oy[q];
            int lo = ke + 1, hi = kb - 1;
            for (int k = kb; k <= ke; k++)
               if (side || oz[k - b.kb]) {
                  lo = std::min(lo, k);
                  hi = std::max(hi, k);
               }
            if (lo <= hi) out.push_back(DampTile{tix, tjy, lo, hi});
         }
   }
   return out;
}

int launch_addsgd4_T_march(const Box& b, const double* y, const double* u, const double* um, const double* rho,
                           const SgCoef& sg, double beta, double* dbar, double* grho, hipStream_t s, bool inv_rho,
                           const DampTile* tiles, int ntiles)
{
   const DampCtx D{b, b.ie - b.ib - 3, b.je - b.jb - 3, b.ke - b.kb - 3, sg, rho, beta};
   const int ntx = (int)cdiv(b.ni, GTX), nty = (int)cdiv(b.je - b.jb + 1, GTY);
   const int kc = march_kchunk((long)ntx * nty, b.ke - b.kb + 1, 2), nkc = (int)cdiv(b.ke - b.kb + 1, kc);
   if (tiles) {
      if (ntiles > 0) k_addsgd4_T_march<<<ntiles, GNT, 0, s>>>(D, ntx, nty, kc, y, u, um, dbar, grho, tiles);
   } else {
      k_addsgd4_T_march<<<ntx * nty * nkc, GNT, 0, s>>>(D, ntx, nty, kc, y, u, um, dbar, grho, nullptr);
   }
   if (inv_rho) {
      const dim3 blk(64, 4, 1);
      k_damp_invrho<<<dim3(cdiv(D.nx, 64), cdiv(D.ny, 4), D.nz), blk, 0, s>>>(D, y, u, um, grho);
   }
   return check_hip(hipGetLastError(), "addsgd4 transpose launch");
}

int launch_bcfortsg_T(const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd, double* y,
                      const double* fwd, double h, const SbpTables* d_sbp, const double* mu, const double* la,
                      const double* strx, const double* stry, double* gmu, double* gla, double* plane,
                      hipStream_t s)
{
   (void)nx;
   (void)ny;
   for (int side = 5; side >= 0; side--) {
      const int* w = wind + 6 * side;
      const WinT win{w[0], w[1], w[2], w[3], w[4], w[5]};
      const long n = (long)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1);
      if (bccnd[side] == 1 || bccnd[side] == 2) {
         if (n > 0) k_bc_zero_T<<<cdiv(n, 256), 256, 0, s>>>(b, win, y);
      } else if (bccnd[side] == 0) {
         if (side != 4 && side != 5) return set_error(E_INVAL, "bcfortsg transpose: free surface only on sides 4, 5");
         const int k = side == 4 ? 1 : nz, kl = side == 4 ? 1 : -1;
         dim3 blk(64, 4, 1), grd(cdiv(b.ni, 64), cdiv(b.je - b.jb + 1, 4), 1);
         k_freesurf_take<<<grd, blk, 0, s>>>(b, k, kl, y, plane);
         k_freesurf_push<<<grd, blk, 0, s>>>(b, k, kl, h, d_sbp, plane, fwd, mu, la, strx, stry, y, gmu, gla);
      } else if (bccnd[side] == 3) {
         return set_error(E_INVAL, "the adjoint does not support periodic sides");
      }
   }
   return check_hip(hipGetLastError(), "bcfortsg transpose launch");
}

int launch_axpy(long n, double a, const double* x, double* y, hipStream_t s)
{
   if (n > 0) k_axpy<<<cdiv(n, 256), 256, 0, s>>>(n, a, x, y);
   return check_hip(hipGetLastError(), "axpy launch");
}

int launch_vjp_corr_pw(const Box& b, const RowMap& m, const double* strx, const double* stry, const double* strz,
                       double c4, const double* rho, const double* y, const double* u1, const double* up, double* z,
                       double* grho, const AdjPad& P, hipStream_t s)
{
   const AdjCtx A{b, m, nullptr, nullptr, nullptr, strx, stry, strz};
   k_vjp_corr_pw<<<cdiv(b.npts, 256), 256, 0, s>>>(A, P, c4, rho, y, u1, up, z, grho);
   return check_hip(hipGetLastError(), "reverse corrector launch");
}

int launch_vjp_pred_pw(const Box& b, const RowMap& m, const double* strx, const double* stry, const double* strz,
                       double dt2, const double* rho, const double* yp, const double* acc, double* z, double* grho,
                       const AdjPad& P, hipStream_t s)
{
   const AdjCtx A{b, m, nullptr, nullptr, nullptr, strx, stry, strz};
   k_vjp_pred_pw<<<cdiv(b.npts, 256), 256, 0, s>>>(A, P, dt2, rho, yp, acc, z, grho);
   return check_hip(hipGetLastError(), "reverse predictor launch");
}

int launch_vjp_axpy_pad(const Box& b, double dt2i, const AdjPad& P, double* y, hipStream_t s)
{
   k_vjp_axpy_pad<<<cdiv(b.npts, 256), 256, 0, s>>>(b, P.pb, dt2i, P.o, y);
   return check_hip(hipGetLastError(), "reverse step axpy launch");
}

int launch_vjp_lam(const Box& b, double dt2i, const double* db, const AdjPad& P1, const double* yp,
                   const AdjPad& P2, double* lb, double* lc, hipStream_t s)
{
   k_vjp_lam<<<cdiv(b.npts, 256), 256, 0, s>>>(b, P1.pb, dt2i, db, P1.o, yp, P2.o, lb, lc);
   return check_hip(hipGetLastError(), "reverse step adjoint launch");
}

int launch_receiver_sources_T(const ReceiverAdj& R, int s, int nsteps, double twodt, double* lam, hipStream_t st)
{
   if (R.nu <= 0 || s < 1 || s > nsteps) return OK;
   k_receiver_sources_T<<<cdiv(R.nu, 256), 256, 0, st>>>(R, s, nsteps, twodt, lam);
   return check_hip(hipGetLastError(), "receiver adjoint source launch");
}

}  // namespace sw4amd
