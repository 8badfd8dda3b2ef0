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
#include "kernels_adj.h"
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

// Accessor of rhs4_interior_body / rhs4_xy_body at (i,j,k) over S z:
// INPLANE = false: interior rows, S = sx sy sz; true: SBP rows, S = sx sy.
template <bool INPLANE>
struct TSten {
   const AdjCtx& A;
   const double* z;
   int i, j, k;
   __device__ __forceinline__ double u(int c, int di, int dj, int dk) const
   {
      const int ii = i + di, jj = j + dj, kk = k + dk;
      const double v = INPLANE ? A.zb(z, c, ii, jj, kk) : A.zi(z, c, ii, jj, kk);
      return v * A.sx(ii) * A.sy(jj) * (INPLANE ? 1.0 : A.sz(kk));
   }
   __device__ __forceinline__ double mu(int di, int dj, int dk) const { return A.vmu(i + di, j + dj, k + dk); }
   __device__ __forceinline__ double la(int di, int dj, int dk) const { return A.vla(i + di, j + dj, k + dk); }
   __device__ __forceinline__ double sx(int d) const { return A.sx(i + d); }
   __device__ __forceinline__ double sy(int d) const { return A.sy(j + d); }
   __device__ __forceinline__ double sz(int d) const { return A.sz(k + d); }
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
#pragma unroll 1
   for (int kr = 1; kr <= 6; kr++) {
      const int krg = kbase + sgn * (kr - 1);
      if (col) {
         // acof / ghcof: u_c(q) enters row kr of the same column
         const double z0 = A.zb(z, 0, i, j, krg), z1 = A.zb(z, 1, i, j, krg), z2 = A.zb(z, 2, i, j, krg);
         double amu = 0, a2ml = 0;
         if (q == 0) {
            const int k1g = kbase;  // plane m = 1
            const double gh = sbp->ghcof[kr - 1];
            amu = gh * A.vmu(i, j, k1g);
            a2ml = gh * (2 * A.vmu(i, j, k1g) + A.vla(i, j, k1g));
         } else {
#pragma unroll 1
            for (int mm = 1; mm <= 8; mm++) {
               const double a = sbp->acof[(kr - 1) + 6 * (q - 1) + 48 * (mm - 1)];
               const int kmg = kbase + sgn * (mm - 1);
               amu += a * A.vmu(i, j, kmg);
               a2ml += a * (2 * A.vmu(i, j, kmg) + A.vla(i, j, kmg));
            }
         }
         r[0] += z0 * amu;
         r[1] += z1 * amu;
         r[2] += z2 * a2ml;
      }
      if (q < 1) continue;
      const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
      double s2 = 0, s0 = 0, s1 = 0;
      const double lar = A.vla(i, j, krg), mur = A.vmu(i, j, krg);
#pragma unroll
      for (int t = 0; t < 4; t++) {
         const int a = t < 2 ? t - 2 : t - 1;
         const double ca = cd4(a);
         const double wx0 = A.sx(i - a) * A.zb(z, 0, i - a, j, krg);
         const double wy1 = A.sy(j - a) * A.zb(z, 1, i, j - a, krg);
         const double wx2 = A.sx(i - a) * A.zb(z, 2, i - a, j, krg);
         const double wy2 = A.sy(j - a) * A.zb(z, 2, i, j - a, krg);
         // (la w_z)_x, (la w_z)_y: coefficient la at this column, row plane
         s2 += ca * lar * (wx0 + wy1);
         // (mu w_x)_z, (mu w_y)_z: coefficient mu at the row's column, plane q
         s2 += ca * (A.vmu(i - a, j, k) * wx0 + A.vmu(i, j - a, k) * wy1);
         // (mu u_z)_x and (la u_x)_z
         s0 += ca * (mur * wx2 + A.vla(i - a, j, k) * wx2);
         // (mu v_z)_y and (la v_y)_z
         s1 += ca * (mur * wy2 + A.vla(i, j - a, k) * wy2);
      }
      r[0] += bq * i12 * s0;
      r[1] += bq * i12 * s1;
      r[2] += bq * i12 * s2;
   }
}

// W = S zI: the adjoint field on the interior-formula rows times S = sx sy sz
__global__ void k_adj_weight(AdjCtx A, const double* __restrict__ z, double* __restrict__ w)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const Box& b = A.b;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const double s = int_row(A.m, i, j, k) ? A.sx(i) * A.sy(j) * A.sz(k) : 0.0;
   for (int c = 0; c < 3; c++) w[p + b.npts * c] = s * z[p + b.npts * c];
}

// out /= S on the interior points (where the forward interior kernel wrote)
__global__ void k_adj_unweight(AdjCtx A, double* __restrict__ out)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const Box& b = A.b;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   if (i < b.ib + 2 || i > b.ie - 2 || j < b.jb + 2 || j > b.je - 2 || k < b.kb + 2 || k > b.ke - 2) return;
   const double s = 1.0 / (A.sx(i) * A.sy(j) * A.sz(k));
   for (int c = 0; c < 3; c++) out[p + b.npts * c] *= s;
}

// Accessor of rhs4_interior_body over a weighted field W, 0 outside the box.
struct WSten {
   const AdjCtx& A;
   const double* w;
   int i, j, k;
   __device__ __forceinline__ double u(int c, int di, int dj, int dk) const { return A.f(w, c, i + di, j + dj, k + dk); }
   __device__ __forceinline__ double mu(int di, int dj, int dk) const { return A.vmu(i + di, j + dj, k + dk); }
   __device__ __forceinline__ double la(int di, int dj, int dk) const { return A.vla(i + di, j + dj, k + dk); }
   __device__ __forceinline__ double sx(int d) const { return A.sx(i + d); }
   __device__ __forceinline__ double sy(int d) const { return A.sy(j + d); }
   __device__ __forceinline__ double sz(int d) const { return A.sz(k + d); }
};

// the interior-row part of L^T on the ghost shell: cof S^{-1} L(W)
__global__ void __launch_bounds__(256) k_rhs4sg_T_shell(AdjCtx A, const double* __restrict__ w,
                                                        double* __restrict__ out, double cof)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   int i, j, k;
   if (!shell_point(A.b, q, i, j, k)) return;
   const WSten s{A, w, i, j, k};
   double r[3];
   rhs4_interior_body(s, r);
   const double sc = cof / (A.sx(i) * A.sy(j) * A.sz(k));
   const long p = sidx(A.b, i, j, k);
   for (int c = 0; c < 3; c++) out[p + A.b.npts * c] = sc * r[c];
}

// SBP rows of one side: out += in-plane part on the row planes and the k
// terms on the planes 0..8 of the row space (thread per point of those planes)
__global__ void __launch_bounds__(256) k_rhs4sg_T_sbp(AdjCtx A, int kbase, int sgn, const double* __restrict__ z,
                                                      double* __restrict__ out, double cof)
{
   const Box& b = A.b;
   const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (e >= 9 * b.nij) return;
   const int q = (int)(e / b.nij);
   const long r0 = e - q * b.nij;
   const int i = b.ib + (int)(r0 % b.ni), j = b.jb + (int)(r0 / b.ni);
   const int k = kbase + sgn * (q - 1);
   if (k < b.kb || k > b.ke) return;
   double r[3] = {0, 0, 0};
   if (q >= 1 && q <= 6) {
      const TSten<true> s{A, z, i, j, k};
      double t[3];
      rhs4_xy_body(s, t);
      const double wgt = 1.0 / (A.sx(i) * A.sy(j));
      for (int c = 0; c < 3; c++) r[c] += wgt * t[c];
   }
   sbp_col_T(A, z, i, j, k, kbase, sgn, r);
   const long p = sidx(b, i, j, k);
   for (int c = 0; c < 3; c++) out[p + b.npts * c] += cof * r[c];
}

// ---------------------------------------------------------------------------
// Material derivatives
// ---------------------------------------------------------------------------
// Psi_t(f) = sum_q C[q][t] (f[o_q] - f[0]): the factor of a[t] in varsum4 with
// the coefficients varcoef4 builds from a[0..4] (offsets -2..2).
__device__ __forceinline__ double psi(int t, const double f[5])
{
   const double d0 = f[0] - f[2], d1 = f[1] - f[2], d3 = f[3] - f[2], d4v = f[4] - f[2];
   switch (t) {
   case 0: return -0.75 * d0 + d1;
   case 1: return d0 + 3 * d1 + d3;
   case 2: return -0.75 * d0 + 3 * d1 + 3 * d3 - 0.75 * d4v;
   case 3: return d1 + 3 * d3 + d4v;
   default: return d3 - 0.75 * d4v;
   }
}

// Second-difference terms along axis AX of the rows selected by ROWS (0
// interior formula, 1 SBP rows): contribution to (gmu, gla) at p.
template <int AX, int ROWS>
__device__ __forceinline__ void grad_second(const AdjCtx& A, const double* z, const double* u, int i, int j, int k,
                                            double& gm, double& gl)
{
   const double sp = AX == 0 ? A.sx(i) : AX == 1 ? A.sy(j) : A.sz(k);
#pragma unroll
   for (int t = 0; t < 5; t++) {
      const int o = 2 - t;  // row = p + o e_AX
      const int ir = i + (AX == 0) * o, jr = j + (AX == 1) * o, kr = k + (AX == 2) * o;
      const bool ok = ROWS == 0 ? int_row(A.m, ir, jr, kr) : sbp_row(A.m, ir, jr, kr);
      if (!ok) continue;
      const double srow = AX == 0 ? A.sx(ir) : AX == 1 ? A.sy(jr) : A.sz(kr);
      const double w = srow * sp * (1.0 / 6);
#pragma unroll
      for (int c = 0; c < 3; c++) {
         double f[5];
#pragma unroll
         for (int s = 0; s < 5; s++)
            f[s] = A.f(u, c, ir + (AX == 0) * (s - 2), jr + (AX == 1) * (s - 2), kr + (AX == 2) * (s - 2));
         const double val = z[sidx(A.b, ir, jr, kr) + A.b.npts * c] * w * psi(t, f);
         gm += (c == AX ? 2.0 : 1.0) * val;
         if (c == AX) gl += val;
      }
   }
}

// Mixed term (coef D_DB u_C)_DA of row component RC: coef at p.
template <int RC, int DA, int DB, int C, int ROWS>
__device__ __forceinline__ double grad_mixed(const AdjCtx& A, const double* z, const double* u, int i, int j, int k)
{
   double D = 0, W = 0;
#pragma unroll
   for (int t = 0; t < 4; t++) {
      const int a = t < 2 ? t - 2 : t - 1;
      D += cd4(a) * A.f(u, C, i + (DB == 0) * a, j + (DB == 1) * a, k + (DB == 2) * a);
      const int ir = i - (DA == 0) * a, jr = j - (DA == 1) * a, kr = k - (DA == 2) * a;
      const double zr = ROWS == 0 ? A.zi(z, RC, ir, jr, kr) : A.zb(z, RC, ir, jr, kr);
      const double sa = DA == 0 ? A.sx(ir) : DA == 1 ? A.sy(jr) : A.sz(kr);
      const double sb = DB == 0 ? A.sx(ir) : DB == 1 ? A.sy(jr) : A.sz(kr);
      W += cd4(a) * sa * sb * zr;
   }
   return D * W * (1.0 / 144);
}

// acof / ghcof / bope terms of the SBP rows of one side at p
__device__ void sbp_col_grad(const AdjCtx& A, const double* z, const double* u, int i, int j, int k, int kbase,
                             int sgn, double& gm, double& gl)
{
   const int q = sgn * (k - kbase) + 1;
   if (q < 1 || q > 8) return;
   const SbpTables* sbp = A.sbp;
   const double i12 = 1.0 / 12;
   const bool col = i >= 1 && i <= A.m.nx && j >= 1 && j <= A.m.ny;
#pragma unroll 1
   for (int kr = 1; kr <= 6; kr++) {
      const int krg = kbase + sgn * (kr - 1);
      if (col) {
         const double z0 = A.zb(z, 0, i, j, krg), z1 = A.zb(z, 1, i, j, krg), z2 = A.zb(z, 2, i, j, krg);
         // mu, la at plane m = q enter row kr through acof(kr, q', q) u(q')
         double sm = 0, sl = 0;
#pragma unroll 1
         for (int qq = 1; qq <= 8; qq++) {
            const double a = sbp->acof[(kr - 1) + 6 * (qq - 1) + 48 * (q - 1)];
            const int kq = kbase + sgn * (qq - 1);
            const double u0 = A.f(u, 0, i, j, kq), u1 = A.f(u, 1, i, j, kq), u2 = A.f(u, 2, i, j, kq);
            sm += a * (z0 * u0 + z1 * u1 + 2 * z2 * u2);
            sl += a * z2 * u2;
         }
         if (q == 1) {  // ghost term gh mu(1) u(0)
            const double gh = sbp->ghcof[kr - 1];
            const int k0g = kbase - sgn;
            const double u0 = A.f(u, 0, i, j, k0g), u1 = A.f(u, 1, i, j, k0g), u2 = A.f(u, 2, i, j, k0g);
            sm += gh * (z0 * u0 + z1 * u1 + 2 * z2 * u2);
            sl += gh * z2 * u2;
         }
         gm += sm;
         gl += sl;
      }
      // (mu w_x)_z, (mu w_y)_z, (la u_x)_z, (la v_y)_z: coefficient at (i,j) plane q, row (i,j,kr)
      {
         const double bq = sgn * sbp->bope[(kr - 1) + 6 * (q - 1)];
         double dx2 = 0, dy2 = 0, dx0 = 0, dy1 = 0;
#pragma unroll
         for (int t = 0; t < 4; t++) {
            const int a = t < 2 ? t - 2 : t - 1;
            dx2 += cd4(a) * A.f(u, 2, i + a, j, k);
            dy2 += cd4(a) * A.f(u, 2, i, j + a, k);
            dx0 += cd4(a) * A.f(u, 0, i + a, j, k);
            dy1 += cd4(a) * A.f(u, 1, i, j + a, k);
         }
         const double z0 = A.zb(z, 0, i, j, krg), z1 = A.zb(z, 1, i, j, krg), z2 = A.zb(z, 2, i, j, krg);
         gm += bq * i12 * (A.sx(i) * dx2 * z0 + A.sy(j) * dy2 * z1);
         gl += bq * i12 * (A.sx(i) * dx0 + A.sy(j) * dy1) * z2;
      }
   }
   // (la w_z)_x, (la w_z)_y, (mu u_z)_x, (mu v_z)_y: coefficient at p on row plane kr = q
   if (q <= 6) {
      const int kr = q;
      double bu0 = 0, bu1 = 0, bu2 = 0;
#pragma unroll 1
      for (int qq = 1; qq <= 8; qq++) {
         const double bq = sgn * sbp->bope[(kr - 1) + 6 * (qq - 1)];
         const int kq = kbase + sgn * (qq - 1);
         bu0 += bq * A.f(u, 0, i, j, kq);
         bu1 += bq * A.f(u, 1, i, j, kq);
         bu2 += bq * A.f(u, 2, i, j, kq);
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

__global__ void __launch_bounds__(256) k_rhs4sg_matgrad(AdjCtx A, const double* __restrict__ z,
                                                        const double* __restrict__ u, double* __restrict__ gmu,
                                                        double* __restrict__ gla, double cof)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const Box& b = A.b;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   double gm = 0, gl = 0;
   // interior rows
   grad_second<0, 0>(A, z, u, i, j, k, gm, gl);
   grad_second<1, 0>(A, z, u, i, j, k, gm, gl);
   grad_second<2, 0>(A, z, u, i, j, k, gm, gl);
   gl += grad_mixed<0, 0, 1, 1, 0>(A, z, u, i, j, k) + grad_mixed<0, 0, 2, 2, 0>(A, z, u, i, j, k) +
         grad_mixed<1, 1, 0, 0, 0>(A, z, u, i, j, k) + grad_mixed<1, 1, 2, 2, 0>(A, z, u, i, j, k) +
         grad_mixed<2, 2, 0, 0, 0>(A, z, u, i, j, k) + grad_mixed<2, 2, 1, 1, 0>(A, z, u, i, j, k);
   gm += grad_mixed<0, 1, 0, 1, 0>(A, z, u, i, j, k) + grad_mixed<0, 2, 0, 2, 0>(A, z, u, i, j, k) +
         grad_mixed<1, 0, 1, 0, 0>(A, z, u, i, j, k) + grad_mixed<1, 2, 1, 2, 0>(A, z, u, i, j, k) +
         grad_mixed<2, 0, 2, 0, 0>(A, z, u, i, j, k) + grad_mixed<2, 1, 2, 1, 0>(A, z, u, i, j, k);
   // SBP rows: in-plane part, then the k-direction terms
   if (A.m.top || A.m.bot) {
      grad_second<0, 1>(A, z, u, i, j, k, gm, gl);
      grad_second<1, 1>(A, z, u, i, j, k, gm, gl);
      gl += grad_mixed<0, 0, 1, 1, 1>(A, z, u, i, j, k) + grad_mixed<1, 1, 0, 0, 1>(A, z, u, i, j, k);
      gm += grad_mixed<0, 1, 0, 1, 1>(A, z, u, i, j, k) + grad_mixed<1, 0, 1, 0, 1>(A, z, u, i, j, k);
      if (A.m.top) sbp_col_grad(A, z, u, i, j, k, 1, 1, gm, gl);
      if (A.m.bot) sbp_col_grad(A, z, u, i, j, k, A.m.nz, -1, gm, gl);
   }
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
   // g^ax = beta w_ax / rho y at interior points (w_x = strx coy coz, ...)
   __device__ __forceinline__ double g(int ax, const double* y, int c, int i, int j, int k) const
   {
      if (!interior(i, j, k)) return 0.0;
      const int ii = i - b.ib, jj = j - b.jb, kk = k - b.kb;
      const double w = ax == 0 ? sg.strx[ii] * sg.coy[jj] * sg.coz[kk]
                     : ax == 1 ? sg.stry[jj] * sg.cox[ii] * sg.coz[kk]
                               : sg.strz[kk] * sg.cox[ii] * sg.coy[jj];
      const long p = sidx(b, i, j, k);
      return beta * w / rho[p] * y[p + b.npts * c];
   }
   // r^ax = rho dc_ax, 0 outside the box
   __device__ __forceinline__ double r(int ax, int i, int j, int k) const
   {
      if (!inbox(b, i, j, k)) return 0.0;
      const double dc = ax == 0 ? sg.dcx[i - b.ib] : ax == 1 ? sg.dcy[j - b.jb] : sg.dcz[k - b.kb];
      return rho[sidx(b, i, j, k)] * dc;
   }
};

template <int AX>
__device__ __forceinline__ void step3(int& i, int& j, int& k, int d)
{
   if (AX == 0) i += d;
   if (AX == 1) j += d;
   if (AX == 2) k += d;
}

// (D2 g)(l) along AX
template <int AX>
__device__ __forceinline__ double d2g(const DampCtx& D, const double* y, int c, int i, int j, int k)
{
   int im = i, jm = j, km = k, ip = i, jp = j, kp = k;
   step3<AX>(im, jm, km, -1);
   step3<AX>(ip, jp, kp, 1);
   return D.g(AX, y, c, ip, jp, kp) - 2 * D.g(AX, y, c, i, j, k) + D.g(AX, y, c, im, jm, km);
}

__global__ void __launch_bounds__(256) k_addsgd4_T(DampCtx D, const double* __restrict__ y,
                                                   const double* __restrict__ u, const double* __restrict__ um,
                                                   double* __restrict__ dbar, double* __restrict__ grho)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   const Box& b = D.b;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   auto dd = [&](int c, int ii, int jj, int kk) {
      if (!inbox(b, ii, jj, kk)) return 0.0;
      const long q = sidx(b, ii, jj, kk) + b.npts * c;
      return u[q] - um[q];
   };
   double gr = 0;
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double acc = 0;
      // D^T y = sum_ax D2 (r D2 g) at p; rho derivative through r at p
      {
         double s = 0;
         s += D.r(0, i + 1, j, k) * d2g<0>(D, y, c, i + 1, j, k) - 2 * D.r(0, i, j, k) * d2g<0>(D, y, c, i, j, k) +
              D.r(0, i - 1, j, k) * d2g<0>(D, y, c, i - 1, j, k);
         s += D.r(1, i, j + 1, k) * d2g<1>(D, y, c, i, j + 1, k) - 2 * D.r(1, i, j, k) * d2g<1>(D, y, c, i, j, k) +
              D.r(1, i, j - 1, k) * d2g<1>(D, y, c, i, j - 1, k);
         s += D.r(2, i, j, k + 1) * d2g<2>(D, y, c, i, j, k + 1) - 2 * D.r(2, i, j, k) * d2g<2>(D, y, c, i, j, k) +
              D.r(2, i, j, k - 1) * d2g<2>(D, y, c, i, j, k - 1);
         acc = s;
      }
      dbar[p + b.npts * c] = -acc;
      const double d0 = dd(c, i, j, k);
      const double dxx = dd(c, i + 1, j, k) - 2 * d0 + dd(c, i - 1, j, k);
      const double dyy = dd(c, i, j + 1, k) - 2 * d0 + dd(c, i, j - 1, k);
      const double dzz = dd(c, i, j, k + 1) - 2 * d0 + dd(c, i, j, k - 1);
      gr -= D.sg.dcx[i - b.ib] * dxx * d2g<0>(D, y, c, i, j, k) + D.sg.dcy[j - b.jb] * dyy * d2g<1>(D, y, c, i, j, k) +
            D.sg.dcz[k - b.kb] * dzz * d2g<2>(D, y, c, i, j, k);
   }
   if (D.interior(i, j, k)) {
      double dm[3];
      addsgd4_point<1>(b, i, j, k, u, um, D.rho, D.sg, D.beta, dm);
      const double rp = D.rho[p];
      gr += (y[p] * dm[0] + y[p + b.npts] * dm[1] + y[p + 2 * b.npts] * dm[2]) / rp;
   }
   grho[p] += gr;
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

__global__ void k_rows_scale(Box b, double coef, const double* __restrict__ rho, const double* __restrict__ y,
                             double* __restrict__ z)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   const bool in = i >= b.ib + 2 && i <= b.ie - 2 && j >= b.jb + 2 && j <= b.je - 2 && k >= b.kb + 2 && k <= b.ke - 2;
   const double s = in ? coef / rho[p] : 0.0;
   for (int c = 0; c < 3; c++) z[p + b.npts * c] = s * y[p + b.npts * c];
}

__global__ void k_rho_grad(Box b, double coef, const double* __restrict__ y, const double* __restrict__ lu,
                           const double* __restrict__ f, const double* __restrict__ rho, double* __restrict__ grho)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   double s = 0;
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      s += y[q] * (lu[q] + f[q]);
   }
   grho[p] -= coef * s / (rho[p] * rho[p]);
}

__global__ void k_scatter_add(int n, const int* __restrict__ idx, const double* __restrict__ val, double* __restrict__ y)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q < n) y[idx[q]] += val[q];
}

// ===========================================================================
// launchers
// ===========================================================================
int launch_rhs4sg_T(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* mu,
                    const double* la, double h, const double* strx, const double* stry, const double* strz,
                    double* out, double* work, hipStream_t s)
{
   const AdjCtx A{b, m, d_sbp, mu, la, strx, stry, strz};
   const double cof = 1.0 / (h * h);
   // interior rows: S^{-1} L(S zI), the interior formula on every interior
   // point (the forward kernel, no SBP closure) and on the ghost shell
   k_adj_weight<<<cdiv(b.npts, 256), 256, 0, s>>>(A, z, work);
   const int plain[6] = {0, 0, 0, 0, 0, 0};
   int rc = launch_rhs4sg(1, b, m.nz, plain, d_sbp, out, work, mu, la, h, strx, stry, strz, s);
   if (rc) return rc;
   k_adj_unweight<<<cdiv(b.npts, 256), 256, 0, s>>>(A, out);
   const long ns = shell_count(b);
   k_rhs4sg_T_shell<<<cdiv(ns, 256), 256, 0, s>>>(A, work, out, cof);
   // SBP rows
   if (m.top) k_rhs4sg_T_sbp<<<cdiv(9 * b.nij, 256), 256, 0, s>>>(A, 1, 1, z, out, cof);
   if (m.bot) k_rhs4sg_T_sbp<<<cdiv(9 * b.nij, 256), 256, 0, s>>>(A, m.nz, -1, z, out, cof);
   return check_hip(hipGetLastError(), "rhs4sg transpose launch");
}

int launch_rhs4sg_matgrad(const Box& b, const RowMap& m, const SbpTables* d_sbp, const double* z, const double* u,
                          const double* mu, const double* la, double h, const double* strx, const double* stry,
                          const double* strz, double* gmu, double* gla, hipStream_t s)
{
   const AdjCtx A{b, m, d_sbp, mu, la, strx, stry, strz};
   k_rhs4sg_matgrad<<<cdiv(b.npts, 256), 256, 0, s>>>(A, z, u, gmu, gla, 1.0 / (h * h));
   return check_hip(hipGetLastError(), "rhs4sg material-derivative launch");
}

int launch_addsgd4_T(const Box& b, const double* y, const double* u, const double* um, const double* rho,
                     const SgCoef& sg, double beta, double* dbar, double* grho, hipStream_t s)
{
   const DampCtx D{b, b.ie - b.ib - 3, b.je - b.jb - 3, b.ke - b.kb - 3, sg, rho, beta};
   k_addsgd4_T<<<cdiv(b.npts, 256), 256, 0, s>>>(D, y, u, um, dbar, grho);
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

int launch_rows_scale(const Box& b, double coef, const double* rho, const double* y, double* z, hipStream_t s)
{
   k_rows_scale<<<cdiv(b.npts, 256), 256, 0, s>>>(b, coef, rho, y, z);
   return check_hip(hipGetLastError(), "row scaling launch");
}

int launch_rho_grad(const Box& b, double coef, const double* y, const double* lu, const double* f,
                    const double* rho, double* grho, hipStream_t s)
{
   k_rho_grad<<<cdiv(b.npts, 256), 256, 0, s>>>(b, coef, y, lu, f, rho, grho);
   return check_hip(hipGetLastError(), "density derivative launch");
}

int launch_scatter_add(int n, const int* idx, const double* val, double* y, hipStream_t s)
{
   if (n > 0) k_scatter_add<<<cdiv(n, 128), 128, 0, s>>>(n, idx, val, y);
   return check_hip(hipGetLastError(), "scatter launch");
}

}  // namespace sw4amd
