// kernels_curv.hip -- gfx950 kernels of the curvilinear (topography-following)
// grid, behind the reference's kernel-level entry points:
//   metric / metric_rev            curvilinear-c.C:53-172, :175-295
//   rhs4sgcurv / rhs4sgcurv_rev    rhs4sgcurv.C:34-1406, rhs4sgcurv_rev.C:34-1395
//   addsgd4cfort(_indrev)          ew-cfromfort.C:1160-1240, :1323-1404
//   freesurfcurvisg(_rev)          curvilinear-c.C:465-618, :621-764
// Both Sarray layouts through the CO template parameter; the metric is a
// 4-component array in the same layout as the displacement.
//
// The divergence of stress on the curvilinear grid, with M1..M4 the metric,
// K = 2mu + lambda, sx = strx(i), sy = stry(j), D_p, D_q centred 4th-order
// first differences along i and j and D_r the k-derivative of the row (SBP
// bope on the top six rows, centred inside), is L(u)_a = sx sy / jac * r_a,
//   r_a = pp + qq + rr + pq + qp + pr + qr + rp + rq,
//   pp = (1/6)(A u_a,p)_p / sy with A = (a==0 ? K : mu) M1^2 strx,
//   qq = (1/6)(B u_a,q)_q / sx with B = (a==1 ? K : mu) M1^2 stry,
//   rr = sum_b S_ab (R_ab u_b,r)_r, R the 3x3 coupling of mu, lambda and the
//        metric, S the stretching scalings; acof/ghcof SBP rows at the top,
//   pq, qp = centred mixed products in the plane (u and v equations),
//   pr, qr = D_r of T_a(D_p u, D_q u) on the neighbouring planes,
//   rp, rq = D_p of P_a(D_r u), D_q of Q_a(D_r u) on the neighbouring columns.
// The terms are those of the reference; the grouping of sums is this file's.
#include "kernels_common.h"
#include "kernels_curv.h"

namespace sw4amd {

namespace {
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}  // namespace

constexpr double CC1 = 2.0 / 3, CC2 = -1.0 / 12;

template <int CO>
__device__ __forceinline__ long midx(const Box& b, int c, int i, int j, int k)
{
   const long p = sidx(b, i, j, k);
   return CO ? p + b.npts * c : 4 * p + c;
}

// ---------------------------------------------------------------------------
// Metric
// ---------------------------------------------------------------------------
// Derivative along a grid line of n points at position s (0-based): centred
// 4th order inside, 6-point one-sided closures on the first and last two
// points (curvilinear-c.C:78-146).
__device__ __forceinline__ double line_derivative(const double* __restrict__ f, long st, int s, int n)
{
#define F(t) f[(long)(t) * st]
   if (s >= 2 && s <= n - 3) return CC2 * (F(s + 2) - F(s - 2)) + CC1 * (F(s + 1) - F(s - 1));
   if (s == 0)
      return -2.25 * F(0) + (4 + 5.0 / 6) * F(1) - (14.0 / 3) * F(2) + 3 * F(3) - (1 + 1.0 / 12) * F(4) +
             (1.0 / 6) * F(5);
   if (s == 1)
      return -(1.0 / 6) * F(0) - 1.25 * F(1) + (1 + 4.0 / 3) * F(2) - (4.0 / 3) * F(3) + 0.5 * F(4) -
             (1.0 / 12) * F(5);
   if (s == n - 2)
      return (1.0 / 6) * F(n - 1) + 1.25 * F(n - 2) - (1 + 4.0 / 3) * F(n - 3) + (4.0 / 3) * F(n - 4) -
             0.5 * F(n - 5) + (1.0 / 12) * F(n - 6);
   return 2.25 * F(n - 1) - (4 + 5.0 / 6) * F(n - 2) + (14.0 / 3) * F(n - 3) - 3 * F(n - 4) +
          (1 + 1.0 / 12) * F(n - 5) - (1.0 / 6) * F(n - 6);
#undef F
}

template <int CO>
__global__ void __launch_bounds__(256)
k_metric(Box b, double h, const double* __restrict__ z, double* __restrict__ met, double* __restrict__ jac,
         int* __restrict__ bad)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int ni = (int)b.ni, nj = b.je - b.jb + 1, nk = b.ke - b.kb + 1;
   const int ii = (int)(p % ni);
   const long r = p / ni;
   const int jj = (int)(r % nj), kk = (int)(r / nj);
   const double zr = line_derivative(z + (p - b.nij * kk), b.nij, kk, nk);
   const double zq = line_derivative(z + (p - b.ni * jj), b.ni, jj, nj);
   const double zp = line_derivative(z + (p - ii), 1, ii, ni);
   if (!(zr > 0)) {
      bad[0] = 1;
      return;
   }
   const double sq = sqrt(zr);
   jac[p] = h * h * zr;
   const long m0 = CO ? p : 4 * p, ms = CO ? b.npts : 1;
   met[m0] = sq;
   met[m0 + ms] = -zp / sq;
   met[m0 + 2 * ms] = -zq / sq;
   met[m0 + 3 * ms] = h / sq;
}

// ---------------------------------------------------------------------------
// Divergence of stress
// ---------------------------------------------------------------------------
template <int CO>
struct CurvG {
   Box b;
   const double* __restrict__ u;
   const double* __restrict__ mu;
   const double* __restrict__ la;
   const double* __restrict__ met;
   const double* __restrict__ strx;
   const double* __restrict__ stry;
   __device__ __forceinline__ double U(int c, int i, int j, int k) const { return u[vidx<CO>(b, c, i, j, k)]; }
   __device__ __forceinline__ double M(int c, int i, int j, int k) const { return met[midx<CO>(b, c, i, j, k)]; }
   __device__ __forceinline__ double MU(int i, int j, int k) const { return mu[sidx(b, i, j, k)]; }
   __device__ __forceinline__ double LA(int i, int j, int k) const { return la[sidx(b, i, j, k)]; }
   __device__ __forceinline__ double SX(int i) const { return strx[i - b.ib]; }
   __device__ __forceinline__ double SY(int j) const { return stry[j - b.jb]; }
   __device__ __forceinline__ double Dp(int c, int i, int j, int k) const
   {
      return CC2 * (U(c, i + 2, j, k) - U(c, i - 2, j, k)) + CC1 * (U(c, i + 1, j, k) - U(c, i - 1, j, k));
   }
   __device__ __forceinline__ double Dq(int c, int i, int j, int k) const
   {
      return CC2 * (U(c, i, j + 2, k) - U(c, i, j - 2, k)) + CC1 * (U(c, i, j + 1, k) - U(c, i, j - 1, k));
   }
};

__device__ __forceinline__ void vc4(const double a[5], double m[4])
{
   m[0] = a[1] - 0.75 * (a[2] + a[0]);
   m[1] = a[0] + a[3] + 3 * (a[2] + a[1]);
   m[2] = a[1] + a[4] + 3 * (a[3] + a[2]);
   m[3] = a[3] - 0.75 * (a[2] + a[4]);
}
__device__ __forceinline__ double vs4(const double m[4], const double f[5])
{
   return m[0] * (f[0] - f[2]) + m[1] * (f[1] - f[2]) + m[2] * (f[3] - f[2]) + m[3] * (f[4] - f[2]);
}

// rr coupling at (i,j,m): R[0..5] = 00, 11, 22, 01, 02, 12
template <class S>
__device__ __forceinline__ void rr_coef(const S& s, int i, int j, int m, double sx, double sy, double R[6])
{
   const double mu = s.MU(i, j, m), la = s.LA(i, j, m);
   const double m2 = s.M(1, i, j, m), m3 = s.M(2, i, j, m), m4 = s.M(3, i, j, m);
   const double x2 = m2 * sx * m2 * sx, y2 = m3 * sy * m3 * sy, z2 = m4 * m4;
   R[0] = (2 * mu + la) * x2 + mu * (y2 + z2);
   R[1] = (2 * mu + la) * y2 + mu * (x2 + z2);
   R[2] = (2 * mu + la) * z2 + mu * (x2 + y2);
   R[3] = (mu + la) * m2 * m3;
   R[4] = (mu + la) * m2 * m4;
   R[5] = (mu + la) * m3 * m4;
}

// traction-like combination of in-plane derivatives on plane m, column (i,j):
// the pr/qr terms of the operator and the tangential part of the free-
// surface traction (curvilinear-c.C:520-578)
template <class S>
__device__ __forceinline__ void t_terms(const S& s, int i, int j, int m, double sx, double sy, double isx,
                                        double isy, double T[3])
{
   const double mu = s.MU(i, j, m), la = s.LA(i, j, m), K = 2 * mu + la;
   const double m1 = s.M(0, i, j, m), m2 = s.M(1, i, j, m), m3 = s.M(2, i, j, m), m4 = s.M(3, i, j, m);
   const double p0 = s.Dp(0, i, j, m), p1 = s.Dp(1, i, j, m), p2 = s.Dp(2, i, j, m);
   const double q0 = s.Dq(0, i, j, m), q1 = s.Dq(1, i, j, m), q2 = s.Dq(2, i, j, m);
   const double a2 = m2 * m1, a3 = m3 * m1, a4 = m4 * m1;
   T[0] = K * a2 * p0 * sx * isy + mu * a3 * p1 + mu * a4 * p2 * isy + mu * a3 * q0 * sy * isx + la * a2 * q1;
   T[1] = la * a3 * p0 + mu * a2 * p1 * sx * isy + mu * a2 * q0 + K * a3 * q1 * sy * isx + mu * a4 * q2 * isx;
   T[2] = la * a4 * p0 * isy + mu * a2 * p2 * sx * isy + mu * a3 * q2 * sy * isx + la * a4 * q1 * isx;
}

template <bool BND, class S>
__device__ __forceinline__ void curv_point(const S& s, int i, int j, int k, const SbpTables* __restrict__ sbp,
                                           double r[3])
{
   const double i6 = 1.0 / 6;
   const double sx = s.SX(i), sy = s.SY(j);
   const double isx = 1 / sx, isy = 1 / sy, isxy = isy * isx;
   const int off[4] = {-2, -1, 1, 2};
   const double wc[4] = {-CC2, -CC1, CC1, CC2};
   r[0] = r[1] = r[2] = 0;

   // pp and qq
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double a[5], m[4], f[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
         const int ii = i + t - 2;
         const double mu = s.MU(ii, j, k), m1 = s.M(0, ii, j, k);
         a[t] = (c == 0 ? 2 * mu + s.LA(ii, j, k) : mu) * m1 * m1 * s.SX(ii);
         f[t] = s.U(c, ii, j, k);
      }
      vc4(a, m);
      r[c] += i6 * vs4(m, f) * isy;
#pragma unroll
      for (int t = 0; t < 5; t++) {
         const int jj = j + t - 2;
         const double mu = s.MU(i, jj, k), m1 = s.M(0, i, jj, k);
         a[t] = (c == 1 ? 2 * mu + s.LA(i, jj, k) : mu) * m1 * m1 * s.SY(jj);
         f[t] = s.U(c, i, jj, k);
      }
      vc4(a, m);
      r[c] += i6 * vs4(m, f) * isx;
   }

   // rr
   const double S3[3][3] = {{isxy, 1, isy}, {1, isxy, isx}, {isy, isx, isxy}};
   const int PAIR[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
   if (!BND) {
      double R[5][6];
#pragma unroll
      for (int t = 0; t < 5; t++) rr_coef(s, i, j, k + t - 2, sx, sy, R[t]);
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
         for (int d = 0; d < 3; d++) {
            double a[5], m[4], f[5];
#pragma unroll
            for (int t = 0; t < 5; t++) {
               a[t] = R[t][PAIR[c][d]];
               f[t] = s.U(d, i, j, k + t - 2);
            }
            vc4(a, m);
            r[c] += i6 * vs4(m, f) * S3[c][d];
         }
   } else {
      // C_ab(q) = sum_m acof(k,q,m) R_ab(m), planes 1..8, then the ghost term
#pragma unroll 1
      for (int q = 1; q <= 8; q++) {
         double Cq[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
         for (int mm = 1; mm <= 8; mm++) {
            const double a = sbp->acof[(k - 1) + 6 * (q - 1) + 48 * (mm - 1)];
            if (a == 0) continue;
            double R[6];
            rr_coef(s, i, j, mm, sx, sy, R);
#pragma unroll
            for (int e = 0; e < 6; e++) Cq[e] += a * R[e];
         }
#pragma unroll
         for (int c = 0; c < 3; c++)
#pragma unroll
            for (int d = 0; d < 3; d++) r[c] += S3[c][d] * Cq[PAIR[c][d]] * s.U(d, i, j, q);
      }
      double R[6];
      rr_coef(s, i, j, 1, sx, sy, R);
      const double g = sbp->ghcof[k - 1];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
         for (int d = 0; d < 3; d++) r[c] += S3[c][d] * (g * R[PAIR[c][d]]) * s.U(d, i, j, 0);
   }

   // pq and qp (u and v equations)
#pragma unroll
   for (int t = 0; t < 4; t++) {
      const int d = off[t];
      const double mq = s.M(0, i, j + d, k), mp = s.M(0, i + d, j, k);
      r[0] += wc[t] * (s.MU(i, j + d, k) * mq * mq * s.Dp(1, i, j + d, k) +
                       s.LA(i + d, j, k) * mp * mp * s.Dq(1, i + d, j, k));
      r[1] += wc[t] * (s.LA(i, j + d, k) * mq * mq * s.Dp(0, i, j + d, k) +
                       s.MU(i + d, j, k) * mp * mp * s.Dq(0, i + d, j, k));
   }

   // k-derivative weights of this row
   constexpr int NQ = BND ? 8 : 4;
   int kq[NQ];
   double wq[NQ];
#pragma unroll
   for (int e = 0; e < NQ; e++) {
      kq[e] = BND ? e + 1 : k + off[e];
      wq[e] = BND ? sbp->bope[(k - 1) + 6 * e] : wc[e];
   }

   // pr and qr
#pragma unroll
   for (int e = 0; e < NQ; e++) {
      double T[3];
      t_terms(s, i, j, kq[e], sx, sy, isx, isy, T);
#pragma unroll
      for (int c = 0; c < 3; c++) r[c] += wq[e] * T[c];
   }

   // rp and rq
#pragma unroll
   for (int t = 0; t < 4; t++) {
      const int d = off[t];
      double dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0};
#pragma unroll
      for (int e = 0; e < NQ; e++)
#pragma unroll
         for (int c = 0; c < 3; c++) {
            dx[c] += wq[e] * s.U(c, i + d, j, kq[e]);
            dy[c] += wq[e] * s.U(c, i, j + d, kq[e]);
         }
      {
         const int ii = i + d;
         const double mu = s.MU(ii, j, k), la = s.LA(ii, j, k), sxi = s.SX(ii);
         const double m1 = s.M(0, ii, j, k), m2 = s.M(1, ii, j, k), m3 = s.M(2, ii, j, k), m4 = s.M(3, ii, j, k);
         r[0] += wc[t] * ((2 * mu + la) * m2 * m1 * sxi * dx[0] + la * m3 * m1 * dx[1] * sy + la * m4 * m1 * dx[2]) *
                 isy;
         r[1] += wc[t] * (mu * m3 * m1 * dx[0] + mu * m2 * m1 * dx[1] * sxi * isy);
         r[2] += wc[t] * (mu * m4 * m1 * dx[0] + mu * m2 * m1 * dx[2] * sxi) * isy;
      }
      {
         const int jj = j + d;
         const double mu = s.MU(i, jj, k), la = s.LA(i, jj, k), syj = s.SY(jj);
         const double m1 = s.M(0, i, jj, k), m2 = s.M(1, i, jj, k), m3 = s.M(2, i, jj, k), m4 = s.M(3, i, jj, k);
         r[0] += wc[t] * (mu * m3 * m1 * dy[0] * syj * isx + mu * m2 * m1 * dy[1]);
         r[1] += wc[t] * (la * m2 * m1 * dy[0] + (2 * mu + la) * m3 * m1 * dy[1] * syj * isx +
                          la * m4 * m1 * dy[2] * isx);
         r[2] += wc[t] * (mu * m3 * m1 * dy[2] * syj + mu * m4 * m1 * dy[1]) * isx;
      }
   }
}

// Component-blocked layout: capped at 256 registers for 2 waves/SIMD (2-4
// spilled; 262 registers and 1 wave/SIMD without the cap).  The interleaved
// layout's index arithmetic would spill ~120 registers under the cap.
template <int CO, bool BND>
__global__ void __launch_bounds__(256, CO ? 2 : 1)
k_rhs4curv(Box b, int klo, int khi, const SbpTables* __restrict__ sbp, CurvG<CO> s,
           const double* __restrict__ jac, double* __restrict__ lu)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = klo + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > khi) return;
   double r[3];
   curv_point<BND>(s, i, j, k, sbp, r);
   const double ij = s.SX(i) * s.SY(j) / jac[sidx(b, i, j, k)];
#pragma unroll
   for (int c = 0; c < 3; c++) lu[vidx<CO>(b, c, i, j, k)] = r[c] * ij;
}

// ---------------------------------------------------------------------------
// Supergrid damping in x and y with rho -> rho jac
// ---------------------------------------------------------------------------
template <int CO>
__global__ void __launch_bounds__(256)
k_addsgd4c(Box b, double* __restrict__ up, const double* __restrict__ u, const double* __restrict__ um,
           const double* __restrict__ rho, const double* __restrict__ jac, SgCoef sg, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = b.kb + 2 + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const int ii = i - b.ib, jj = j - b.jb;
   // every term carries one of these damping coefficients: where all six
   // vanish (the columns away from the x and y supergrid layers) the update
   // subtracts zero, and the point is left as it is
   if (sg.dcx[ii - 1] == 0 && sg.dcx[ii] == 0 && sg.dcx[ii + 1] == 0 && sg.dcy[jj - 1] == 0 && sg.dcy[jj] == 0 &&
       sg.dcy[jj + 1] == 0)
      return;
   const long p = sidx(b, i, j, k);
   const double w = beta / (rho[p] * jac[p]);
   double rx[3], ry[3];
#pragma unroll
   for (int t = 0; t < 3; t++) {
      const long px = sidx(b, i + t - 1, j, k), py = sidx(b, i, j + t - 1, k);
      rx[t] = rho[px] * sg.dcx[ii + t - 1] * jac[px];
      ry[t] = rho[py] * sg.dcy[jj + t - 1] * jac[py];
   }
   const double wx = sg.strx[ii] * sg.coy[jj], wy = sg.stry[jj] * sg.cox[ii];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double dx[5], dy[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
         long q = vidx<CO>(b, c, i + t - 2, j, k);
         dx[t] = u[q] - um[q];
         q = vidx<CO>(b, c, i, j + t - 2, k);
         dy[t] = u[q] - um[q];
      }
      const double ax = rx[2] * (dx[4] - 2 * dx[3] + dx[2]) - 2 * rx[1] * (dx[3] - 2 * dx[2] + dx[1]) +
                        rx[0] * (dx[2] - 2 * dx[1] + dx[0]);
      const double ay = ry[2] * (dy[4] - 2 * dy[3] + dy[2]) - 2 * ry[1] * (dy[3] - 2 * dy[2] + dy[1]) +
                        ry[0] * (dy[2] - 2 * dy[1] + dy[0]);
      up[vidx<CO>(b, c, i, j, k)] -= w * (wx * ax + wy * ay);
   }
}

// ---------------------------------------------------------------------------
// Traction-free ghost points on the curved boundary plane
// ---------------------------------------------------------------------------
// The normal stress on the plane is N (D_r u) + g with N = mu ac I +
// (mu+lambda) v v^T, v = (sqrt(sx/sy) M2, sqrt(sy/sx) M3, M4/sqrt(sx sy)),
// ac = |v|^2, g the tangential part minus the forcing.  D_r u = s0 u_ghost +
// sum_w s_w u_w, so u_ghost = -(sum_w s_w u_w + N^-1 g)/s0 with
// N^-1 = I/(mu ac) - (mu+lambda)/((2mu+lambda) mu ac^2) v v^T.
template <int CO>
__global__ void __launch_bounds__(256)
k_freesurf_curv(Box b, int k, int kl, CurvG<CO> s, const SbpTables* __restrict__ sbp,
                const double* __restrict__ forcing, double* __restrict__ u)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const double sx = s.SX(i), sy = s.SY(j), isx = 1 / sx, isy = 1 / sy;
   double g[3];
   t_terms(s, i, j, k, sx, sy, isx, isy, g);
   if (forcing) {
      const long f = 3 * ((i - b.ib) + b.ni * (j - b.jb));
#pragma unroll
      for (int c = 0; c < 3; c++) g[c] -= forcing[f + c];
   }
   const double m2 = s.M(1, i, j, k), m3 = s.M(2, i, j, k), m4 = s.M(3, i, j, k);
   const double mu = s.MU(i, j, k), la = s.LA(i, j, k);
   const double ac = sx * isy * m2 * m2 + sy * isx * m3 * m3 + m4 * m4 * isy * isx;
   const double bc = 1 / (mu * ac);
   const double cc = (mu + la) / (2 * mu + la) * bc / ac;
   const double xy = sqrt(sx * isy);
   const double v[3] = {xy * m2, m3 / xy, isx * xy * m4};
   const double dc = cc * (v[0] * g[0] + v[1] * g[1] + v[2] * g[2]);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      double acc = 0;
#pragma unroll
      for (int w = 1; w <= 4; w++) acc += sbp->sbop[w] * s.U(c, i, j, k + kl * (w - 1));
      u[vidx<CO>(b, c, i, j, k - kl)] = -(acc + bc * g[c] - dc * v[c]) / sbp->sbop[0];
   }
}

// ---------------------------------------------------------------------------
// Time step of the curvilinear grid (EW::computeDT, EW.C:5063-5136)
// ---------------------------------------------------------------------------
// Per point the symmetric matrix A (EW.C:5089-5102) of the curvilinear
// operator's principal part; the time step follows from its most negative
// eigenvalue: max over points of -lambda_min(A)/(4 rho).  The reference calls
// LAPACK dspev; here the eigenvalue comes from the closed form for symmetric
// 3x3 matrices (O. K. Smith, "Eigenvalues of a symmetric 3x3 matrix",
// Comm. ACM 4(4):168, 1961).
constexpr int CRED = 256;

__device__ __forceinline__ double sym3_min_eig(double a11, double a12, double a13, double a22, double a23,
                                               double a33)
{
   const double p1 = a12 * a12 + a13 * a13 + a23 * a23;
   const double q = (a11 + a22 + a33) / 3;
   if (p1 == 0) return fmin(a11, fmin(a22, a33));
   const double b11 = a11 - q, b22 = a22 - q, b33 = a33 - q;
   const double p2 = b11 * b11 + b22 * b22 + b33 * b33 + 2 * p1;
   const double p = sqrt(p2 / 6);
   const double ip = 1 / p;
   // r = det(B)/2 with B = (A - qI)/p
   const double det = b11 * (b22 * b33 - a23 * a23) - a12 * (a12 * b33 - a23 * a13) + a13 * (a12 * a23 - b22 * a13);
   const double r = fmin(1.0, fmax(-1.0, det * ip * ip * ip / 2));
   const double phi = acos(r) / 3;
   // eigenvalues q + 2p cos(phi + 2 pi m / 3); the smallest is m = 1
   return q + 2 * p * cos(phi + 2.0943951023931954923);
}

template <int CO>
__global__ void __launch_bounds__(CRED)
k_maxeig_curv(Box b, const double* __restrict__ met, const double* __restrict__ jac, const double* __restrict__ mu,
              const double* __restrict__ la, const double* __restrict__ rho, double* __restrict__ partial)
{
   __shared__ double sh[CRED];
   double m = -1;
   for (long p = blockIdx.x * (long)CRED + threadIdx.x; p < b.npts; p += (long)gridDim.x * CRED) {
      const long m0 = CO ? p : 4 * p, ms = CO ? b.npts : 1;
      const double m1 = met[m0], m2 = met[m0 + ms], m3 = met[m0 + 2 * ms], m4 = met[m0 + 3 * ms];
      const double l = la[p], u = mu[p], l2 = l + 2 * u, ji = 1 / jac[p];
      const double s1 = m1 * m1, s2 = m2 * m2, s3 = m3 * m3, s4 = m4 * m4;
      const double a11 = -4 * (s1 * l2 + s1 * u + s2 * l2 + s3 * u + s4 * u) * ji;
      const double a12 = -4 * m2 * m3 * (u + l) * ji;
      const double a13 = -4 * m2 * m4 * (u + l) * ji;
      const double a22 = -4 * (s1 * u + s1 * l2 + s2 * u + s3 * l2 + s4 * u) * ji;
      const double a23 = -4 * m3 * m4 * (u + l) * ji;
      const double a33 = -4 * (s1 * u + s1 * u + s2 * u + s3 * u + s4 * l2) * ji;
      const double w0 = sym3_min_eig(a11, a12, a13, a22, a23, a33);
      m = fmax(m, -w0 / (4 * rho[p]));
   }
   sh[threadIdx.x] = m;
   __syncthreads();
   for (int o = CRED / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      __syncthreads();
   }
   if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(CRED) k_final_max_curv(int n, const double* __restrict__ partial,
                                                         double* __restrict__ out)
{
   __shared__ double sh[CRED];
   double m = -1;
   for (int q = threadIdx.x; q < n; q += CRED) m = fmax(m, partial[q]);
   sh[threadIdx.x] = m;
   __syncthreads();
   for (int o = CRED / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      __syncthreads();
   }
   if (threadIdx.x == 0) out[0] = sh[0];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int launch_metric(int co, const Box& b, double h, const double* z, double* met, double* jac, int* bad,
                  hipStream_t s)
{
   const unsigned g = cdiv(b.npts, 256);
   if (co)
      k_metric<1><<<g, 256, 0, s>>>(b, h, z, met, jac, bad);
   else
      k_metric<0><<<g, 256, 0, s>>>(b, h, z, met, jac, bad);
   return check_hip(hipGetLastError(), "metric launch");
}

template <int CO>
static void rhs4curv_co(const Box& b, const int* onesided, const SbpTables* d_sbp, const CurvG<CO>& cg,
                        const double* jac, double* lu, hipStream_t s)
{
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
   dim3 blk(64, 4, 1);
   const int top = onesided[4] == 1;
   const int k1 = top ? 7 : b.kb + 2, k2 = b.ke - 2;
   if (top) k_rhs4curv<CO, true><<<dim3(cdiv(nii, 64), cdiv(njj, 4), 6), blk, 0, s>>>(b, 1, 6, d_sbp, cg, jac, lu);
   if (k2 >= k1)
      k_rhs4curv<CO, false><<<dim3(cdiv(nii, 64), cdiv(njj, 4), k2 - k1 + 1), blk, 0, s>>>(b, k1, k2, d_sbp, cg, jac,
                                                                                           lu);
}

int launch_rhs4sgcurv(int co, const Box& b, const int* onesided, const SbpTables* d_sbp, const double* u,
                      const double* mu, const double* la, const double* met, const double* jac, const double* strx,
                      const double* stry, double* lu, hipStream_t s)
{
   if (co)
      rhs4curv_co<1>(b, onesided, d_sbp, CurvG<1>{b, u, mu, la, met, strx, stry}, jac, lu, s);
   else
      rhs4curv_co<0>(b, onesided, d_sbp, CurvG<0>{b, u, mu, la, met, strx, stry}, jac, lu, s);
   return check_hip(hipGetLastError(), "rhs4sgcurv launch");
}

int launch_addsgd4c(int co, const Box& b, double* up, const double* u, const double* um, const double* rho,
                    const double* jac, const SgCoef& sg, double beta, hipStream_t s)
{
   if (beta == 0) return OK;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3, nkk = b.ke - b.kb - 3;
   dim3 grd(cdiv(nii, 64), cdiv(njj, 4), nkk), blk(64, 4, 1);
   if (co)
      k_addsgd4c<1><<<grd, blk, 0, s>>>(b, up, u, um, rho, jac, sg, beta);
   else
      k_addsgd4c<0><<<grd, blk, 0, s>>>(b, up, u, um, rho, jac, sg, beta);
   return check_hip(hipGetLastError(), "addsgd4c launch");
}

int launch_freesurf_curv(int co, const Box& b, int nz, int side, double* u, const double* mu, const double* la,
                         const double* met, const SbpTables* d_sbp, const double* forcing, const double* strx,
                         const double* stry, hipStream_t s)
{
   const int k = side == 4 ? 1 : nz, kl = side == 4 ? 1 : -1;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
   dim3 grd(cdiv(nii, 64), cdiv(njj, 4), 1), blk(64, 4, 1);
   if (co)
      k_freesurf_curv<1><<<grd, blk, 0, s>>>(b, k, kl, CurvG<1>{b, u, mu, la, met, strx, stry}, d_sbp, forcing, u);
   else
      k_freesurf_curv<0><<<grd, blk, 0, s>>>(b, k, kl, CurvG<0>{b, u, mu, la, met, strx, stry}, d_sbp, forcing, u);
   return check_hip(hipGetLastError(), "curvilinear free-surface launch");
}

template <int CO>
__global__ void k_enforce_cart_topo(Box cart, double* __restrict__ ucart, Box curv, double* __restrict__ ucurv, int g)
{
   const int i = cart.ib + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = cart.jb + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > cart.ie || j > cart.je) return;
   const int kbc = cart.kb, kec = curv.ke;
   for (int c = 0; c < 3; c++) {
      for (int q = 0; q < g; q++) ucart[vidx<CO>(cart, c, i, j, kbc + q)] = ucurv[vidx<CO>(curv, c, i, j, kec - 2 * g + q)];
      for (int q = 0; q <= g; q++) ucurv[vidx<CO>(curv, c, i, j, kec - q)] = ucart[vidx<CO>(cart, c, i, j, kbc + 2 * g - q)];
   }
}

int launch_enforce_cart_topo(int co, const Box& cart, double* ucart, const Box& curv, double* ucurv, int g,
                             hipStream_t s)
{
   const dim3 blk(64, 4, 1), grd(cdiv(cart.ni, 64), cdiv(cart.je - cart.jb + 1, 4), 1);
   if (co)
      k_enforce_cart_topo<1><<<grd, blk, 0, s>>>(cart, ucart, curv, ucurv, g);
   else
      k_enforce_cart_topo<0><<<grd, blk, 0, s>>>(cart, ucart, curv, ucurv, g);
   return check_hip(hipGetLastError(), "grid interface launch");
}

int launch_maxeig_curv(int co, const Box& b, const double* met, const double* jac, const double* mu, const double* la,
                       const double* rho, double* partial, int npartial, double* out, hipStream_t s)
{
   if (co)
      k_maxeig_curv<1><<<npartial, CRED, 0, s>>>(b, met, jac, mu, la, rho, partial);
   else
      k_maxeig_curv<0><<<npartial, CRED, 0, s>>>(b, met, jac, mu, la, rho, partial);
   k_final_max_curv<<<1, CRED, 0, s>>>(npartial, partial, out);
   return check_hip(hipGetLastError(), "curvilinear maxeig launch");
}

}  // namespace sw4amd
