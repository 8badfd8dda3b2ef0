// kernels_step.hip -- kernels of the fused time step and its helpers:
//   * predictor sweep:  Up  = 2U - Um + dt^2/rho L(U)            (+ sources)
//   * corrector sweep:  Up' = Up + dt^4/(12 rho) L((Up-2U+Um)/dt^2)
//                             - supergrid damping(U, Um)           (+ sources)
//     i.e. evalRHS+evalPredictor (EW.C:2569-2584) and evalDpDmInTime+
//     evalRHS+evalCorrector+addSuperGridDamping (EW.C:2662-2715) in two
//     passes over HBM instead of eight;
//   * ghost-shell updates that reproduce what the unfused path leaves in the
//     ghost points;
//   * the sparse point-source update, the CFL reduction (EW::computeDT,
//     EW.C:5041-5146), the analytic point-source solution
//     (EW::get_exact_point_source, EW.C:3899-4523) and the error norms
//     (EW::normOfDifference, EW.C:4526-4610 / solerr3fort).
#include "kernels_common.h"
#include "kernels_fast.h"
#include "kernels_step.h"

namespace sw4amd {

namespace {
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}

// ---------------------------------------------------------------------------
// Fused sweeps, thread per point (component-blocked layout)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_pred_interior(Box b, int klo, int khi, const double* __restrict__ u, const double* __restrict__ um,
                const double* __restrict__ mu, const double* __restrict__ la,
                const double* __restrict__ rho, const double* __restrict__ strx,
                const double* __restrict__ stry, const double* __restrict__ strz,
                double* __restrict__ up, double cof, double dt2)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = klo + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > khi) return;
   double r[3];
   rhs_interior_point<1>(b, i, j, k, Plain{u}, mu, la, strx, stry, strz, r);
   const long p = sidx(b, i, j, k);
   const double s = dt2 / rho[p];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      up[q] = 2 * u[q] - um[q] + s * (cof * r[c]);
   }
}

__global__ void __launch_bounds__(256)
k_pred_sbp(Box b, int top, int nk, const SbpTables* __restrict__ sbp, const double* __restrict__ u,
           const double* __restrict__ um, const double* __restrict__ mu,
           const double* __restrict__ la, const double* __restrict__ rho,
           const double* __restrict__ strx, const double* __restrict__ stry,
           double* __restrict__ up, double cof, double dt2)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int kr = 1 + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const int k = top ? kr : nk - kr + 1;
   double r[3];
   rhs_sbp_point<1>(b, i, j, k, kr, top ? 1 : nk, top ? 1 : -1, sbp, Plain{u}, mu, la, strx, stry, r);
   const long p = sidx(b, i, j, k);
   const double s = dt2 / rho[p];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      up[q] = 2 * u[q] - um[q] + s * (cof * r[c]);
   }
}

__global__ void __launch_bounds__(256)
k_corr_interior(Box b, int klo, int khi, const double* __restrict__ up, const double* __restrict__ u,
                const double* __restrict__ um, const double* __restrict__ mu,
                const double* __restrict__ la, const double* __restrict__ rho, SgCoef sg,
                double* __restrict__ upo, double cof, double dt2i, double dt4i12, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = klo + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > khi) return;
   double r[3], d[3];
   rhs_interior_point<1>(b, i, j, k, Accel{up, u, um, dt2i}, mu, la, sg.strx, sg.stry, sg.strz, r);
   addsgd4_point<1>(b, i, j, k, u, um, rho, sg, beta, d);
   const long p = sidx(b, i, j, k);
   const double s = dt4i12 / rho[p];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      upo[q] = (up[q] + s * (cof * r[c])) - d[c];
   }
}

// Acceleration (up - 2u + um)/dt^2 on planes ks..ks+SLAB_PLANES-1, all (i,j)
// of the box, component-blocked: the values the SBP rows of the corrector
// read (evalDpDmInTime restricted to those planes, same arithmetic).
constexpr int SLAB_PLANES = 9;
__global__ void __launch_bounds__(256)
k_accel_slab(Box b, int ks, const double* __restrict__ up, const double* __restrict__ u,
             const double* __restrict__ um, double dt2i, double* __restrict__ acc)
{
   const long ns = b.nij * SLAB_PLANES;
   const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (e >= 3 * ns) return;
   const int c = (int)(e / ns);
   const long r = e - c * ns;
   const long q = r + b.nij * (ks - b.kb) + b.npts * c;
   acc[e] = dt2i * (up[q] - 2 * u[q] + um[q]);
}

__global__ void __launch_bounds__(256)
k_corr_sbp(Box b, int top, int nk, const SbpTables* __restrict__ sbp, const double* __restrict__ up,
           const double* __restrict__ acc, int ks, const double* __restrict__ u, const double* __restrict__ um,
           const double* __restrict__ mu, const double* __restrict__ la, const double* __restrict__ rho, SgCoef sg,
           double* __restrict__ upo, double cof, double dt4i12, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int kr = 1 + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const int k = top ? kr : nk - kr + 1;
   double r[3], d[3];
   const Slab a{acc, b.npts, b.nij * (ks - b.kb), b.nij * SLAB_PLANES};
   rhs_sbp_point<1>(b, i, j, k, kr, top ? 1 : nk, top ? 1 : -1, sbp, a, mu, la, sg.strx, sg.stry, r);
   addsgd4_point<1>(b, i, j, k, u, um, rho, sg, beta, d);
   const long p = sidx(b, i, j, k);
   const double s = dt4i12 / rho[p];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      upo[q] = (up[q] + s * (cof * r[c])) - d[c];
   }
}

// Supergrid damping of the corrector sweep on planes klo..khi (the planes the
// LDS corrector covers): up -= damping(u, um) (ew-cfromfort.C:985-1041).
__global__ void __launch_bounds__(256)
k_addsgd4_planes(Box b, int klo, int khi, double* __restrict__ up, const double* __restrict__ u,
                 const double* __restrict__ um, const double* __restrict__ rho, SgCoef sg, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = klo + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > khi) return;
   double d[3];
   addsgd4_point<1>(b, i, j, k, u, um, rho, sg, beta, d);
   const long p = sidx(b, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) up[p + b.npts * c] -= d[c];
}

// ---------------------------------------------------------------------------
// Ghost shell (shell_point, kernels_common.h).  mode 0: up = 2u - um (the
// predictor where L(u) and F vanish); mode 1: out = in.
// ---------------------------------------------------------------------------
__global__ void k_shell(Box b, int mode, const double* __restrict__ a, const double* __restrict__ u,
                        const double* __restrict__ um, double* __restrict__ out, int zfill,
                        double* __restrict__ acc = nullptr, double dt2i = 0)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   int i, j, k;
   if (!shell_point(b, q, i, j, k)) return;
   // zfill (bit s): side s takes zero ghost values (Dirichlet or supergrid,
   // no data) in this pass instead of in the boundary conditions
   const bool zero = ((zfill & 1) && i <= b.ib + 1) || ((zfill & 2) && i >= b.ie - 1) ||
                     ((zfill & 4) && j <= b.jb + 1) || ((zfill & 8) && j >= b.je - 1) ||
                     ((zfill & 16) && k <= b.kb + 1) || ((zfill & 32) && k >= b.ke - 1);
   const long p = sidx(b, i, j, k);
   if (mode == 0 && acc) {
      // predictor with the stored acceleration: also (up - 2u + um)/dt^2 here,
      // the expression the corrector stages
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const long v = p + b.npts * c;
         const double uv = u[v], mv = um[v], o = zero ? 0.0 : 2 * uv - mv;
         out[v] = o;
         acc[v] = dt2i * (o - 2 * uv + mv);
      }
      return;
   }
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      out[v] = zero ? 0.0 : mode == 0 ? 2 * u[v] - um[v] : a[v];
   }
}

long shell_count(const Box& b)
{
   const long ni = b.ni, nj = b.je - b.jb + 1, nk = b.ke - b.kb + 1;
   return 4 * ni * nj + 4 * ni * (nk - 4) + 4 * (nj - 4) * (nk - 4);
}

// ---------------------------------------------------------------------------
// Sparse source update: up(:,p) += coef/rho(p) * sum_s f_s g_s(t)
// ---------------------------------------------------------------------------
__global__ void k_src_apply(Box b, SourceTable st, double t, int deriv2, double coef,
                            const double* __restrict__ rho, double* __restrict__ up, const double* __restrict__ tdev)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= st.nunique) return;
   if (tdev) t = *tdev;
   const int s0 = st.first[r], s1 = st.first[r + 1];
   const int i = st.ijk[3 * s0], j = st.ijk[3 * s0 + 1], k = st.ijk[3 * s0 + 2];
   double f0 = 0, f1 = 0, f2 = 0;
   for (int s = s0; s < s1; s++) {
      const double g = time_function(st.tfkind[s], deriv2, st.freq[s], t - st.t0[s]);
      f0 += st.force[3 * s] * g;
      f1 += st.force[3 * s + 1] * g;
      f2 += st.force[3 * s + 2] * g;
   }
   const long p = sidx(b, i, j, k);
   const double sc = coef / rho[p];
   up[p] += sc * f0;
   up[p + b.npts] += sc * f1;
   up[p + 2 * b.npts] += sc * f2;
}

__global__ void k_accel_points(Box b, SourceTable st, const double* __restrict__ up, const double* __restrict__ u,
                               const double* __restrict__ um, double dt2i, double* __restrict__ acc)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= st.nunique) return;
   const int s0 = st.first[r];
   const long p = sidx(b, st.ijk[3 * s0], st.ijk[3 * s0 + 1], st.ijk[3 * s0 + 2]);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      acc[q] = dt2i * (up[q] - 2 * u[q] + um[q]);  // the expression the corrector stages
   }
}

int launch_accel_points(const Box& b, const SourceTable& st, const double* up, const double* u, const double* um,
                        double dt, double* acc, hipStream_t s)
{
   if (st.nunique <= 0) return OK;
   k_accel_points<<<cdiv(st.nunique, 256), 256, 0, s>>>(b, st, up, u, um, 1.0 / (dt * dt), acc);
   return check_hip(hipGetLastError(), "acceleration at source points");
}

void fused_planes(const Box& b, int nk, const int* onesided, int* k1, int* k2)
{
   *k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   *k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
}

// ---------------------------------------------------------------------------
// Deterministic two-stage reductions
// ---------------------------------------------------------------------------
constexpr int RED_BLOCK = 256;

__device__ __forceinline__ double block_max(double v, double* sh)
{
   sh[threadIdx.x] = v;
   __syncthreads();
   for (int o = RED_BLOCK / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      __syncthreads();
   }
   const double r = sh[0];
   __syncthreads();
   return r;
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
   sh[threadIdx.x] = v;
   __syncthreads();
   for (int o = RED_BLOCK / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
   }
   const double r = sh[0];
   __syncthreads();
   return r;
}

// max over the whole box of (4 mu + la)/rho  (EW.C:5050-5058)
__global__ void __launch_bounds__(RED_BLOCK)
k_maxeig(long n, const double* __restrict__ mu, const double* __restrict__ la,
         const double* __restrict__ rho, double* __restrict__ partial)
{
   __shared__ double sh[RED_BLOCK];
   double m = -1;
   for (long q = blockIdx.x * (long)RED_BLOCK + threadIdx.x; q < n; q += (long)gridDim.x * RED_BLOCK)
      m = fmax(m, (4 * mu[q] + la[q]) / rho[q]);
   m = block_max(m, sh);
   if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

__global__ void __launch_bounds__(RED_BLOCK) k_final_max(int n, const double* __restrict__ partial, double* __restrict__ out)
{
   __shared__ double sh[RED_BLOCK];
   double m = -1;
   for (int q = threadIdx.x; q < n; q += RED_BLOCK) m = fmax(m, partial[q]);
   m = block_max(m, sh);
   if (threadIdx.x == 0) out[0] = m;
}

// ---------------------------------------------------------------------------
// Analytic whole-space solution of a moment-tensor point source
// ---------------------------------------------------------------------------
// Time functions of the exact solution, with s = time - R/c and "time" the
// scaled time of EW.C:3929-3934.  The C6 bump and its antiderivative moments
// are evaluated in Horner form; the Gaussian uses erf.
__device__ __forceinline__ double c6_g(double s)
{
   if (!(s > 0 && s < 1)) return 0.0;
   const double p = s * (1 - s), p2 = p * p, p4 = p2 * p2;
   return 51480.0 * (p4 * p2 * p);
}
__device__ __forceinline__ double c6_gp(double s)
{
   if (!(s > 0 && s < 1)) return 0.0;
   const double p = s * (1 - s), p2 = p * p;
   return 51480.0 * 7 * (1 - 2 * s) * (p2 * p2 * p2);
}
// G1(x) = int_0^x g, G2(x) = int_0^x s g(s) ds for g = 51480 s^7 (1-s)^7
// (unit mass, mean 1/2), x in [0,1].  The binomial series
// (1-s)^7 = sum_m (-1)^m C(7,m) s^m alternates, so it is only summed for
// arguments <= 1/2; larger x use g(s) = g(1-s):
//   G1(x) = 1 - G1(1-x),  G2(x) = 1/2 - G1(1-x) + G2(1-x).
__device__ __forceinline__ void c6_moments(double x, double& g1, double& g2)
{
   const bool flip = x > 0.5;
   const double y = flip ? 1 - x : x;  // exact for x in [1/2,1]
   const double binom[8] = {1, 7, 21, 35, 35, 21, 7, 1};
   double a1 = 0, a2 = 0;
   for (int m = 7; m >= 0; m--) {  // Horner over y^m/(8+m) and y^m/(9+m)
      const double sg = (m & 1) ? -binom[m] : binom[m];
      a1 = a1 * y + sg / (8 + m);
      a2 = a2 * y + sg / (9 + m);
   }
   const double y2 = y * y, y4 = y2 * y2, y8 = y4 * y4;
   const double G1 = 51480.0 * y8 * a1, G2 = 51480.0 * y8 * y * a2;
   if (flip) {
      g1 = 1 - G1;
      g2 = 0.5 - G1 + G2;
   } else {
      g1 = G1;
      g2 = G2;
   }
}
// int_{lo}^{hi} tau g(time - tau) dtau, clipped to the support of g
__device__ __forceinline__ double c6_tau_integral(double time, double R, double alpha, double beta)
{
   const double lo = fmax(R / alpha, time - 1), hi = fmin(R / beta, time);
   if (!(lo < time && hi > time - 1)) return 0.0;
   // = int_{time-hi}^{time-lo} (time - s) g(s) ds
   double g1a, g2a, g1b, g2b;
   c6_moments(time - hi, g1a, g2a);
   c6_moments(time - lo, g1b, g2b);
   return time * (g1b - g1a) - (g2b - g2a);
}

__device__ void exact_moment(const ExactParams& P, double x, double y, double z, double out[3])
{
   const double X[3] = {x - P.x0, y - P.y0, z - P.z0};
   const double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
   out[0] = out[1] = out[2] = 0;
   if (R < P.eps) return;
   const double al2 = P.alpha * P.alpha, be2 = P.beta * P.beta;
   double A, B, C, D, E;
   if (P.tfkind == 2) {
      A = c6_g(P.time - R / P.alpha);
      B = c6_g(P.time - R / P.beta);
      C = c6_tau_integral(P.time, R, P.alpha, P.beta);
      D = c6_gp(P.time - R / P.alpha) / (al2 * P.alpha) / R;
      E = c6_gp(P.time - R / P.beta) / (be2 * P.beta) / R;
   } else {
      const double f = P.fr, nrm = 1 / (f * sqrt(2 * M_PI));
      const double sa = P.time - R / P.alpha, sb = P.time - R / P.beta;
      A = nrm * exp(-sa * sa / (2 * f * f));
      B = nrm * exp(-sb * sb / (2 * f * f));
      C = -0.5 * P.time * (erf(sb / (sqrt(2.0) * f)) - erf(sa / (sqrt(2.0) * f))) -
          f / sqrt(2 * M_PI) * (exp(-sb * sb / (2 * f * f)) - exp(-sa * sa / (2 * f * f)));
      D = -A * sa / (f * f) / (al2 * P.alpha) / R;
      E = -B * sb / (f * f) / (be2 * P.beta) / R;
   }
   const double Aa = A / al2, Bb = B / be2, AmB = Aa - Bb;
   const double r2 = R * R, ir = 1 / R, ir3 = ir / r2, ir5 = ir3 / r2, ir7 = ir5 / r2;
   // u_i = -1/(4 pi rho) sum_kj M_kj d_j G_ik, G_ik = T1 + T2 + T3:
   //  T1 = (3 X_i X_k/r^5 - d_ik/r^3) C,  T2 = X_i X_k/r^3 A/alpha^2,
   //  T3 = -(X_i X_k/r^3 - d_ik/r) B/beta^2,
   //  d_j C = X_j (B/beta^2 - A/alpha^2), d_j(A/alpha^2) = -X_j D, d_j(B/beta^2) = -X_j E
   const double Mt[3][3] = {{P.m[0], P.m[1], P.m[2]}, {P.m[1], P.m[3], P.m[4]}, {P.m[2], P.m[4], P.m[5]}};
   for (int i = 0; i < 3; i++) {
      double acc = 0;
      for (int k = 0; k < 3; k++)
         for (int j = 0; j < 3; j++) {
            const double dij = i == j, dkj = k == j, dik = i == k;
            const double xx = X[i] * X[k];
            const double dxx = dij * X[k] + dkj * X[i];
            const double t1 = (3 * dxx * ir5 - 15 * xx * X[j] * ir7 + 3 * dik * X[j] * ir5) * C -
                              (3 * xx * ir5 - dik * ir3) * X[j] * AmB;
            const double t2 = (dxx * ir3 - 3 * xx * X[j] * ir5) * Aa - xx * ir3 * X[j] * D;
            const double t3 = -(dxx * ir3 - 3 * xx * X[j] * ir5 + dik * X[j] * ir3) * Bb +
                              (xx * ir3 - dik * ir) * X[j] * E;
            acc += Mt[k][j] * (t1 + t2 + t3);
         }
      out[i] = -acc / (4 * M_PI * P.rho);
   }
}

__global__ void k_exact(Box b, ExactParams P, double* __restrict__ out)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= b.npts) return;
   const int i = b.ib + (int)(p % b.ni);
   const int j = b.jb + (int)((p / b.ni) % (b.je - b.jb + 1));
   const int k = b.kb + (int)(p / b.nij);
   double e[3];
   exact_moment(P, (i - 1) * P.h, (j - 1) * P.h, (k - 1) * P.h + P.zmin, e);
   out[p] = e[0];
   out[p + b.npts] = e[1];
   out[p + 2 * b.npts] = e[2];
}

// Linf, h^3 L2 and max(uex) of u - uex over a window, excluding a ball
// (solerr3fort, ew-cfromfort.C:144-202).  Partials per block, 3 per block.
__global__ void __launch_bounds__(RED_BLOCK)
k_solerr(Box b, NormWindow w, const double* __restrict__ uex, const double* __restrict__ u,
         double* __restrict__ partial)
{
   __shared__ double sh[RED_BLOCK];
   const long wi = w.imax - w.imin + 1, wj = w.jmax - w.jmin + 1, wk = w.kmax - w.kmin + 1;
   const long n = wi * wj * wk;
   const double h3 = w.h * w.h * w.h;
   double li = 0, l2 = 0, xl = 0;
   for (long q = blockIdx.x * (long)RED_BLOCK + threadIdx.x; q < n; q += (long)gridDim.x * RED_BLOCK) {
      const int i = w.imin + (int)(q % wi), j = w.jmin + (int)((q / wi) % wj), k = w.kmin + (int)(q / (wi * wj));
      const double dx = (i - 1) * w.h - w.x0, dy = (j - 1) * w.h - w.y0, dz = (k - 1) * w.h + w.zmin - w.z0;
      if (dx * dx + dy * dy + dz * dz <= w.r2) continue;
      const long p = sidx(b, i, j, k);
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const double ex = uex[p + b.npts * c], e = u[p + b.npts * c] - ex;
         li = fmax(li, fabs(e));
         xl = fmax(xl, ex);
         l2 += h3 * e * e;
      }
   }
   li = block_max(li, sh);
   xl = block_max(xl, sh);
   l2 = block_sum(l2, sh);
   if (threadIdx.x == 0) {
      partial[3 * blockIdx.x] = li;
      partial[3 * blockIdx.x + 1] = l2;
      partial[3 * blockIdx.x + 2] = xl;
   }
}

__global__ void __launch_bounds__(RED_BLOCK) k_final_norm(int n, const double* __restrict__ partial, double* __restrict__ out)
{
   __shared__ double sh[RED_BLOCK];
   double li = 0, l2 = 0, xl = 0;
   for (int q = threadIdx.x; q < n; q += RED_BLOCK) {
      li = fmax(li, partial[3 * q]);
      l2 += partial[3 * q + 1];
      xl = fmax(xl, partial[3 * q + 2]);
   }
   li = block_max(li, sh);
   xl = block_max(xl, sh);
   l2 = block_sum(l2, sh);
   if (threadIdx.x == 0) {
      out[0] = li;
      out[1] = l2;
      out[2] = xl;
   }
}

// ---------------------------------------------------------------------------
// Receivers, EW::extractRecordData (EW.C:5854-5875): displacement (mode 0) or
// centred velocity (up - um)/(2 dt) (mode 1) at the nearest grid point.
// rec[4r..4r+3] = i, j, k, mode; out[3r+c] for this sample.
// ---------------------------------------------------------------------------
__global__ void k_record(Box b, int nrec, const int* __restrict__ rec, const double* __restrict__ up,
                         const double* __restrict__ um, double dt, double* __restrict__ out,
                         const int* __restrict__ sdev, int maxs)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= nrec) return;
   if (sdev) {  // sample index from the device clock
      const int smp = *sdev;
      if (smp >= maxs) return;
      out += (size_t)3 * nrec * smp;
   }
   const long p = sidx(b, rec[4 * r], rec[4 * r + 1], rec[4 * r + 2]);
   const int mode = rec[4 * r + 3];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      out[3 * r + c] = mode == 1 ? (up[q] - um[q]) / (2 * dt) : up[q];
   }
}

// ===========================================================================
// launchers
// ===========================================================================
int launch_record(const Box& b, int nrec, const int* rec, const double* up, const double* um, double dt,
                  double* out, hipStream_t s, const int* sdev, int maxs)
{
   if (nrec <= 0) return OK;
   k_record<<<cdiv(nrec, 64), 64, 0, s>>>(b, nrec, rec, up, um, dt, out, sdev, maxs);
   return check_hip(hipGetLastError(), "receiver launch");
}

__global__ void k_clock_tick(int n, double* __restrict__ t, int* __restrict__ smp, const double* __restrict__ dt,
                             const int* __restrict__ maxs)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q >= n) return;
   t[q] += dt[q];
   if (smp[q] < maxs[q]) smp[q] += 1;
}

int launch_clock_tick(int n, double* t, int* smp, const double* dt, const int* maxs, hipStream_t s)
{
   if (n > 0) k_clock_tick<<<cdiv(n, 64), 64, 0, s>>>(n, t, smp, dt, maxs);
   return check_hip(hipGetLastError(), "clock launch");
}

int launch_pred_fused(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                      const double* u, const double* um, const double* mu, const double* la,
                      const double* rho, const SgCoef& sg, double* up, double h, double dt,
                      hipStream_t s, const std::function<int()>& after_interior, int variant, int kchunk,
                      int zfill, double* accf)
{
   const double cof = 1.0 / (h * h), dt2 = dt * dt;
   const int k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   const int k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
   dim3 blk(64, 4, 1);
   if (k2 >= k1 && variant >= 1) {
      const int rc = launch_sweep_lds(false, variant, b, k1, k2, kchunk, nullptr, u, um, mu, la, rho, sg, up, h, dt,
                                      0.0, s, accf);
      if (rc) return rc;
   } else if (k2 >= k1) {
      dim3 grd(cdiv(nii, 64), cdiv(njj, 4), k2 - k1 + 1);
      k_pred_interior<<<grd, blk, 0, s>>>(b, k1, k2, u, um, mu, la, rho, sg.strx, sg.stry, sg.strz, up, cof, dt2);
   }
   if (after_interior) {
      const int rc = after_interior();
      if (rc) return rc;
   }
   if (variant >= 1)
      return launch_pred_tail(b, nk, onesided, d_sbp, u, um, mu, la, rho, sg, up, h, dt, s, zfill, accf);
   for (int side = 4; side <= 5; side++) {
      if (onesided[side] != 1) continue;
      dim3 grd(cdiv(nii, 64), cdiv(njj, 4), 6);
      k_pred_sbp<<<grd, blk, 0, s>>>(b, side == 4, nk, d_sbp, u, um, mu, la, rho, sg.strx, sg.stry, up, cof, dt2);
   }
   const long ns = shell_count(b);
   k_shell<<<cdiv(ns, 256), 256, 0, s>>>(b, 0, nullptr, u, um, up, zfill);
   return check_hip(hipGetLastError(), "fused predictor launch");
}

long accel_slab_size(const Box& b) { return 3 * b.nij * SLAB_PLANES; }

int launch_pred_tail(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp, const double* u,
                     const double* um, const double* mu, const double* la, const double* rho, const SgCoef& sg,
                     double* up, double h, double dt, hipStream_t s, int zfill, double* accf)
{
   for (int side = 4; side <= 5; side++) {  // LDS slab of the six rows
      if (onesided[side] != 1) continue;
      const int rc = launch_sbp_rows(false, b, side, nk, d_sbp, nullptr, u, um, mu, la, rho, sg, up, h, dt, 0.0, s);
      if (rc) return rc;
   }
   const long ns = shell_count(b);
   k_shell<<<cdiv(ns, 256), 256, 0, s>>>(b, 0, nullptr, u, um, up, zfill, accf, 1.0 / (dt * dt));
   return check_hip(hipGetLastError(), "predictor rows and shell launch");
}

int launch_corr_tail(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp, const double* up,
                     const double* u, const double* um, const double* mu, const double* la, const double* rho,
                     const SgCoef& sg, double beta, double* upo, double h, double dt, hipStream_t s,
                     int zfill, const double* accf, bool split_damp)
{
   for (int side = 4; side <= 5; side++) {  // LDS slab, acceleration formed while staging (or stored: accf)
      if (onesided[side] != 1) continue;
      int rc = launch_sbp_rows(true, b, side, nk, d_sbp, up, u, um, mu, la, rho, sg, upo, h, dt,
                               split_damp ? 0.0 : beta, s, accf);
      if (rc == OK && split_damp) rc = launch_rows_damp(b, side, nk, upo, u, um, rho, sg, beta, s);
      if (rc) return rc;
   }
   const long ns = shell_count(b);
   k_shell<<<cdiv(ns, 256), 256, 0, s>>>(b, 1, up, nullptr, nullptr, upo, zfill);
   return check_hip(hipGetLastError(), "corrector rows and shell launch");
}

int launch_corr_fused(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                      const double* up, const double* u, const double* um, const double* mu,
                      const double* la, const double* rho, const SgCoef& sg, double beta,
                      double* upo, double* acc, double h, double dt, hipStream_t s,
                      const std::function<int()>& after_interior, const std::function<int()>& after_damping,
                      int variant, int kchunk, int zfill, const double* accf, const DampTile* dtiles,
                      int ndtiles)
{
   const double cof = 1.0 / (h * h), dt2i = 1.0 / (dt * dt), dt4i12 = dt * dt * dt * dt / 12;
   const int k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   const int k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
   dim3 blk(64, 4, 1);
   if (k2 >= k1 && variant >= 1) {
      int rc = launch_sweep_lds(true, variant, b, k1, k2, kchunk, up, u, um, mu, la, rho, sg, upo, h, dt, beta, s,
                                const_cast<double*>(accf));
      if (rc) return rc;
      if (after_interior && (rc = after_interior())) return rc;
      if (beta != 0 && variant >= 6) {
         if ((rc = launch_damp_lds(b, k1, k2, kchunk, upo, u, um, rho, sg, beta, s, variant >= 9 ? dtiles : nullptr,
                                   ndtiles)))
            return rc;
      } else if (beta != 0 && variant < 4) {  // variants 4, 5 damp inside the sweep
         dim3 grd(cdiv(nii, 64), cdiv(njj, 4), k2 - k1 + 1);
         k_addsgd4_planes<<<grd, blk, 0, s>>>(b, k1, k2, upo, u, um, rho, sg, beta);
      }
   } else {
      if (k2 >= k1) {
         dim3 grd(cdiv(nii, 64), cdiv(njj, 4), k2 - k1 + 1);
         k_corr_interior<<<grd, blk, 0, s>>>(b, k1, k2, up, u, um, mu, la, rho, sg, upo, cof, dt2i, dt4i12, beta);
      }
      int rc;
      if (after_interior && (rc = after_interior())) return rc;
   }
   if (after_damping) {
      const int rc = after_damping();
      if (rc) return rc;
   }
   // kernel variant 9: the rows' damping in a pass of its own after them
   // (inside the rows kernel its registers cost that kernel's occupancy)
   if (variant >= 1)
      return launch_corr_tail(b, nk, onesided, d_sbp, up, u, um, mu, la, rho, sg, beta, upo, h, dt, s, zfill, accf,
                              variant >= 9);
   for (int side = 4; side <= 5; side++) {
      if (onesided[side] != 1) continue;
      const int ks = side == 4 ? 0 : nk - SLAB_PLANES + 2;  // planes KK(0)..KK(8) of the rows
      k_accel_slab<<<cdiv(accel_slab_size(b), 256), 256, 0, s>>>(b, ks, up, u, um, dt2i, acc);
      dim3 grd(cdiv(nii, 64), cdiv(njj, 4), 6);
      k_corr_sbp<<<grd, blk, 0, s>>>(b, side == 4, nk, d_sbp, up, acc, ks, u, um, mu, la, rho, sg, upo, cof,
                                     dt4i12, beta);
   }
   const long ns = shell_count(b);
   k_shell<<<cdiv(ns, 256), 256, 0, s>>>(b, 1, up, nullptr, nullptr, upo, zfill);
   return check_hip(hipGetLastError(), "fused corrector launch");
}

int launch_src_apply(const Box& b, const SourceTable& st, double t, int deriv2, double coef,
                     const double* rho, double* up, hipStream_t s, const double* tdev)
{
   if (st.nunique <= 0) return OK;
   k_src_apply<<<cdiv(st.nunique, 128), 128, 0, s>>>(b, st, t, deriv2, coef, rho, up, tdev);
   return check_hip(hipGetLastError(), "source update launch");
}

int launch_maxeig(long n, const double* mu, const double* la, const double* rho, double* partial,
                  int npartial, double* out, hipStream_t s)
{
   k_maxeig<<<npartial, RED_BLOCK, 0, s>>>(n, mu, la, rho, partial);
   k_final_max<<<1, RED_BLOCK, 0, s>>>(npartial, partial, out);
   return check_hip(hipGetLastError(), "maxeig launch");
}

int launch_exact(const Box& b, const ExactParams& P, double* out, hipStream_t s)
{
   k_exact<<<cdiv(b.npts, 256), 256, 0, s>>>(b, P, out);
   return check_hip(hipGetLastError(), "exact solution launch");
}

int launch_solerr(const Box& b, const NormWindow& w, const double* uex, const double* u,
                  double* partial, int npartial, double* out, hipStream_t s)
{
   k_solerr<<<npartial, RED_BLOCK, 0, s>>>(b, w, uex, u, partial);
   k_final_norm<<<1, RED_BLOCK, 0, s>>>(npartial, partial, out);
   return check_hip(hipGetLastError(), "error norm launch");
}

}  // namespace sw4amd
