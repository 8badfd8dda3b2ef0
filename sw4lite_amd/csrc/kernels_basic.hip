// kernels_basic.hip -- one-thread-per-point gfx950 kernels for every stage of
// the sw4lite time step, behind the reference's kernel-level entry points:
//   rhs4sg / rhs4sg_rev       rhs4sg.C:38-849, rhs4sg_rev.C:43-864
//   predfort/corrfort/dpdmt   ew-cfromfort.C:40-141
//   addsgd4fort(_indrev)      ew-cfromfort.C:748-849, :952-1056
//   bcfortsg(_indrev)         ew-cfromfort.C:205-747
//   Force (point sources)     EW.C:3085-3123, GridPointSource.C:330-586
// Both Sarray layouts are supported through the CO template parameter.  The
// fused, LDS-staged kernels of the fast step live in kernels_fused.hip and
// are checked against these.
#include "kernels_common.h"

namespace sw4amd {

// ---------------------------------------------------------------------------
// Divergence of stress
// ---------------------------------------------------------------------------
template <int CO>
__global__ void __launch_bounds__(256)
k_rhs4_interior(Box b, int klo, int khi, const double* __restrict__ u,
                const double* __restrict__ mu, const double* __restrict__ la,
                const double* __restrict__ strx, const double* __restrict__ stry,
                const double* __restrict__ strz, double* __restrict__ lu, double cof)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = klo + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > khi) return;
   double r[3];
   rhs_interior_point<CO>(b, i, j, k, Plain{u}, mu, la, strx, stry, strz, r);
   for (int c = 0; c < 3; c++) lu[vidx<CO>(b, c, i, j, k)] = cof * r[c];
}

template <int CO>
__global__ void __launch_bounds__(256)
k_rhs4_sbp(Box b, int top, int nk, const SbpTables* __restrict__ sbp, const double* __restrict__ u,
           const double* __restrict__ mu, const double* __restrict__ la,
           const double* __restrict__ strx, const double* __restrict__ stry,
           double* __restrict__ lu, double cof)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int kr = 1 + blockIdx.z;  // boundary row 1..6
   if (i > b.ie - 2 || j > b.je - 2) return;
   const int k = top ? kr : nk - kr + 1;
   double r[3];
   rhs_sbp_point<CO>(b, i, j, k, kr, top ? 1 : nk, top ? 1 : -1, sbp, Plain{u}, mu, la, strx, stry, r);
   for (int c = 0; c < 3; c++) lu[vidx<CO>(b, c, i, j, k)] = cof * r[c];
}

// ---------------------------------------------------------------------------
// Pointwise time update
// ---------------------------------------------------------------------------
template <int CO>
__global__ void k_pred(long npts, double* __restrict__ up, const double* __restrict__ u,
                       const double* __restrict__ um, const double* __restrict__ lu,
                       const double* __restrict__ fo, const double* __restrict__ rho, double dt2)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= npts) return;
   const double s = dt2 / rho[p];
   for (int c = 0; c < 3; c++) {
      const long q = CO ? p + npts * c : 3 * p + c;
      up[q] = 2 * u[q] - um[q] + s * (lu[q] + fo[q]);
   }
}

template <int CO>
__global__ void k_corr(long npts, double* __restrict__ up, const double* __restrict__ lu,
                       const double* __restrict__ fo, const double* __restrict__ rho, double dt4i12)
{
   const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (p >= npts) return;
   const double s = dt4i12 / rho[p];
   for (int c = 0; c < 3; c++) {
      const long q = CO ? p + npts * c : 3 * p + c;
      up[q] += s * (lu[q] + fo[q]);
   }
}

__global__ void k_dpdmt(long n, const double* __restrict__ up, const double* __restrict__ u,
                        const double* __restrict__ um, double* __restrict__ uacc, double dt2i)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= n) return;
   uacc[q] = dt2i * (up[q] - 2 * u[q] + um[q]);
}

// ---------------------------------------------------------------------------
// Supergrid damping
// ---------------------------------------------------------------------------
template <int CO>
__global__ void __launch_bounds__(256)
k_addsgd4(Box b, double* __restrict__ up, const double* __restrict__ u,
          const double* __restrict__ um, const double* __restrict__ rho, SgCoef sg, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   const int k = b.kb + 2 + blockIdx.z;
   if (i > b.ie - 2 || j > b.je - 2 || k > b.ke - 2) return;
   // every term carries one of these damping coefficients: where all nine
   // vanish (away from the supergrid layers) the update subtracts zero for
   // finite data, and the point is left as it is
   const int ii = i - b.ib, jj = j - b.jb, kk = k - b.kb;
   if (sg.dcx[ii - 1] == 0 && sg.dcx[ii] == 0 && sg.dcx[ii + 1] == 0 && sg.dcy[jj - 1] == 0 && sg.dcy[jj] == 0 &&
       sg.dcy[jj + 1] == 0 && sg.dcz[kk - 1] == 0 && sg.dcz[kk] == 0 && sg.dcz[kk + 1] == 0)
      return;
   double d[3];
   addsgd4_point<CO>(b, i, j, k, u, um, rho, sg, beta, d);
   for (int c = 0; c < 3; c++) up[vidx<CO>(b, c, i, j, k)] -= d[c];
}

// ---------------------------------------------------------------------------
// Ghost-point boundary conditions
// ---------------------------------------------------------------------------
struct Window {
   int i0, i1, j0, j1, k0, k1;
};

template <int CO>
__global__ void k_bc_fill(Box b, Window w, double* __restrict__ u, const double* __restrict__ bf)
{
   const long wi = w.i1 - w.i0 + 1, wj = w.j1 - w.j0 + 1, wk = w.k1 - w.k0 + 1;
   const long n = wi * wj * wk;
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= n) return;
   const int i = w.i0 + (int)(q % wi);
   const int j = w.j0 + (int)((q / wi) % wj);
   const int k = w.k0 + (int)(q / (wi * wj));
   for (int c = 0; c < 3; c++) u[vidx<CO>(b, c, i, j, k)] = bf ? bf[3 * q + c] : 0.0;
}

template <int CO>
__global__ void k_bc_periodic(Box b, Window w, int pi, int pj, int pk, double* __restrict__ u)
{
   const long wi = w.i1 - w.i0 + 1, wj = w.j1 - w.j0 + 1, wk = w.k1 - w.k0 + 1;
   const long n = wi * wj * wk;
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= n) return;
   const int i = w.i0 + (int)(q % wi);
   const int j = w.j0 + (int)((q / wi) % wj);
   const int k = w.k0 + (int)(q / (wi * wj));
   for (int c = 0; c < 3; c++) u[vidx<CO>(b, c, i, j, k)] = u[vidx<CO>(b, c, i + pi, j + pj, k + pk)];
}

template <int CO>
__global__ void k_bc_freesurf(Box b, int k, int kl, double h, const SbpTables* __restrict__ sbp,
                              double* __restrict__ u, const double* __restrict__ mu,
                              const double* __restrict__ la, const double* __restrict__ bf,
                              const double* __restrict__ strx, const double* __restrict__ stry)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const long q = (i - b.ib) + b.ni * (j - b.jb);
   double t[3] = {0, 0, 0};
   if (bf) {
      t[0] = bf[3 * q];
      t[1] = bf[3 * q + 1];
      t[2] = bf[3 * q + 2];
   }
   double g[3];
   freesurf_point<CO>(b, i, j, k, kl, h, sbp->sbop, u, mu, la, t, strx, stry, g);
   for (int c = 0; c < 3; c++) u[vidx<CO>(b, c, i, j, k - kl)] = g[c];
}

// ---------------------------------------------------------------------------
// Point-source forcing: F(:, unique point) = sum over its sources.
// ---------------------------------------------------------------------------
__global__ void k_zero(long n, double* __restrict__ a)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q < n) a[q] = 0.0;
}

template <int CO>
__global__ void k_force(Box b, SourceTable st, double t, int deriv2, double* __restrict__ F)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= st.nunique) return;
   const int s0 = st.first[r], s1 = st.first[r + 1];
   const int i = st.ijk[3 * s0], j = st.ijk[3 * s0 + 1], k = st.ijk[3 * s0 + 2];
   double f0 = F[vidx<CO>(b, 0, i, j, k)], f1 = F[vidx<CO>(b, 1, i, j, k)],
          f2 = F[vidx<CO>(b, 2, i, j, k)];
   for (int s = s0; s < s1; s++) {
      const double g = time_function(st.tfkind[s], deriv2, st.freq[s], t - st.t0[s]);
      f0 += st.force[3 * s] * g;
      f1 += st.force[3 * s + 1] * g;
      f2 += st.force[3 * s + 2] * g;
   }
   F[vidx<CO>(b, 0, i, j, k)] = f0;
   F[vidx<CO>(b, 1, i, j, k)] = f1;
   F[vidx<CO>(b, 2, i, j, k)] = f2;
}

// ===========================================================================
// Host launchers (all stream-ordered, no synchronisation)
// ===========================================================================
namespace {
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}

int launch_rhs4sg(int co, const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                  double* lu, const double* u, const double* mu, const double* la, double h,
                  const double* strx, const double* stry, const double* strz, hipStream_t s)
{
   const double cof = 1.0 / (h * h);
   const int k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   const int k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
   if (nii <= 0 || njj <= 0) return OK;
   dim3 blk(64, 4, 1);
   if (k2 >= k1) {
      dim3 grd(cdiv(nii, 64), cdiv(njj, 4), k2 - k1 + 1);
      if (co)
         k_rhs4_interior<1><<<grd, blk, 0, s>>>(b, k1, k2, u, mu, la, strx, stry, strz, lu, cof);
      else
         k_rhs4_interior<0><<<grd, blk, 0, s>>>(b, k1, k2, u, mu, la, strx, stry, strz, lu, cof);
   }
   for (int side = 4; side <= 5; side++) {
      if (onesided[side] != 1) continue;
      dim3 grd(cdiv(nii, 64), cdiv(njj, 4), 6);
      if (co)
         k_rhs4_sbp<1><<<grd, blk, 0, s>>>(b, side == 4, nk, d_sbp, u, mu, la, strx, stry, lu, cof);
      else
         k_rhs4_sbp<0><<<grd, blk, 0, s>>>(b, side == 4, nk, d_sbp, u, mu, la, strx, stry, lu, cof);
   }
   return check_hip(hipGetLastError(), "rhs4sg launch");
}

int launch_pred(int co, const Box& b, double* up, const double* u, const double* um,
                const double* lu, const double* fo, const double* rho, double dt2, hipStream_t s)
{
   if (co)
      k_pred<1><<<cdiv(b.npts, 256), 256, 0, s>>>(b.npts, up, u, um, lu, fo, rho, dt2);
   else
      k_pred<0><<<cdiv(b.npts, 256), 256, 0, s>>>(b.npts, up, u, um, lu, fo, rho, dt2);
   return check_hip(hipGetLastError(), "predfort launch");
}

int launch_corr(int co, const Box& b, double* up, const double* lu, const double* fo,
                const double* rho, double dt4, hipStream_t s)
{
   const double c = dt4 / 12;
   if (co)
      k_corr<1><<<cdiv(b.npts, 256), 256, 0, s>>>(b.npts, up, lu, fo, rho, c);
   else
      k_corr<0><<<cdiv(b.npts, 256), 256, 0, s>>>(b.npts, up, lu, fo, rho, c);
   return check_hip(hipGetLastError(), "corrfort launch");
}

int launch_dpdmt(const Box& b, const double* up, const double* u, const double* um, double* uacc,
                 double dt2i, hipStream_t s)
{
   const long n = 3 * b.npts;
   k_dpdmt<<<cdiv(n, 256), 256, 0, s>>>(n, up, u, um, uacc, dt2i);
   return check_hip(hipGetLastError(), "dpdmtfort launch");
}

int launch_addsgd4(int co, const Box& b, double* up, const double* u, const double* um,
                   const double* rho, const SgCoef& sg, double beta, hipStream_t s)
{
   if (beta == 0) return OK;
   const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3, nkk = b.ke - b.kb - 3;
   if (nii <= 0 || njj <= 0 || nkk <= 0) return OK;
   dim3 blk(64, 4, 1), grd(cdiv(nii, 64), cdiv(njj, 4), nkk);
   if (co)
      k_addsgd4<1><<<grd, blk, 0, s>>>(b, up, u, um, rho, sg, beta);
   else
      k_addsgd4<0><<<grd, blk, 0, s>>>(b, up, u, um, rho, sg, beta);
   return check_hip(hipGetLastError(), "addsgd4 launch");
}

int launch_bcfortsg(int co, const Box& b, const int* wind, int nx, int ny, int nz, double* u,
                    double h, const int* bccnd, const SbpTables* d_sbp, const double* mu,
                    const double* la, const double* const* bforce, const double* strx,
                    const double* stry, hipStream_t s)
{
   for (int side = 0; side < 6; side++) {
      const int* w = wind + 6 * side;
      Window win{w[0], w[1], w[2], w[3], w[4], w[5]};
      const long n = (long)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1);
      if (bccnd[side] == 1 || bccnd[side] == 2) {
         if (n <= 0) continue;
         if (co)
            k_bc_fill<1><<<cdiv(n, 256), 256, 0, s>>>(b, win, u, bforce[side]);
         else
            k_bc_fill<0><<<cdiv(n, 256), 256, 0, s>>>(b, win, u, bforce[side]);
      } else if (bccnd[side] == 3) {
         if (n <= 0) continue;
         const int pi = side == 0 ? nx : side == 1 ? -nx : 0;
         const int pj = side == 2 ? ny : side == 3 ? -ny : 0;
         const int pk = side == 4 ? nz : side == 5 ? -nz : 0;
         if (co)
            k_bc_periodic<1><<<cdiv(n, 256), 256, 0, s>>>(b, win, pi, pj, pk, u);
         else
            k_bc_periodic<0><<<cdiv(n, 256), 256, 0, s>>>(b, win, pi, pj, pk, u);
      } else if (bccnd[side] == 0) {
         if (side != 4 && side != 5)
            return set_error(E_INVAL, "bcfortsg: free surface is only implemented on sides 4 and 5");
         const int k = side == 4 ? 1 : nz, kl = side == 4 ? 1 : -1;
         const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
         dim3 blk(64, 4, 1), grd(cdiv(nii, 64), cdiv(njj, 4), 1);
         if (co)
            k_bc_freesurf<1><<<grd, blk, 0, s>>>(b, k, kl, h, d_sbp, u, mu, la, bforce[side], strx, stry);
         else
            k_bc_freesurf<0><<<grd, blk, 0, s>>>(b, k, kl, h, d_sbp, u, mu, la, bforce[side], strx, stry);
      }
   }
   return check_hip(hipGetLastError(), "bcfortsg launch");
}

int launch_force(int co, const Box& b, double* F, const SourceTable& st, double t, int deriv2,
                 bool zero_first, hipStream_t s)
{
   if (zero_first) k_zero<<<cdiv(3 * b.npts, 256), 256, 0, s>>>(3 * b.npts, F);
   if (st.nunique > 0) {
      if (co)
         k_force<1><<<cdiv(st.nunique, 128), 128, 0, s>>>(b, st, t, deriv2, F);
      else
         k_force<0><<<cdiv(st.nunique, 128), 128, 0, s>>>(b, st, t, deriv2, F);
   }
   return check_hip(hipGetLastError(), "force launch");
}

int launch_zero(double* a, long n, hipStream_t s)
{
   if (n > 0) k_zero<<<cdiv(n, 256), 256, 0, s>>>(n, a);
   return check_hip(hipGetLastError(), "zero launch");
}

__global__ void k_fill(long n, double* __restrict__ a, double v)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q < n) a[q] = v;
}

int launch_fill(double* a, long n, double v, hipStream_t s)
{
   if (n > 0) k_fill<<<cdiv(n, 256), 256, 0, s>>>(n, a, v);
   return check_hip(hipGetLastError(), "fill launch");
}

}  // namespace sw4amd
