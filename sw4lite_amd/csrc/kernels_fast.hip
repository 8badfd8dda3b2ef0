// kernels_fast.hip -- LDS-staged sweeps of the fused time step.
//
// A workgroup owns a 32 x 8 tile of (i,j) columns and marches along k over a
// chunk of planes.  The five displacement planes k-2..k+2 the 4th-order
// stencil touches are kept, with their 2-point (i,j) halo, in an LDS ring;
// each new plane is loaded once from HBM (coalesced rows of 36 doubles) and
// read ~135 times from LDS.  mu and lambda of the current plane sit in LDS for
// the in-plane terms, the column values k+-1, k+-2 rotate in registers.  The
// corrector stages (Up - 2U + Um)/dt^2 instead of U, so the acceleration is
// never stored.  The arithmetic is rhs4_interior_body, shared with the
// global-memory kernels (same results to rounding).  The supergrid damping
// of the corrector sweep runs as k_addsgd4_planes right after it.
#include <algorithm>

#include "kernels_common.h"
#include "kernels_fast.h"
#include "sbp_row_body.h"
#include "stencil_body.h"

namespace sw4amd {

namespace {
constexpr int TX = 32, TY = 8, HX = TX + 4, HY = TY + 4, NT = TX * TY, PL = HX * HY;
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}  // namespace


struct LdsSten {
   const double* p0;  // ring slot of plane k-2, offset to this thread's point
   const double* p1;
   const double* p2;
   const double* p3;
   const double* p4;
   const double* pm;  // mu of plane k at this point
   const double* pl;
   const double* psx;  // strx at this thread's i (LDS copy over the tile)
   const double* psy;
   double muc[5], lac[5];
   double szv[5];
   __device__ __forceinline__ double u(int c, int di, int dj, int dk) const
   {
      const double* p = dk == -2 ? p0 : dk == -1 ? p1 : dk == 0 ? p2 : dk == 1 ? p3 : p4;
      return p[c * PL + dj * HX + di];
   }
   __device__ __forceinline__ double mu(int di, int dj, int dk) const { return dk ? muc[dk + 2] : pm[dj * HX + di]; }
   __device__ __forceinline__ double la(int di, int dj, int dk) const { return dk ? lac[dk + 2] : pl[dj * HX + di]; }
   __device__ __forceinline__ double sx(int d) const { return psx[d]; }
   __device__ __forceinline__ double sy(int d) const { return psy[d]; }
   __device__ __forceinline__ double sz(int d) const { return szv[d + 2]; }
};

// CORR = false: out = 2u - um + dt2/rho cof L(u)                  (predictor)
// CORR = true:  out = up + dt4/(12 rho) cof L((up - 2u + um) dt2i)  (corrector)
// DAMP (corrector only): the supergrid damping of the corrector is applied in
// the same sweep.  d = U - Um is kept for the planes k..k+2 in an LDS ring
// (in-plane stencil at plane k), the thread's own column of d for k-2..k+2
// and rho at k-1, k+1 in registers, rho of plane k in LDS.
// PIPE: software pipeline.  Every global load an iteration needs (plane k+3
// of the staged arrays, the material of plane k+1, the column and centre
// values of the next plane) is issued right after the barrier and consumed in
// the next iteration.  The compute phase itself reads only LDS and registers:
// the load counter is in order, so a global load issued after the prefetch
// and waited on in the same iteration would wait for the prefetch as well.
// TR (predictor only): out = cof L(u)/S with S = strx stry strz at the point
// (no um, no rho): the transposed operator's L^T z = S^{-1} L(S z).
template <bool CORR, bool PIPE, bool DAMP, bool TR = false>
__device__ __forceinline__ void sweep_lds_body(unsigned bid, Box b, TileMap map, int kbeg, int kend, int kc,
                                               const double* __restrict__ upin, const double* __restrict__ u,
                                               const double* __restrict__ um, const double* __restrict__ mu,
                                               const double* __restrict__ la, const double* __restrict__ rho,
                                               SgCoef sg, double* __restrict__ out, double cof, double dt2i,
                                               double cupd, double beta)
{
   static_assert(!DAMP || CORR, "damping belongs to the corrector");
   constexpr int NU = (3 * PL + NT - 1) / NT;  // staged displacement values per thread
   constexpr int NM = (PL + NT - 1) / NT;      // staged material values per thread
   __shared__ double su[5 * 3 * PL];            // ring: plane kp in slot (kp - k0 + 2) % 5
   __shared__ double sm[PL], sl[PL];
   __shared__ double sd[DAMP ? 3 * 3 * PL : 1];  // d = U - Um, plane kp in slot (kp - k0) % 3
   __shared__ double sr[DAMP ? PL : 1];          // rho of plane k
   __shared__ double ssx[HX], ssy[HY];            // strx, stry over the tile with halo
   __shared__ double sgx[DAMP ? 5 * TX : 1], sgy[DAMP ? 5 * TY : 1];  // dc(-1,0,+1), str, co per column
   __shared__ double sct[3 * NT];                 // centre values of plane k (see load_centre)
   // XCD-aware tile order: workgroups go to the 8 XCDs round-robin by their
   // id, so id -> (xcd = id % 8, slot = id / 8) and each XCD takes one
   // contiguous run of tiles (i fastest, then j, then k-chunk).  Tiles that
   // share halos then run on the same XCD at the same time and meet in its L2.
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(bid % 8) * per + (int)(bid / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;  // halo corner of the tile
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const int k0 = kbeg + tkz * kc;
   const int k1 = min(kend, k0 + kc - 1);
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;

   // element e = tid + p*NT of the staged (c, jj, ii) tile, its global offset
   long goff[NU];
   bool gok[NU];
#pragma unroll
   for (int p = 0; p < NU; p++) {
      const int e = tid + p * NT;
      const int c = e / PL, r = e - c * PL, jj = r / HX, ii = r - jj * HX;
      const int gi = ia + ii, gj = ja + jj;
      gok[p] = e < 3 * PL && gi <= b.ie && gj <= b.je;
      goff[p] = gok[p] ? (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
   }
   long moff[NM];
   bool mok[NM];
#pragma unroll
   for (int p = 0; p < NM; p++) {
      const int e = tid + p * NT;
      const int jj = e / HX, ii = e - jj * HX;
      const int gi = ia + ii, gj = ja + jj;
      mok[p] = e < PL && gi <= b.ie && gj <= b.je;
      moff[p] = mok[p] ? (gi - b.ib) + ni * (gj - b.jb) : 0;
   }
   double ru[NU], rp[NU], rq[NU], rmu[NM], rla[NM], rrh[DAMP ? NM : 1];
   // Out-of-tile elements have offset 0 (a valid address): every load is
   // issued unconditionally and masked afterwards, so no branches.
   auto load_plane = [&](int kp) {
      const long pb = nij * (kp - b.kb);
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const double vu = u[pb + goff[p]];
         ru[p] = gok[p] ? vu : 0.0;
         if (CORR) {
            const double vp = upin[pb + goff[p]], vq = um[pb + goff[p]];
            rp[p] = gok[p] ? vp : 0.0;
            rq[p] = gok[p] ? vq : 0.0;
         }
      }
   };
   auto store_plane = [&](int slot, int dslot) {
      double* dst = su + slot * 3 * PL;
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const int e = tid + p * NT;
         if (e < 3 * PL) {
            dst[e] = CORR ? dt2i * (rp[p] - 2 * ru[p] + rq[p]) : ru[p];
            if (DAMP) sd[dslot * 3 * PL + e] = ru[p] - rq[p];
         }
      }
   };
   auto load_mat = [&](int kp) {
      const long pb = nij * (kp - b.kb);
#pragma unroll
      for (int p = 0; p < NM; p++) {
         const double vm = mu[pb + moff[p]], vl = la[pb + moff[p]];
         rmu[p] = mok[p] ? vm : 0.0;
         rla[p] = mok[p] ? vl : 0.0;
         if (DAMP) {
            const double vr = rho[pb + moff[p]];
            rrh[p] = mok[p] ? vr : 0.0;
         }
      }
   };
   auto store_mat = [&]() {
#pragma unroll
      for (int p = 0; p < NM; p++) {
         const int e = tid + p * NT;
         if (e < PL) {
            sm[e] = rmu[p];
            sl[e] = rla[p];
            if (DAMP) sr[e] = rrh[p];
         }
      }
   };
   // values of plane kp at this thread's point used after the stencil: the
   // centre of up (corrector) or um (predictor), rho (without DAMP)
   double ctr[3] = {0, 0, 0}, crho = 1.0;
   auto load_centre = [&](int kp, double c3[3], double& r) {
      const long p = col + nij * (kp - b.kb);  // col = 0 where inactive: a valid address
      if (!TR) {
#pragma unroll
         for (int c = 0; c < 3; c++) c3[c] = (CORR ? upin : um)[p + npts * c];
      }
      if (!DAMP && !TR) {
         const double v = rho[p];
         r = active ? v : 1.0;
      }
   };

   LdsSten s;
   // damping: own column of d at k-2..k+2, rho at k-1 and k+1, and the x/y
   // supergrid coefficients of this column (constant over the sweep)
   double dcol[5][3], rkm = 0, rkp = 0, rk0 = 0;
   if (DAMP) {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const long g = col + nij * (k0 - 2 + q - b.kb) + npts * c;
            dcol[q][c] = active ? u[g] - um[g] : 0.0;
         }
      rkm = active ? rho[col + nij * (k0 - 1 - b.kb)] : 0.0;
      rkp = active ? rho[col + nij * (k0 + 1 - b.kb)] : 0.0;
      // x/y supergrid coefficients of the tile's columns (interior columns only)
      if (tid < TX) {
         const int ii = ia + 2 + tid - b.ib;
         const bool ok = ia + 2 + tid <= b.ie - 2;
         sgx[0 * TX + tid] = ok ? sg.dcx[ii - 1] : 0.0;
         sgx[1 * TX + tid] = ok ? sg.dcx[ii] : 0.0;
         sgx[2 * TX + tid] = ok ? sg.dcx[ii + 1] : 0.0;
         sgx[3 * TX + tid] = ok ? sg.strx[ii] : 0.0;
         sgx[4 * TX + tid] = ok ? sg.cox[ii] : 0.0;
      } else if (tid >= 64 && tid < 64 + TY) {
         const int q = tid - 64, jj = ja + 2 + q - b.jb;
         const bool ok = ja + 2 + q <= b.je - 2;
         sgy[0 * TY + q] = ok ? sg.dcy[jj - 1] : 0.0;
         sgy[1 * TY + q] = ok ? sg.dcy[jj] : 0.0;
         sgy[2 * TY + q] = ok ? sg.dcy[jj + 1] : 0.0;
         sgy[3 * TY + q] = ok ? sg.stry[jj] : 0.0;
         sgy[4 * TY + q] = ok ? sg.coy[jj] : 0.0;
      }
   }
   // stretching over the tile with its halo (1 beyond the array ends)
   if (tid >= 128 && tid < 128 + HX) ssx[tid - 128] = ia + tid - 128 <= b.ie ? sg.strx[ia + tid - 128 - b.ib] : 1.0;
   if (tid >= 192 && tid < 192 + HY) ssy[tid - 192] = ja + tid - 192 <= b.je ? sg.stry[ja + tid - 192 - b.jb] : 1.0;
#pragma unroll
   for (int q = 0; q < 4; q++) {
      load_plane(k0 - 2 + q);
      store_plane(q, q >= 2 ? q - 2 : 2);  // d ring: planes k0 (slot 0), k0+1 (slot 1)
      const long pq = nij * (k0 - 2 + q - b.kb);
      s.muc[q] = active ? mu[col + pq] : 0.0;
      s.lac[q] = active ? la[col + pq] : 0.0;
      s.szv[q] = sg.strz[k0 - 2 + q - b.kb];
   }
   s.pm = sm + o;
   s.pl = sl + o;
   s.psx = ssx + tx + 2;
   s.psy = ssy + ty + 2;
   // PIPE: prefetch plane k0+2, material of plane k0, column values of k0+2,
   // centre values of plane k0
   double cmu = 0, cla = 0, csz = 0;
   double nctr[3] = {0, 0, 0}, ncrho = 1.0;
   if (PIPE) {
      load_plane(k0 + 2);
      load_mat(k0);
      cmu = active ? mu[col + nij * (k0 + 2 - b.kb)] : 0.0;
      cla = active ? la[col + nij * (k0 + 2 - b.kb)] : 0.0;
      csz = sg.strz[k0 + 2 - b.kb];
      load_centre(k0, nctr, ncrho);
   }

   for (int k = k0; k <= k1; k++) {
      const int sl0 = (k - k0) % 5;  // slot of plane k-2
      if (!PIPE) {
         load_plane(k + 2);
         load_mat(k);
         cmu = active ? mu[col + nij * (k + 2 - b.kb)] : 0.0;
         cla = active ? la[col + nij * (k + 2 - b.kb)] : 0.0;
         csz = sg.strz[k + 2 - b.kb];
         load_centre(k, ctr, crho);
      } else {
#pragma unroll
         for (int c = 0; c < 3; c++) sct[c * NT + tid] = nctr[c];
         crho = ncrho;
      }
      if (!PIPE)
#pragma unroll
         for (int c = 0; c < 3; c++) sct[c * NT + tid] = ctr[c];
      store_plane((sl0 + 4) % 5, (k + 2 - k0) % 3);
      store_mat();
      s.muc[4] = cmu;
      s.lac[4] = cla;
      s.szv[4] = csz;
      // z supergrid coefficients of plane k (uniform over the workgroup)
      const int kz = k - b.kb;
      double gz[5] = {0, 0, 0, 0, 0};
      if (DAMP) {
         gz[0] = sg.dcz[kz - 1]; gz[1] = sg.dcz[kz]; gz[2] = sg.dcz[kz + 1]; gz[3] = sg.strz[kz]; gz[4] = sg.coz[kz];
      }
      __syncthreads();
      if (PIPE && k < k1) {
         load_plane(k + 3);
         load_mat(k + 1);
         const long pq = nij * (k + 3 - b.kb);
         cmu = mu[col + pq];  // col = 0 where inactive: a valid address, value unused
         cla = la[col + pq];
         csz = sg.strz[k + 3 - b.kb];
         load_centre(k + 1, nctr, ncrho);
      }
      if (active) {
         s.p0 = su + (sl0 * 3) * PL + o;
         s.p1 = su + (((sl0 + 1) % 5) * 3) * PL + o;
         s.p2 = su + (((sl0 + 2) % 5) * 3) * PL + o;
         s.p3 = su + (((sl0 + 3) % 5) * 3) * PL + o;
         s.p4 = su + (((sl0 + 4) % 5) * 3) * PL + o;
         double r[3];
         rhs4_interior_body(s, r);
         const long p = col + nij * (k - b.kb);
         const double r0 = DAMP ? sr[o] : crho;
         if (DAMP) rk0 = r0;
         const double sc = TR ? 1.0 / (s.psx[0] * s.psy[0] * s.szv[2]) : cupd / r0;
         double dmp[3] = {0, 0, 0};
         if (DAMP) {
            const double* dpl = sd + ((k - k0) % 3) * 3 * PL + o;  // d of plane k at this point
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[4][c] = sd[((k + 2 - k0) % 3) * 3 * PL + c * PL + o];
            const double birho = beta / r0;
            const double* gx = sgx + tx;
            const double* gy = sgy + ty;
            const double rxp = sr[o + 1] * gx[2 * TX], rx0 = r0 * gx[TX], rxm = sr[o - 1] * gx[0];
            const double ryp = sr[o + HX] * gy[2 * TY], ry0 = r0 * gy[TY], rym = sr[o - HX] * gy[0];
            const double rzp = rkp * gz[2], rz0 = r0 * gz[1], rzm = rkm * gz[0];
            const double wx = gx[3 * TX] * gy[4 * TY] * gz[4];
            const double wy = gy[3 * TY] * gx[4 * TX] * gz[4];
            const double wz = gz[3] * gx[4 * TX] * gy[4 * TY];
#pragma unroll
            for (int c = 0; c < 3; c++) {
               const double* q = dpl + c * PL;
               const double ax = rxp * (q[2] - 2 * q[1] + q[0]) - 2 * rx0 * (q[1] - 2 * q[0] + q[-1]) +
                                 rxm * (q[0] - 2 * q[-1] + q[-2]);
               const double ay = ryp * (q[2 * HX] - 2 * q[HX] + q[0]) - 2 * ry0 * (q[HX] - 2 * q[0] + q[-HX]) +
                                 rym * (q[0] - 2 * q[-HX] + q[-2 * HX]);
               const double az = rzp * (dcol[4][c] - 2 * dcol[3][c] + dcol[2][c]) -
                                 2 * rz0 * (dcol[3][c] - 2 * dcol[2][c] + dcol[1][c]) +
                                 rzm * (dcol[2][c] - 2 * dcol[1][c] + dcol[0][c]);
               dmp[c] = birho * (wx * ax + wy * ay + wz * az);
            }
         }
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const long q = p + npts * c;
            const double cv = sct[c * NT + tid];
            if (TR)
               out[q] = sc * (cof * r[c]);
            else if (!CORR)
               out[q] = 2 * s.p2[c * PL] - cv + sc * (cof * r[c]);
            else if (DAMP)
               out[q] = (cv + sc * (cof * r[c])) - dmp[c];
            else
               out[q] = cv + sc * (cof * r[c]);
         }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++) {
         s.muc[q] = s.muc[q + 1];
         s.lac[q] = s.lac[q + 1];
         s.szv[q] = s.szv[q + 1];
      }
      if (DAMP) {
#pragma unroll
         for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[q][c] = dcol[q + 1][c];
         rkm = rk0;  // rho of plane k becomes k-1 of the next row (0 where inactive)
         rkp = active && k < k1 ? rho[col + nij * (k + 2 - b.kb)] : 0.0;
      }
   }
}

template <bool CORR, bool PIPE, int MINW, bool DAMP = false, bool TR = false>
__global__ void __launch_bounds__(NT, MINW)
k_sweep_lds(Box b, TileMap map, int kbeg, int kend, int kc, const double* __restrict__ upin,
            const double* __restrict__ u, const double* __restrict__ um, const double* __restrict__ mu,
            const double* __restrict__ la, const double* __restrict__ rho, SgCoef sg, double* __restrict__ out,
            double cof, double dt2i, double cupd, double beta)
{
   sweep_lds_body<CORR, PIPE, DAMP, TR>(blockIdx.x, b, map, kbeg, kend, kc, upin, u, um, mu, la, rho, sg, out, cof, dt2i,
                                    cupd, beta);
}

// Modes of the asynchronous-load sweep: the predictor (out = Up from u = U,
// um = Um; with acc, also the acceleration (Up - 2U + Um)/dt^2 at the points
// it computes, the expression the corrector stages), the transposed operator
// (TR above: u = W, no um, no rho), and the corrector from the stored
// acceleration (u = A, um = Up for the centre values, cupd = dt^4/12).
enum DmaMode { DMA_PRED = 0, DMA_TR = 1, DMA_CORR = 2 };

// The acceleration the corrector applies its stencil to (dpdmtfort,
// ew-cfromfort.C:116-141), one expression wherever it is formed so that a
// stored value equals the one staged on the fly.
__device__ __forceinline__ double accel_of(double up, double u, double um, double dt2i)
{
   return dt2i * (up - 2 * u + um);
}

// Sweep (mode MODE above) whose staged planes reach
// LDS by asynchronous global-to-LDS loads (global_load_lds_dwordx4: lane l of
// a wave moves the two doubles at its address to LDS at M0 + 16 l).  The
// load of plane k+3 is issued right after the barrier of plane k and lands
// while plane k is computed, with no registers held for it; the column
// values of the next plane are prefetched into registers alongside.
// Six ring slots (planes k-2..k+3) and two material buffers leave one
// barrier per plane.  Every staged pair of doubles must lie in one row and be
// 16-byte aligned: an even row length ni and 16-byte aligned arrays (the
// launcher checks both).  Pairs off the box load from the plane's first
// point instead (a valid address); only inactive points would read them.
// DW: rows of any length and 8-byte aligned arrays, by four-byte loads: each
// staged row of the tile (HX doubles, 2 HX dwords) is moved by one wave, lane
// l taking dwords l and, for l < 2 HX - 64, 64 + l; doubles off the box load
// the row's first one.  BATCH: the batched kernels' instantiations, the same
// code kept apart from the single-grid kernels' (hipcc's host pass rejects a
// specialization that a non-template kernel instantiated when a kernel
// template uses it as well).
#define SW4_LDS(p) ((__attribute__((address_space(3))) void*)(p))
template <int MODE, bool DW = false, bool BATCH = false>
__device__ __forceinline__ void sweep_dma_body(unsigned bid, Box b, TileMap map, int kbeg, int kend, int kc,
                                               const double* __restrict__ u, const double* __restrict__ um,
                                               const double* __restrict__ mu, const double* __restrict__ la,
                                               const double* __restrict__ rho, SgCoef sg, double* __restrict__ out,
                                               double* __restrict__ acc, double cof, double cupd, double dt2i)
{
   constexpr bool TR = MODE == DMA_TR;
   constexpr int PP = 3 * PL / 2;          // pairs of a staged displacement plane
   constexpr int NP = (PP + NT - 1) / NT;  // asynchronous loads per thread and plane
   constexpr int MP = PL / 2;              // pairs of a material plane
   static_assert(HX % 2 == 0 && MP <= NT, "pairs must tile the rows; one material load per thread");
   static_assert(2 * HX > 64 && 2 * HX <= 128, "a row is moved by one wave in two loads");
   constexpr int NW = NT / 64;  // waves of the workgroup
   __shared__ __align__(16) double su[6 * 3 * PL];  // plane kp in slot (kp - k0 + 2) % 6
   __shared__ __align__(16) double sm[2 * PL];      // mu of plane kp in half (kp - k0) & 1
   __shared__ __align__(16) double sl[2 * PL];
   __shared__ double ssx[HX], ssy[HY];
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(bid % 8) * per + (int)(bid / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave of the workgroup
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const int k0 = kbeg + tkz * kc;
   const int k1 = min(kend, k0 + kc - 1);
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;

   // pair m = tid + p NT of the staged (c, jj, ii) tile: offset of its first double
   long goff[NP];
   long moff = 0;
   // DW: this lane's dword offsets in a staged row (first and second load)
   int dwo0 = 0, dwo1 = 0;
   if constexpr (DW) {
      const int lane = tid & 63, d0 = lane >> 1, d1 = 32 + (lane >> 1);
      dwo0 = 2 * (ia + d0 <= b.ie ? d0 : 0) + (lane & 1);
      dwo1 = 2 * (ia + d1 <= b.ie ? d1 : 0) + (lane & 1);
   } else {
#pragma unroll
      for (int p = 0; p < NP; p++) {
         const int e = 2 * (tid + p * NT);
         const int c = e / PL, r = e - c * PL, jj = r / HX, ii = r - jj * HX;
         const int gi = ia + ii, gj = ja + jj;
         const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
         goff[p] = ok ? (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
      }
      const int e = 2 * tid, jj = e / HX, ii = e - jj * HX;
      const int gi = ia + ii, gj = ja + jj;
      moff = e < PL && gi <= b.ie && gj <= b.je ? (gi - b.ib) + ni * (gj - b.jb) : 0;
   }
   // DW: one staged row from its first double (lds: the row's slot, wave-uniform)
   auto dma_row = [&](const double* row, double* lds) {
      const float* g = reinterpret_cast<const float*>(row);
      __builtin_amdgcn_global_load_lds(g + dwo0, SW4_LDS(lds), 4, 0, 0);
      if ((tid & 63) < 2 * HX - 64) __builtin_amdgcn_global_load_lds(g + dwo1, SW4_LDS(lds + 32), 4, 0, 0);
   };
   auto dma_plane = [&](int kp, int slot) {
      const double* src = u + nij * (kp - b.kb);
      if constexpr (DW) {
         // rows r = (c, jj) of the plane, r = wv, wv + NW, ...; rows off the box load the last one
         double* dst = su + slot * 3 * PL;
#pragma unroll
         for (int t = 0; t < (3 * HY + NW - 1) / NW; t++) {
            const int r = wv + NW * t;
            if (r < 3 * HY) {
               const int c = r / HY, jj = r - c * HY;
               dma_row(src + npts * c + ni * (min(ja + jj, b.je) - b.jb) + (ia - b.ib), dst + r * HX);
            }
         }
      } else {
         double* dst = su + slot * 3 * PL + 2 * 64 * wv;
#pragma unroll
         for (int p = 0; p < NP; p++)
            if (tid + p * NT < PP)
               __builtin_amdgcn_global_load_lds(src + goff[p], SW4_LDS(dst + 2 * p * NT), 16, 0, 0);
      }
   };
   auto dma_mat = [&](int kp, int half) {
      const long pb = nij * (kp - b.kb);
      if constexpr (DW) {
#pragma unroll
         for (int t = 0; t < (HY + NW - 1) / NW; t++) {
            const int r = wv + NW * t;
            if (r < HY) {
               const long ro = pb + ni * (min(ja + r, b.je) - b.jb) + (ia - b.ib);
               dma_row(mu + ro, sm + half * PL + r * HX);
               dma_row(la + ro, sl + half * PL + r * HX);
            }
         }
      } else if (tid < MP) {
         __builtin_amdgcn_global_load_lds(mu + pb + moff, SW4_LDS(sm + half * PL + 2 * 64 * wv), 16, 0, 0);
         __builtin_amdgcn_global_load_lds(la + pb + moff, SW4_LDS(sl + half * PL + 2 * 64 * wv), 16, 0, 0);
      }
   };

   LdsSten s;
   if (tid >= 128 && tid < 128 + HX) ssx[tid - 128] = ia + tid - 128 <= b.ie ? sg.strx[ia + tid - 128 - b.ib] : 1.0;
   if (tid >= 192 && tid < 192 + HY) ssy[tid - 192] = ja + tid - 192 <= b.je ? sg.stry[ja + tid - 192 - b.jb] : 1.0;
#pragma unroll
   for (int q = 0; q < 5; q++) dma_plane(k0 - 2 + q, q);
   dma_mat(k0, 0);
#pragma unroll
   for (int q = 0; q < 4; q++) {
      const long pq = nij * (k0 - 2 + q - b.kb);
      s.muc[q] = active ? mu[col + pq] : 0.0;
      s.lac[q] = active ? la[col + pq] : 0.0;
      s.szv[q] = sg.strz[k0 - 2 + q - b.kb];
   }
   s.psx = ssx + tx + 2;
   s.psy = ssy + ty + 2;
   // column values of plane k+2, loaded one plane ahead
   double cmu = mu[col + nij * (k0 + 2 - b.kb)], cla = la[col + nij * (k0 + 2 - b.kb)];
   double csz = sg.strz[k0 + 2 - b.kb];

   // the results of plane k are stored after the barrier of plane k+1, so
   // that this barrier (which waits for every outstanding memory operation)
   // does not wait for stores issued just before it
   double ov[3] = {0, 0, 0}, av[3] = {0, 0, 0};
   long op = -1;  // offset of the held results (-1: none)
   const bool wacc = MODE == DMA_PRED && acc != nullptr;
   for (int k = k0; k <= k1; k++) {
      const int sl0 = (k - k0) % 6;  // slot of plane k-2
      s.muc[4] = active ? cmu : 0.0;
      s.lac[4] = active ? cla : 0.0;
      s.szv[4] = csz;
      __syncthreads();  // plane k+2 and the material of plane k have landed; plane k-1 is done
      if (op >= 0) {
#pragma unroll
         for (int c = 0; c < 3; c++) {
            out[op + npts * c] = ov[c];
            if (wacc) acc[op + npts * c] = av[c];
         }
      }
      // centre values (um, or up for the corrector; rho) of plane k: issued first, needed only after
      // the stencil, so the stencil covers their latency
      double ctr[3] = {0, 0, 0}, crho = 1.0;
      if (!TR) {
         const long p = col + nij * (k - b.kb);  // col = 0 where inactive: a valid address
#pragma unroll
         for (int c = 0; c < 3; c++) ctr[c] = um[p + npts * c];
         crho = rho[p];
      }
      if (k < k1) {
         dma_plane(k + 3, (sl0 + 5) % 6);
         dma_mat(k + 1, (k + 1 - k0) & 1);
         const long pq = nij * (k + 3 - b.kb);
         cmu = mu[col + pq];  // value unused where inactive
         cla = la[col + pq];
         csz = sg.strz[k + 3 - b.kb];
      }
      if (active) {
         s.p0 = su + (sl0 * 3) * PL + o;
         s.p1 = su + (((sl0 + 1) % 6) * 3) * PL + o;
         s.p2 = su + (((sl0 + 2) % 6) * 3) * PL + o;
         s.p3 = su + (((sl0 + 3) % 6) * 3) * PL + o;
         s.p4 = su + (((sl0 + 4) % 6) * 3) * PL + o;
         s.pm = sm + ((k - k0) & 1) * PL + o;
         s.pl = sl + ((k - k0) & 1) * PL + o;
         double r[3];
         rhs4_interior_body(s, r);
         const double sc = TR ? 1.0 / (s.psx[0] * s.psy[0] * s.szv[2]) : cupd / crho;
#pragma unroll
         for (int c = 0; c < 3; c++) {
            if (MODE == DMA_TR)
               ov[c] = sc * (cof * r[c]);
            else if (MODE == DMA_CORR)
               ov[c] = ctr[c] + sc * (cof * r[c]);
            else
               ov[c] = 2 * s.p2[c * PL] - ctr[c] + sc * (cof * r[c]);
            if (MODE == DMA_PRED) av[c] = accel_of(ov[c], s.p2[c * PL], ctr[c], dt2i);
         }
         op = col + nij * (k - b.kb);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
         s.muc[q] = s.muc[q + 1];
         s.lac[q] = s.lac[q + 1];
         s.szv[q] = s.szv[q + 1];
      }
   }
   if (op >= 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
         out[op + npts * c] = ov[c];
         if (wacc) acc[op + npts * c] = av[c];
      }
   }
}

template <int MODE>
__global__ void __launch_bounds__(NT, 2)
k_sweep_dma(Box b, TileMap map, int kbeg, int kend, int kc, const double* __restrict__ u,
            const double* __restrict__ um, const double* __restrict__ mu, const double* __restrict__ la,
            const double* __restrict__ rho, SgCoef sg, double* __restrict__ out, double* __restrict__ acc, double cof,
            double cupd, double dt2i)
{
   sweep_dma_body<MODE>(blockIdx.x, b, map, kbeg, kend, kc, u, um, mu, la, rho, sg, out, acc, cof, cupd, dt2i);
}

// The same sweep with four-byte loads: rows of any length (DW above).
template <int MODE>
__global__ void __launch_bounds__(NT, 2)
k_sweep_dma_dw(Box b, TileMap map, int kbeg, int kend, int kc, const double* __restrict__ u,
               const double* __restrict__ um, const double* __restrict__ mu, const double* __restrict__ la,
               const double* __restrict__ rho, SgCoef sg, double* __restrict__ out, double* __restrict__ acc,
               double cof, double cupd, double dt2i)
{
   sweep_dma_body<MODE, true>(blockIdx.x, b, map, kbeg, kend, kc, u, um, mu, la, rho, sg, out, acc, cof, cupd, dt2i);
}

// The asynchronous loads need whole 16-byte pairs: even row length, aligned arrays.
static bool dma_fits(const Box& b, const double* u, const double* mu, const double* la)
{
   auto al = [](const double* p) { return ((uintptr_t)p & 15) == 0; };
   return b.ni % 2 == 0 && al(u) && al(mu) && al(la);
}

// Supergrid damping of the corrector on planes kbeg..kend as a k-marching LDS
// sweep (ew-cfromfort.C:985-1041): up -= damping(u - um).  d = u - um of plane
// k and rho of plane k are staged with their (i,j) halo in LDS; the thread's
// own column of d for k-2..k+2 and rho at k+-1 live in registers.  Small LDS
// and register footprint, so several workgroups share a CU and hide latency.
// Tiles of 32 x DTY columns (DTY * 32 threads); taller tiles re-read less halo.
// Column tile (tix, tjy), planes k0..k1.
template <int DTY>
__device__ __forceinline__ void damp_lds_tile(int tix, int tjy, int k0, int k1, Box b, double* __restrict__ up,
                                              const double* __restrict__ u, const double* __restrict__ um,
                                              const double* __restrict__ rho, SgCoef sg, double beta)
{
   constexpr int TY = DTY, HY = TY + 4, NT = TX * TY, PL = HX * HY;
   constexpr int NU = (3 * PL + NT - 1) / NT;
   constexpr int NM = (PL + NT - 1) / NT;
   __shared__ double sdp[3 * PL];
   __shared__ double srp[PL];
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;
   const int ii = i - b.ib, jj = j - b.jb;
   double dq[5][3];
   double rkm = 0, rkp = 0;
#pragma unroll
   for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const long g = col + nij * (k0 - 2 + q - b.kb) + npts * c;
         const double a = u[g], e = um[g];
         dq[q][c] = active ? a - e : 0.0;
      }
   {
      const double a = rho[col + nij * (k0 - 1 - b.kb)];
      rkm = active ? a : 0.0;
   }
   double gx[5] = {0, 0, 0, 0, 0}, gy[5] = {0, 0, 0, 0, 0};
   if (active) {
      gx[0] = sg.dcx[ii - 1]; gx[1] = sg.dcx[ii]; gx[2] = sg.dcx[ii + 1]; gx[3] = sg.strx[ii]; gx[4] = sg.cox[ii];
      gy[0] = sg.dcy[jj - 1]; gy[1] = sg.dcy[jj]; gy[2] = sg.dcy[jj + 1]; gy[3] = sg.stry[jj]; gy[4] = sg.coy[jj];
   }
   for (int k = k0; k <= k1; k++) {
      const long pb = nij * (k - b.kb);
      // this point's values of up, updated below: loaded with the plane so
      // that their latency overlaps the staging loads'
      double upc[3];
#pragma unroll
      for (int c = 0; c < 3; c++) upc[c] = up[col + pb + npts * c];  // col = 0 where inactive: a valid address
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const int e = tid + p * NT;
         const int c = e / PL, r = e - c * PL, y = r / HX, x = r - y * HX;
         const int gi = ia + x, gj = ja + y;
         const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
         const long g = ok ? pb + (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
         const double a = u[g], m = um[g];
         if (e < 3 * PL) sdp[e] = ok ? a - m : 0.0;
      }
#pragma unroll
      for (int p = 0; p < NM; p++) {
         const int e = tid + p * NT;
         const int y = e / HX, x = e - y * HX;
         const int gi = ia + x, gj = ja + y;
         const bool ok = e < PL && gi <= b.ie && gj <= b.je;
         const double a = rho[ok ? pb + (gi - b.ib) + ni * (gj - b.jb) : 0];
         if (e < PL) srp[e] = ok ? a : 0.0;
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const long g = col + nij * (k + 2 - b.kb) + npts * c;
         const double a = u[g], m = um[g];
         dq[4][c] = active ? a - m : 0.0;
      }
      {
         const double a = rho[col + nij * (k + 1 - b.kb)];
         rkp = active ? a : 0.0;
      }
      const int kz = k - b.kb;
      const double gz0 = sg.dcz[kz - 1], gz1 = sg.dcz[kz], gz2 = sg.dcz[kz + 1], gz3 = sg.strz[kz], gz4 = sg.coz[kz];
      __syncthreads();
      double r0 = 0;
      if (active) {
         r0 = srp[o];
         const double birho = beta / r0;
         const double rxp = srp[o + 1] * gx[2], rx0 = r0 * gx[1], rxm = srp[o - 1] * gx[0];
         const double ryp = srp[o + HX] * gy[2], ry0 = r0 * gy[1], rym = srp[o - HX] * gy[0];
         const double rzp = rkp * gz2, rz0 = r0 * gz1, rzm = rkm * gz0;
         const double wx = gx[3] * gy[4] * gz4, wy = gy[3] * gx[4] * gz4, wz = gz3 * gx[4] * gy[4];
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const double* q = sdp + c * PL + o;
            const double ax = rxp * (q[2] - 2 * q[1] + q[0]) - 2 * rx0 * (q[1] - 2 * q[0] + q[-1]) +
                              rxm * (q[0] - 2 * q[-1] + q[-2]);
            const double ay = ryp * (q[2 * HX] - 2 * q[HX] + q[0]) - 2 * ry0 * (q[HX] - 2 * q[0] + q[-HX]) +
                              rym * (q[0] - 2 * q[-HX] + q[-2 * HX]);
            const double az = rzp * (dq[4][c] - 2 * dq[3][c] + dq[2][c]) -
                              2 * rz0 * (dq[3][c] - 2 * dq[2][c] + dq[1][c]) +
                              rzm * (dq[2][c] - 2 * dq[1][c] + dq[0][c]);
            up[col + pb + npts * c] = upc[c] - birho * (wx * ax + wy * ay + wz * az);
         }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
         for (int c = 0; c < 3; c++) dq[q][c] = dq[q + 1][c];
      rkm = r0;
   }
}

template <int DTY>
__device__ __forceinline__ void damp_lds_body(unsigned bid, Box b, TileMap map, int kbeg, int kend, int kc,
                                              double* __restrict__ up, const double* __restrict__ u,
                                              const double* __restrict__ um, const double* __restrict__ rho,
                                              SgCoef sg, double beta)
{
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(bid % 8) * per + (int)(bid / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int k0 = kbeg + tkz * kc;
   damp_lds_tile<DTY>(tix, tjy, k0, min(kend, k0 + kc - 1), b, up, u, um, rho, sg, beta);
}

// The damping sweep with its planes brought to LDS by asynchronous
// global-to-LDS loads (as sweep_dma_body): raw u, um and rho of plane k+1
// land while plane k is computed, d = u - um is formed from LDS; the column
// values (d at k+3, rho at k+2) and this point's up of plane k+1 are loaded
// one plane ahead into registers, and the results of plane k are stored
// after the next barrier.  Two slots of 7 staged values per point (48 KB):
// three workgroups per CU.  Same conditions as sweep_dma_body (even ni,
// 16-byte aligned arrays).
// Column tile (tix, tjy), planes k0..k1.
__device__ __forceinline__ void damp_dma_tile(int tix, int tjy, int k0, int k1, Box b, double* __restrict__ up,
                                              const double* __restrict__ u, const double* __restrict__ um,
                                              const double* __restrict__ rho, SgCoef sg, double beta)
{
   constexpr int PP = 3 * PL / 2, NP = (PP + NT - 1) / NT, MP = PL / 2;
   static_assert(HX % 2 == 0 && MP <= NT, "pairs must tile the rows; one rho load per thread");
   __shared__ __align__(16) double su[2 * 3 * PL];  // u of plane kp in half (kp - k0) & 1
   __shared__ __align__(16) double sq[2 * 3 * PL];  // um
   __shared__ __align__(16) double sr[2 * PL];      // rho
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;
   const int ii = i - b.ib, jj = j - b.jb;
   long goff[NP];
#pragma unroll
   for (int p = 0; p < NP; p++) {
      const int e = 2 * (tid + p * NT);
      const int c = e / PL, r = e - c * PL, y = r / HX, x = r - y * HX;
      const int gi = ia + x, gj = ja + y;
      const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
      goff[p] = ok ? (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
   }
   long roff;
   {
      const int e = 2 * tid, y = e / HX, x = e - y * HX;
      const int gi = ia + x, gj = ja + y;
      roff = e < PL && gi <= b.ie && gj <= b.je ? (gi - b.ib) + ni * (gj - b.jb) : 0;
   }
   auto dma_plane = [&](int kp, int half) {
      const long pb = nij * (kp - b.kb);
      double* du = su + half * 3 * PL + 2 * 64 * wv;
      double* dq = sq + half * 3 * PL + 2 * 64 * wv;
#pragma unroll
      for (int p = 0; p < NP; p++)
         if (tid + p * NT < PP) {
            __builtin_amdgcn_global_load_lds(u + pb + goff[p], SW4_LDS(du + 2 * p * NT), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(um + pb + goff[p], SW4_LDS(dq + 2 * p * NT), 16, 0, 0);
         }
      if (tid < MP) __builtin_amdgcn_global_load_lds(rho + pb + roff, SW4_LDS(sr + half * PL + 2 * 64 * wv), 16, 0, 0);
   };
   // own column: d at k-2..k+2, rho at k-1 and k+1 (0 where inactive)
   double dq[5][3];
#pragma unroll
   for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const long g = col + nij * (k0 - 2 + q - b.kb) + npts * c;
         const double a = u[g], e = um[g];
         dq[q][c] = active ? a - e : 0.0;
      }
   double rkm = active ? rho[col + nij * (k0 - 1 - b.kb)] : 0.0;
   double gx[5] = {0, 0, 0, 0, 0}, gy[5] = {0, 0, 0, 0, 0};
   if (active) {
      gx[0] = sg.dcx[ii - 1]; gx[1] = sg.dcx[ii]; gx[2] = sg.dcx[ii + 1]; gx[3] = sg.strx[ii]; gx[4] = sg.cox[ii];
      gy[0] = sg.dcy[jj - 1]; gy[1] = sg.dcy[jj]; gy[2] = sg.dcy[jj + 1]; gy[3] = sg.stry[jj]; gy[4] = sg.coy[jj];
   }
   dma_plane(k0, 0);
   // one plane ahead: u, um at k+2 (column), rho at k+1, up at k (this point)
   double nu[3], nm[3], nup[3], nrk;
#pragma unroll
   for (int c = 0; c < 3; c++) {
      nu[c] = u[col + nij * (k0 + 2 - b.kb) + npts * c];  // col = 0 where inactive: a valid address
      nm[c] = um[col + nij * (k0 + 2 - b.kb) + npts * c];
      nup[c] = up[col + nij * (k0 - b.kb) + npts * c];
   }
   nrk = rho[col + nij * (k0 + 1 - b.kb)];
   double ov[3] = {0, 0, 0};
   long op = -1;  // offset of the held results (-1: none)
   for (int k = k0; k <= k1; k++) {
      const int h = (k - k0) & 1;
      __syncthreads();  // plane k has landed; plane k-1 is done
      if (op >= 0) {
#pragma unroll
         for (int c = 0; c < 3; c++) up[op + npts * c] = ov[c];
      }
      double upc[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
         dq[4][c] = active ? nu[c] - nm[c] : 0.0;
         upc[c] = nup[c];
      }
      const double rkp = active ? nrk : 0.0;
      if (k < k1) {
         dma_plane(k + 1, h ^ 1);
#pragma unroll
         for (int c = 0; c < 3; c++) {
            nu[c] = u[col + nij * (k + 3 - b.kb) + npts * c];
            nm[c] = um[col + nij * (k + 3 - b.kb) + npts * c];
            nup[c] = up[col + nij * (k + 1 - b.kb) + npts * c];
         }
         nrk = rho[col + nij * (k + 2 - b.kb)];
      }
      const int kz = k - b.kb;
      const double gz0 = sg.dcz[kz - 1], gz1 = sg.dcz[kz], gz2 = sg.dcz[kz + 1], gz3 = sg.strz[kz], gz4 = sg.coz[kz];
      const double* pr = sr + h * PL + o;
      double r0 = 0;
      if (active) {
         r0 = pr[0];
         const double birho = beta / r0;
         const double rxp = pr[1] * gx[2], rx0 = r0 * gx[1], rxm = pr[-1] * gx[0];
         const double ryp = pr[HX] * gy[2], ry0 = r0 * gy[1], rym = pr[-HX] * gy[0];
         const double rzp = rkp * gz2, rz0 = r0 * gz1, rzm = rkm * gz0;
         const double wx = gx[3] * gy[4] * gz4, wy = gy[3] * gx[4] * gz4, wz = gz3 * gx[4] * gy[4];
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const double* pu = su + (h * 3 + c) * PL + o;
            const double* pm = sq + (h * 3 + c) * PL + o;
            double q[9];  // d at x-2..x+2 (q[0..4]), y-2, y-1, y+1, y+2 (q[5..8])
#pragma unroll
            for (int d = -2; d <= 2; d++) q[d + 2] = pu[d] - pm[d];
            q[5] = pu[-2 * HX] - pm[-2 * HX];
            q[6] = pu[-HX] - pm[-HX];
            q[7] = pu[HX] - pm[HX];
            q[8] = pu[2 * HX] - pm[2 * HX];
            const double ax = rxp * (q[4] - 2 * q[3] + q[2]) - 2 * rx0 * (q[3] - 2 * q[2] + q[1]) +
                              rxm * (q[2] - 2 * q[1] + q[0]);
            const double ay = ryp * (q[8] - 2 * q[7] + q[2]) - 2 * ry0 * (q[7] - 2 * q[2] + q[6]) +
                              rym * (q[2] - 2 * q[6] + q[5]);
            const double az = rzp * (dq[4][c] - 2 * dq[3][c] + dq[2][c]) -
                              2 * rz0 * (dq[3][c] - 2 * dq[2][c] + dq[1][c]) +
                              rzm * (dq[2][c] - 2 * dq[1][c] + dq[0][c]);
            ov[c] = upc[c] - birho * (wx * ax + wy * ay + wz * az);
         }
         op = col + nij * (k - b.kb);
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
         for (int c = 0; c < 3; c++) dq[q][c] = dq[q + 1][c];
      rkm = r0;
   }
   if (op >= 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) up[op + npts * c] = ov[c];
   }
}

__device__ __forceinline__ void damp_dma_body(unsigned bid, Box b, TileMap map, int kbeg, int kend, int kc,
                                              double* __restrict__ up, const double* __restrict__ u,
                                              const double* __restrict__ um, const double* __restrict__ rho,
                                              SgCoef sg, double beta)
{
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(bid % 8) * per + (int)(bid / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int k0 = kbeg + tkz * kc;
   damp_dma_tile(tix, tjy, k0, min(kend, k0 + kc - 1), b, up, u, um, rho, sg, beta);
}

__global__ void __launch_bounds__(NT, 3)
k_damp_dma(Box b, TileMap map, int kbeg, int kend, int kc, double* __restrict__ up, const double* __restrict__ u,
           const double* __restrict__ um, const double* __restrict__ rho, SgCoef sg, double beta)
{
   damp_dma_body(blockIdx.x, b, map, kbeg, kend, kc, up, u, um, rho, sg, beta);
}

// The damping sweep over a list of tiles (damp_tiles): the XCD of a block is
// its id mod 8 and each XCD takes one contiguous run of the list.
__global__ void __launch_bounds__(NT, 3)
k_damp_dma_tiles(Box b, const DampTile* __restrict__ tiles, int ntiles, double* __restrict__ up,
                 const double* __restrict__ u, const double* __restrict__ um, const double* __restrict__ rho,
                 SgCoef sg, double beta)
{
   const int per = (ntiles + 7) / 8;
   const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
   if (t >= ntiles) return;
   const DampTile d = tiles[t];
   damp_dma_tile(d.tix, d.tjy, d.k0, d.k1, b, up, u, um, rho, sg, beta);
}

template <int DTY>
__global__ void __launch_bounds__(32 * DTY, 4)
k_damp_lds(Box b, TileMap map, int kbeg, int kend, int kc, double* __restrict__ up, const double* __restrict__ u,
           const double* __restrict__ um, const double* __restrict__ rho, SgCoef sg, double beta)
{
   damp_lds_body<DTY>(blockIdx.x, b, map, kbeg, kend, kc, up, u, um, rho, sg, beta);
}

// ---------------------------------------------------------------------------
// Batched sweeps: the tiles of several grids in one launch.  Shot s owns the
// block ids [grid0, grid0 + 8 ceil(ntiles/8)), so the XCD of a block (its id
// mod 8) is the same as in a launch of that shot alone.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_shot(const SweepShot* __restrict__ shots, int nshots, unsigned bid)
{
   int s = 0;
   while (s + 1 < nshots && bid >= (unsigned)shots[s + 1].grid0) s++;
   return s;
}

template <bool CORR>
__global__ void __launch_bounds__(NT, 2) k_sweep_batch(const SweepShot* __restrict__ shots, int nshots)
{
   const SweepShot& h = shots[find_shot(shots, nshots, blockIdx.x)];
   sweep_lds_body<CORR, false, false>(blockIdx.x - h.grid0, h.b, h.map, h.kbeg, h.kend, h.kc, h.upin, h.u, h.um, h.mu,
                                      h.la, h.rho, h.sg, h.out, h.cof, h.dt2i, h.cupd, h.beta);
}

// A shot's tile of the batched damping sweep: from its tile list where it
// has one (entry t of the list, each XCD taking one contiguous run), else
// from its tile map.  false: no tile for this block.
__device__ __forceinline__ bool damp_batch_tile(const SweepShot& h, unsigned bid, int& tix, int& tjy, int& k0,
                                                int& k1)
{
   const int per = (int)((h.map.ntiles + 7) / 8);
   const int t = (int)(bid % 8) * per + (int)(bid / 8);
   if (t >= h.map.ntiles) return false;
   if (h.tiles) {
      const DampTile d = h.tiles[t];
      tix = d.tix;
      tjy = d.tjy;
      k0 = d.k0;
      k1 = d.k1;
   } else {
      const int tr = t / h.map.ntx, tkz = tr / h.map.nty;
      tix = t % h.map.ntx;
      tjy = tr % h.map.nty;
      k0 = h.kbeg + tkz * h.kc;
      k1 = min(h.kend, k0 + h.kc - 1);
   }
   return true;
}

__global__ void __launch_bounds__(32 * 8, 4) k_damp_batch(const SweepShot* __restrict__ shots, int nshots)
{
   const SweepShot& h = shots[find_shot(shots, nshots, blockIdx.x)];
   int tix, tjy, k0, k1;
   if (!damp_batch_tile(h, blockIdx.x - h.grid0, tix, tjy, k0, k1)) return;
   damp_lds_tile<8>(tix, tjy, k0, k1, h.b, h.out, h.u, h.um, h.rho, h.sg, h.beta);
}

// The register-staged damping over a tile list of one grid (entry t of the
// list, each XCD taking one contiguous run), for rows of odd length.
__global__ void __launch_bounds__(32 * 8, 4) k_damp_lds_tiles(Box b, const DampTile* __restrict__ tiles, int ntiles,
                                                              double* __restrict__ up, const double* __restrict__ u,
                                                              const double* __restrict__ um,
                                                              const double* __restrict__ rho, SgCoef sg, double beta)
{
   const int per = (ntiles + 7) / 8;
   const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
   if (t >= ntiles) return;
   const DampTile d = tiles[t];
   damp_lds_tile<8>(d.tix, d.tjy, d.k0, d.k1, b, up, u, um, rho, sg, beta);
}

int fill_sweep_shot(bool damping, const Box& b, int k1, int k2, int kc, const double* upin, const double* u,
                    const double* um, const double* mu, const double* la, const double* rho, const SgCoef& sg,
                    double* out, double h, double dt, double beta, bool corr, int grid0, SweepShot* sh,
                    const DampTile* tiles, int ntiles, double* acc)
{
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   if (kc <= 0) kc = 64;
   sh->b = b;
   sh->map.ntx = (int)cdiv(nx, TX);
   sh->map.nty = (int)cdiv(ny, damping ? 8 : TY);
   sh->map.ntiles = k2 >= k1 ? sh->map.ntx * sh->map.nty * (int)cdiv(k2 - k1 + 1, kc) : 0;
   sh->tiles = damping && tiles ? tiles : nullptr;
   if (sh->tiles) sh->map.ntiles = k2 >= k1 ? ntiles : 0;
   sh->kbeg = k1;
   sh->kend = k2;
   sh->kc = kc;
   sh->upin = upin;
   sh->u = u;
   sh->um = um;
   sh->mu = mu;
   sh->la = la;
   sh->rho = rho;
   sh->sg = sg;
   sh->out = out;
   sh->acc = damping ? nullptr : acc;
   sh->cof = 1.0 / (h * h);
   sh->dt2i = 1.0 / (dt * dt);
   sh->cupd = corr ? dt * dt * dt * dt / 12 : dt * dt;
   sh->beta = beta;
   sh->grid0 = grid0;
   return grid0 + 8 * (int)cdiv(sh->map.ntiles, 8);
}

// Batched predictor (kind 0) and damping (kind 2) sweeps of grids whose
// planes take the asynchronous global-to-LDS loads (async_loads_fit).
// (the predictor stores the acceleration in acc where the shot has one)
__global__ void __launch_bounds__(NT, 2) k_sweep_dma_batch(const SweepShot* __restrict__ shots, int nshots)
{
   const SweepShot& h = shots[find_shot(shots, nshots, blockIdx.x)];
   sweep_dma_body<DMA_PRED, false, true>(blockIdx.x - h.grid0, h.b, h.map, h.kbeg, h.kend, h.kc, h.u, h.um, h.mu,
                                         h.la, h.rho, h.sg, h.out, h.acc, h.cof, h.cupd, h.dt2i);
}

// Batched corrector staging the acceleration its predictor stored (acc),
// for shots that store it.
__global__ void __launch_bounds__(NT, 2) k_sweep_dma_corr_batch(const SweepShot* __restrict__ shots, int nshots)
{
   const SweepShot& h = shots[find_shot(shots, nshots, blockIdx.x)];
   sweep_dma_body<DMA_CORR, false, true>(blockIdx.x - h.grid0, h.b, h.map, h.kbeg, h.kend, h.kc, h.acc, h.upin,
                                         h.mu, h.la, h.rho, h.sg, h.out, nullptr, h.cof, h.cupd, h.dt2i);
}

__global__ void __launch_bounds__(NT, 3) k_damp_dma_batch(const SweepShot* __restrict__ shots, int nshots)
{
   const SweepShot& h = shots[find_shot(shots, nshots, blockIdx.x)];
   int tix, tjy, k0, k1;
   if (!damp_batch_tile(h, blockIdx.x - h.grid0, tix, tjy, k0, k1)) return;
   damp_dma_tile(tix, tjy, k0, k1, h.b, h.out, h.u, h.um, h.rho, h.sg, h.beta);
}

bool async_loads_fit(const Box& b, const double* const* arrays, int n)
{
   if (b.ni % 2 != 0) return false;
   for (int q = 0; q < n; q++)
      if (((uintptr_t)arrays[q] & 15) != 0) return false;
   return true;
}

int launch_sweep_batch_async(int kind, const SweepShot* d_shots, int nshots, int nblocks, hipStream_t s)
{
   if (nblocks <= 0) return OK;
   if (kind == 0)
      k_sweep_dma_batch<<<nblocks, NT, 0, s>>>(d_shots, nshots);
   else if (kind == 1)
      k_sweep_dma_corr_batch<<<nblocks, NT, 0, s>>>(d_shots, nshots);
   else
      k_damp_dma_batch<<<nblocks, NT, 0, s>>>(d_shots, nshots);
   return check_hip(hipGetLastError(), "batched asynchronous-load sweep launch");
}

int launch_sweep_batch(int kind, const SweepShot* d_shots, int nshots, int nblocks, hipStream_t s)
{
   if (nblocks <= 0) return OK;
   if (kind == 0)
      k_sweep_batch<false><<<nblocks, NT, 0, s>>>(d_shots, nshots);
   else if (kind == 1)
      k_sweep_batch<true><<<nblocks, NT, 0, s>>>(d_shots, nshots);
   else
      k_damp_batch<<<nblocks, 32 * 8, 0, s>>>(d_shots, nshots);
   return check_hip(hipGetLastError(), "batched sweep launch");
}

// The acceleration where the predictor sweep of planes k1..k2 did not form
// it: the four planes k1-2, k1-1, k2+1, k2+2 whole (all_planes: every plane
// outside k1..k2) and (frame) on the planes k1..k2 the two-point frame of
// columns next to the sides.  Thread q runs
// over the planes' points first, then over the frame's columns (the rows
// jb, jb+1, je-1, je whole, then the four side columns of each row between,
// on consecutive threads, so that a row's two cache lines are read once).
__global__ void k_accel_fix(Box b, int k1, int k2, int nlo, int nhi, bool frame, const double* __restrict__ up,
                            const double* __restrict__ u, const double* __restrict__ um, double dt2i,
                            double* __restrict__ acc)
{
   const long ni = b.ni, nj = b.je - b.jb + 1, nij = b.nij, npts = b.npts;
   const long nplane = (long)(nlo + nhi) * nij;  // planes k1-nlo..k1-1 and k2+1..k2+nhi
   const long nframe = 4 * (ni + nj - 4);  // columns of the frame in one plane
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   long p;
   if (q < nplane) {
      const int s = (int)(q / nij);
      const int k = s < nlo ? k1 - nlo + s : k2 + 1 + (s - nlo);
      p = nij * (k - b.kb) + q % nij;
   } else {
      const long r = q - nplane;
      if (!frame || r >= nframe * (k2 - k1 + 1)) return;
      const int k = k1 + (int)(r / nframe);
      const long f = r % nframe;
      long ii, jj;  // offsets from (ib, jb)
      if (f < 4 * ni) {  // rows jb, jb+1, je-1, je, every i
         const long w = f / ni;
         ii = f % ni;
         jj = w < 2 ? w : nj - 4 + w;
      } else {  // columns ib, ib+1, ie-1, ie on the rows between: four consecutive threads per row
         const long g = f - 4 * ni, w = g % 4;
         jj = 2 + g / 4;
         ii = w < 2 ? w : ni - 4 + w;
      }
      p = nij * (k - b.kb) + ni * jj + ii;
   }
#pragma unroll
   for (int c = 0; c < 3; c++) acc[p + npts * c] = accel_of(up[p + npts * c], u[p + npts * c], um[p + npts * c], dt2i);
}

int launch_accel_fix(const Box& b, int k1, int k2, bool frame, bool all_planes, const double* up, const double* u,
                     const double* um, double dt, double* acc, hipStream_t s)
{
   if (k2 < k1) return set_error(E_INVAL, "acceleration fix: no sweep planes");
   if (k1 - 2 < b.kb || k2 + 2 > b.ke || b.ni < 5 || b.je - b.jb + 1 < 5)
      return set_error(E_INVAL, "acceleration fix: planes or box out of range");
   const int nlo = all_planes ? k1 - b.kb : 2, nhi = all_planes ? b.ke - k2 : 2;
   const long n = (long)(nlo + nhi) * b.nij + (frame ? 4 * (b.ni + (b.je - b.jb + 1) - 4) * (long)(k2 - k1 + 1) : 0);
   k_accel_fix<<<cdiv(n, 256), 256, 0, s>>>(b, k1, k2, nlo, nhi, frame, up, u, um, 1.0 / (dt * dt), acc);
   return check_hip(hipGetLastError(), "acceleration fix launch");
}

std::vector<DampTile> damp_tiles(const Box& b, int k1, int k2, int kc, const double* dcx, const double* dcy,
                                 const double* dcz)
{
   if (kc <= 0) kc = 64;
   // point q (offset from the box start) can be damped along an axis when the
   // profile is nonzero at q-1, q or q+1 (the stencil of addsgd4)
   auto damped = [](const double* dc, int n) {
      std::vector<char> on(n, 0);
      for (int q = 1; q + 1 < n; q++) on[q] = dc[q - 1] != 0.0 || dc[q] != 0.0 || dc[q + 1] != 0.0;
      return on;
   };
   const int ni = b.ie - b.ib + 1, nj = b.je - b.jb + 1, nk = b.ke - b.kb + 1;
   const std::vector<char> ox = damped(dcx, ni), oy = damped(dcy, nj), oz = damped(dcz, nk);
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   const int ntx = (int)cdiv(nx, TX), nty = (int)cdiv(ny, TY);
   std::vector<DampTile> out;
   std::vector<std::vector<DampTile>> per_chunk((size_t)cdiv(k2 - k1 + 1, kc));
   for (int tjy = 0; tjy < nty; tjy++)
      for (int tix = 0; tix < ntx; tix++) {
         bool side = false;  // the tile's columns meet an x or y layer: every plane
         for (int q = 2 + tix * TX; q < std::min(2 + (tix + 1) * TX, ni - 2); q++) side = side || ox[q];
         for (int q = 2 + tjy * TY; q < std::min(2 + (tjy + 1) * TY, nj - 2); q++) side = side || oy[q];
         for (int c = 0; c < (int)per_chunk.size(); c++) {
            const int ka = k1 + c * kc, kb = std::min(k2, ka + kc - 1);
            int lo = kb + 1, hi = ka - 1;  // planes of the chunk to run
            for (int k = ka; k <= kb; k++)
               if (side || oz[k - b.kb]) {
                  lo = std::min(lo, k);
                  hi = std::max(hi, k);
               }
            if (lo <= hi) per_chunk[c].push_back(DampTile{tix, tjy, lo, hi});
         }
      }
   // chunk-major, then j, then i: neighbouring tiles stay neighbours in the list
   for (auto& v : per_chunk) out.insert(out.end(), v.begin(), v.end());
   return out;
}

int launch_damp_lds(const Box& b, int k1, int k2, int kc, double* up, const double* u, const double* um,
                    const double* rho, const SgCoef& sg, double beta, hipStream_t s, const DampTile* tiles,
                    int ntiles)
{
   if (k2 < k1 || beta == 0) return OK;
   if (tiles && dma_fits(b, u, um, rho) && ((uintptr_t)up & 15) == 0) {
      if (ntiles > 0)
         k_damp_dma_tiles<<<8 * cdiv(ntiles, 8), NT, 0, s>>>(b, tiles, ntiles, up, u, um, rho, sg, beta);
      return check_hip(hipGetLastError(), "damping sweep launch");
   }
   if (tiles) {  // odd rows: the register-staged tile over the listed tiles only
      if (ntiles > 0)
         k_damp_lds_tiles<<<8 * cdiv(ntiles, 8), 32 * 8, 0, s>>>(b, tiles, ntiles, up, u, um, rho, sg, beta);
      return check_hip(hipGetLastError(), "damping sweep launch");
   }
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   if (kc <= 0) kc = 64;
   constexpr int DTY = 8;
   TileMap map;
   map.ntx = (int)cdiv(nx, TX);
   map.nty = (int)cdiv(ny, DTY);
   map.ntiles = map.ntx * map.nty * (int)cdiv(k2 - k1 + 1, kc);
   static_assert(DTY == TY, "the asynchronous-load damping sweep uses the sweeps' tile");
   if (dma_fits(b, u, um, rho) && ((uintptr_t)up & 15) == 0)
      k_damp_dma<<<8 * cdiv(map.ntiles, 8), NT, 0, s>>>(b, map, k1, k2, kc, up, u, um, rho, sg, beta);
   else
      k_damp_lds<DTY><<<8 * cdiv(map.ntiles, 8), 32 * DTY, 0, s>>>(b, map, k1, k2, kc, up, u, um, rho, sg, beta);
   return check_hip(hipGetLastError(), "damping sweep launch");
}

// ---------------------------------------------------------------------------
// SBP boundary rows (k = 1..6 below a free surface, or the six rows above
// one at the bottom) of both sweeps, LDS-staged.  Row space: plane q = 0..8
// is grid plane KK(q) = kbase + sgn (q - 1) (rhs4sg_rev.C:353-629); row kr
// lives on q = kr.  One workgroup stages the nine planes of the displacement
// (or of the acceleration, corrector) and the eight material planes of its
// 32 x 8 tile once and evaluates all six rows of its columns from LDS.
// ---------------------------------------------------------------------------
struct SlabSten {
   const double* ub;  // staged plane q = 0, this thread's point
   const double* mb;  // mu plane q = 1
   const double* lb;
   const double* psx;
   const double* psy;
   __device__ __forceinline__ double u(int c, int di, int dj, int q) const
   {
      return ub[(q * 3 + c) * PL + dj * HX + di];
   }
   __device__ __forceinline__ double mu(int di, int dj, int q) const { return mb[(q - 1) * PL + dj * HX + di]; }
   __device__ __forceinline__ double la(int di, int dj, int q) const { return lb[(q - 1) * PL + dj * HX + di]; }
   __device__ __forceinline__ double sx(int d) const { return psx[d]; }
   __device__ __forceinline__ double sy(int d) const { return psy[d]; }
};

// CORR = false: up = 2u - um + dt2/rho cof L(u) on the six rows;
// CORR = true:  upo = up + dt4/(12 rho) cof L(acc) - damping(u, um) with
// acc = (up - 2u + um) dt2i staged in the slab; DAMP = false: without the
// damping (a damping sweep over these rows follows; the damping's registers
// would push this kernel's 256 into 130 accumulator-register copies).
template <bool CORR, bool DAMP = CORR>
__global__ void __launch_bounds__(NT, 1)
k_sbp_rows(Box b, int top, int nk, int ntx, const SbpTables* __restrict__ sbp, const double* __restrict__ upin,
           const double* __restrict__ u, const double* __restrict__ um, const double* __restrict__ mu,
           const double* __restrict__ la, const double* __restrict__ rho, SgCoef sg, double* __restrict__ out,
           double cof, double dt2i, double cupd, double beta, bool dma, const double* __restrict__ acc)
{
   constexpr int NU = (3 * PL + NT - 1) / NT;
   constexpr int NM = (PL + NT - 1) / NT;
   // destinations of 16-byte asynchronous loads: every staged plane starts
   // on a 16-byte boundary
   static_assert((PL * sizeof(double)) % 16 == 0, "staged planes must stay 16-byte aligned");
   __shared__ __align__(16) double su[9 * 3 * PL];  // planes q = 0..8, components
   __shared__ __align__(16) double sm[8 * PL];
   __shared__ __align__(16) double sl[8 * PL];
   __shared__ double ssx[HX], ssy[HY];
   const int tix = blockIdx.x % ntx, tjy = blockIdx.x / ntx;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const int kbase = top ? 1 : nk;
   const int sgn = top ? 1 : -1;
   const int o = (ty + 2) * HX + tx + 2;
   // stage: displacement (or acceleration) planes q = 0..8, material 1..8.
   // The predictor's raw planes go by asynchronous global-to-LDS loads when
   // the rows allow it (dma: even ni, aligned arrays), all issued before one
   // wait; pairs off the box load the plane's first point (read only by
   // inactive points, as every staged value off the box).
   const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
   // (the corrector's planes too when the acceleration is stored: acc)
   const bool async = dma && (!CORR || acc);
   const double* __restrict__ src = CORR ? acc : u;
   if (async) {
      constexpr int PP = 3 * PL / 2, NP = (PP + NT - 1) / NT, MP = PL / 2;
      long goff[NP], moff = 0;
#pragma unroll
      for (int p = 0; p < NP; p++) {
         const int e = 2 * (tid + p * NT);
         const int c = e / PL, r = e - c * PL, y = r / HX, x = r - y * HX;
         const int gi = ia + x, gj = ja + y;
         goff[p] = e < 3 * PL && gi <= b.ie && gj <= b.je ? (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
      }
      {
         const int e = 2 * tid, y = e / HX, x = e - y * HX;
         const int gi = ia + x, gj = ja + y;
         moff = e < PL && gi <= b.ie && gj <= b.je ? (gi - b.ib) + ni * (gj - b.jb) : 0;
      }
#pragma unroll 1
      for (int q = 0; q <= 8; q++) {
         const long pb = nij * (kbase + sgn * (q - 1) - b.kb);
         double* dst = su + q * 3 * PL + 2 * 64 * wv;
#pragma unroll
         for (int p = 0; p < NP; p++)
            if (tid + p * NT < PP)
               __builtin_amdgcn_global_load_lds(src + pb + goff[p], SW4_LDS(dst + 2 * p * NT), 16, 0, 0);
         if (q >= 1 && tid < MP) {
            __builtin_amdgcn_global_load_lds(mu + pb + moff, SW4_LDS(sm + (q - 1) * PL + 2 * 64 * wv), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(la + pb + moff, SW4_LDS(sl + (q - 1) * PL + 2 * 64 * wv), 16, 0, 0);
         }
      }
   }
#pragma unroll 1
   for (int q = 0; q <= 8 && !async; q++) {
      const long pb = nij * (kbase + sgn * (q - 1) - b.kb);
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const int e = tid + p * NT;
         const int c = e / PL, r = e - c * PL, y = r / HX, x = r - y * HX;
         const int gi = ia + x, gj = ja + y;
         const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
         const long g = ok ? pb + (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
         const double vu = u[g];
         double v = vu;
         if (CORR) {
            const double vp = upin[g], vq = um[g];
            v = dt2i * (vp - 2 * vu + vq);
         }
         if (e < 3 * PL) su[(q * 3 + c) * PL + r] = ok ? v : 0.0;
      }
      if (q >= 1) {
#pragma unroll
         for (int p = 0; p < NM; p++) {
            const int e = tid + p * NT;
            const int y = e / HX, x = e - y * HX;
            const int gi = ia + x, gj = ja + y;
            const bool ok = e < PL && gi <= b.ie && gj <= b.je;
            const long g = ok ? pb + (gi - b.ib) + ni * (gj - b.jb) : 0;
            const double vm = mu[g], vl = la[g];
            if (e < PL) {
               sm[(q - 1) * PL + e] = ok ? vm : 0.0;
               sl[(q - 1) * PL + e] = ok ? vl : 0.0;
            }
         }
      }
   }
   if (tid < HX) ssx[tid] = ia + tid <= b.ie ? sg.strx[ia + tid - b.ib] : 1.0;
   if (tid >= 64 && tid < 64 + HY) ssy[tid - 64] = ja + tid - 64 <= b.je ? sg.stry[ja + tid - 64 - b.jb] : 1.0;
   __syncthreads();
   if (!active) return;
   SlabSten st{su + o, sm + o, sl + o, ssx + tx + 2, ssy + ty + 2};
   const long colp = (i - b.ib) + ni * (j - b.jb);
   // every term of the damping carries a damping coefficient at the point or
   // a neighbour along its axis: outside the supergrid layers it subtracts
   // exactly zero and is not evaluated
   const int ii = i - b.ib, jj = j - b.jb;
   const bool dxy = DAMP && beta != 0 &&
                    (sg.dcx[ii - 1] != 0 || sg.dcx[ii] != 0 || sg.dcx[ii + 1] != 0 || sg.dcy[jj - 1] != 0 ||
                     sg.dcy[jj] != 0 || sg.dcy[jj + 1] != 0);
#pragma unroll 1
   for (int kr = 1; kr <= 6; kr++) {
      const int k = kbase + sgn * (kr - 1);
      double r[3];
      sbp_row_body(st, kr, (double)sgn, sbp, r);
      const long p = colp + nij * (k - b.kb);
      const double sc = cupd / rho[p];
      double d[3] = {0, 0, 0};
      const int kk = k - b.kb;
      if (DAMP && beta != 0 && (dxy || sg.dcz[kk - 1] != 0 || sg.dcz[kk] != 0 || sg.dcz[kk + 1] != 0))
         addsgd4_point<1>(b, i, j, k, u, um, rho, sg, beta, d);
#pragma unroll
      for (int c = 0; c < 3; c++) {
         const long q = p + npts * c;
         if (CORR)
            out[q] = (upin[q] + sc * (cof * r[c])) - d[c];
         else
            out[q] = 2 * st.u(c, 0, 0, kr) - um[q] + sc * (cof * r[c]);
      }
   }
}

// The damping of the corrector on the six SBP rows of a side, after
// k_sbp_rows<true, false> wrote them: up -= damping(u, um) where a damping
// coefficient reaches the point (every term carries one; elsewhere the
// term is an exact zero).  The same arithmetic as the damping inside
// k_sbp_rows<true> (addsgd4_point, then the subtraction), so the same bits.
__global__ void __launch_bounds__(256)
k_rows_damp(Box b, int top, int nk, double* __restrict__ up, const double* __restrict__ u,
            const double* __restrict__ um, const double* __restrict__ rho, SgCoef sg, double beta)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const int k = top ? 1 + (int)blockIdx.z : nk - (int)blockIdx.z;
   const int ii = i - b.ib, jj = j - b.jb, kk = k - b.kb;
   const bool damped = sg.dcx[ii - 1] != 0 || sg.dcx[ii] != 0 || sg.dcx[ii + 1] != 0 || sg.dcy[jj - 1] != 0 ||
                       sg.dcy[jj] != 0 || sg.dcy[jj + 1] != 0 || sg.dcz[kk - 1] != 0 || sg.dcz[kk] != 0 ||
                       sg.dcz[kk + 1] != 0;
   if (!damped) return;
   double d[3] = {0, 0, 0};
   if (beta != 0) addsgd4_point<1>(b, i, j, k, u, um, rho, sg, beta, d);
   const long p = sidx(b, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      up[q] = up[q] - d[c];
   }
}

int launch_rows_damp(const Box& b, int side, int nk, double* up, const double* u, const double* um,
                     const double* rho, const SgCoef& sg, double beta, hipStream_t s)
{
   if (beta == 0) return OK;
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   k_rows_damp<<<dim3(cdiv(nx, 64), cdiv(ny, 4), 6), dim3(64, 4, 1), 0, s>>>(b, side == 4, nk, up, u, um, rho, sg,
                                                                             beta);
   return check_hip(hipGetLastError(), "SBP rows damping launch");
}

int launch_sbp_rows(bool corr, const Box& b, int side, int nk, const SbpTables* d_sbp, const double* upin,
                    const double* u, const double* um, const double* mu, const double* la, const double* rho,
                    const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s, const double* acc)
{
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   const int ntx = (int)cdiv(nx, TX), nty = (int)cdiv(ny, TY);
   const double cof = 1.0 / (h * h), dt2i = 1.0 / (dt * dt);
   if (corr && beta == 0)
      k_sbp_rows<true, false><<<ntx * nty, NT, 0, s>>>(b, side == 4, nk, ntx, d_sbp, upin, u, um, mu, la, rho, sg,
                                                       out, cof, dt2i, dt * dt * dt * dt / 12, 0.0,
                                                       acc && dma_fits(b, acc, mu, la), acc);
   else if (corr)
      k_sbp_rows<true><<<ntx * nty, NT, 0, s>>>(b, side == 4, nk, ntx, d_sbp, upin, u, um, mu, la, rho, sg, out, cof,
                                                dt2i, dt * dt * dt * dt / 12, beta, acc && dma_fits(b, acc, mu, la), acc);
   else
      k_sbp_rows<false><<<ntx * nty, NT, 0, s>>>(b, side == 4, nk, ntx, d_sbp, nullptr, u, um, mu, la, rho, sg, out,
                                                 cof, dt2i, dt * dt, beta, dma_fits(b, u, mu, la), nullptr);
   return check_hip(hipGetLastError(), "SBP rows launch");
}

int launch_sweep_lds(bool corr, int variant, const Box& b, int k1, int k2, int kc, const double* upin,
                     const double* u, const double* um, const double* mu, const double* la, const double* rho,
                     const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s, double* acc)
{
   if (k2 < k1) return OK;
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   if (kc <= 0) kc = 32;
   TileMap map;
   map.ntx = (int)cdiv(nx, TX);
   map.nty = (int)cdiv(ny, TY);
   map.ntiles = map.ntx * map.nty * (int)cdiv(k2 - k1 + 1, kc);
   const unsigned grd = 8 * cdiv(map.ntiles, 8);
   const double cof = 1.0 / (h * h), dt2i = 1.0 / (dt * dt);
   const double c4 = dt * dt * dt * dt / 12, c2 = dt * dt;
#define SW4_SWEEP(C, P, W, D)                                                                                    \
   k_sweep_lds<C, P, W, D><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, C ? upin : nullptr, u, um, mu, la, rho, sg, out, \
                                              cof, dt2i, C ? c4 : c2, beta)
   switch (variant) {
   case 2:  // prefetch, 2 waves/SIMD
      if (corr) SW4_SWEEP(true, true, 2, false); else SW4_SWEEP(false, true, 2, false);
      break;
   case 3:  // prefetch, 1 wave/SIMD (no register cap)
      if (corr) SW4_SWEEP(true, true, 1, false); else SW4_SWEEP(false, true, 1, false);
      break;
   case 4:  // as 3, the corrector with the supergrid damping in the same sweep
      if (corr) SW4_SWEEP(true, true, 1, true); else SW4_SWEEP(false, true, 1, false);
      break;
   case 5:  // corrector as 4; predictor without prefetch at 2 waves/SIMD
      if (corr) SW4_SWEEP(true, true, 1, true); else SW4_SWEEP(false, false, 2, false);
      break;
   case 6:  // both sweeps without prefetch at 2 waves/SIMD; damping by launch_damp_lds
      if (corr) SW4_SWEEP(true, false, 2, false); else SW4_SWEEP(false, false, 2, false);
      break;
   case 7:  // as 6, the predictor's planes by asynchronous global-to-LDS loads where they fit
   case 8:  // as 7; with acc, the predictor also stores the acceleration and the corrector stages it
   case 9:   // as 8 (the damping sweep runs only where the supergrid profiles are nonzero)
   case 10:  // as 9, rows of odd length by four-byte asynchronous loads (acceleration stored there too)
      if (variant >= 8 && acc) {
         const bool fit = dma_fits(b, acc, mu, la) && dma_fits(b, corr ? upin : u, mu, la);
         if (!fit && variant < 10)
            return set_error(E_INVAL, "acceleration sweeps need an even row length and aligned arrays");
         if (!fit && corr)
            k_sweep_dma_dw<DMA_CORR><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, acc, upin, mu, la, rho, sg, out, nullptr,
                                                        cof, c4, dt2i);
         else if (!fit)
            k_sweep_dma_dw<DMA_PRED><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, u, um, mu, la, rho, sg, out, acc, cof,
                                                        c2, dt2i);
         else if (corr)
            k_sweep_dma<DMA_CORR><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, acc, upin, mu, la, rho, sg, out, nullptr,
                                                     cof, c4, dt2i);
         else
            k_sweep_dma<DMA_PRED><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, u, um, mu, la, rho, sg, out, acc, cof, c2,
                                                     dt2i);
      } else if (corr)
         SW4_SWEEP(true, false, 2, false);
      else if (dma_fits(b, u, mu, la))
         k_sweep_dma<DMA_PRED><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, u, um, mu, la, rho, sg, out, nullptr, cof, c2,
                                                  dt2i);
      else if (variant >= 10)
         k_sweep_dma_dw<DMA_PRED><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, u, um, mu, la, rho, sg, out, nullptr, cof,
                                                     c2, dt2i);
      else
         SW4_SWEEP(false, false, 2, false);
      break;
   default:  // 1: no prefetch, 2 waves/SIMD
      if (corr) SW4_SWEEP(true, false, 2, false); else SW4_SWEEP(false, false, 2, false);
      break;
   }
#undef SW4_SWEEP
   return check_hip(hipGetLastError(), corr ? "LDS corrector launch" : "LDS predictor launch");
}

int launch_sweep_lds_T(const Box& b, int k1, int k2, int kc, const double* w, const double* mu, const double* la,
                       const SgCoef& sg, double* out, double h, hipStream_t s)
{
   if (k2 < k1) return OK;
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   if (kc <= 0) kc = 32;
   TileMap map;
   map.ntx = (int)cdiv(nx, TX);
   map.nty = (int)cdiv(ny, TY);
   map.ntiles = map.ntx * map.nty * (int)cdiv(k2 - k1 + 1, kc);
   const unsigned grd = 8 * cdiv(map.ntiles, 8);
   // rows of odd length by four-byte asynchronous loads (the same results)
   if (dma_fits(b, w, mu, la))
      k_sweep_dma<DMA_TR><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, w, nullptr, mu, la, nullptr, sg, out, nullptr,
                                             1.0 / (h * h), 1.0, 1.0);
   else
      k_sweep_dma_dw<DMA_TR><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, w, nullptr, mu, la, nullptr, sg, out, nullptr,
                                                1.0 / (h * h), 1.0, 1.0);
   return check_hip(hipGetLastError(), "transposed sweep launch");
}

}  // namespace sw4amd
