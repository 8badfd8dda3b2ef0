// kernels_comp.hip -- the FP32 form of the fused step with a compensated time
// update (state and formulas in kernels_comp.h).
//
// The two sweeps have the tiling of the FP64 sweeps (kernels_fast.hip): a
// workgroup owns a 32 x 8 tile of (i,j) columns and marches along k, the five
// planes of the staged field sit in an LDS ring and the column values of mu,
// lambda and the stretching rotate in registers.  In FP32 the corrector sweep
// also carries the supergrid damping: D = u^n - u^{n-1} is a stored field, so
// its planes are staged in a second LDS ring next to the acceleration.
//   predictor:  reads U, mu, lambda, rho          writes A          (9 words)
//   corrector:  reads A, D, U, C, mu, lambda, rho writes Dn, U, C   (24 words)
#include "kernels_comp.h"
#include "kernels_step.h"
#include "sbp_row_body.h"
#include "stencil_body.h"

namespace sw4amd {

namespace {
constexpr int TX = 32, TY = 8, HX = TX + 4, HY = TY + 4, NT = TX * TY, PL = HX * HY;
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
}  // namespace

// u + d with the low part carried in C: (U, C) <- U + C + d, error-free
// two-sum of U and d + C (Knuth), so u = U + C keeps ~48 bits.
__device__ __forceinline__ void comp_add(float& U, float& C, float d)
{
   const float y = d + C;
   const float s = U + y;
   const float bb = s - U;
   const float e = (U - (s - bb)) + (y - bb);
   U = s;
   C = e;
}

struct CompTiles {
   int ntx, nty, ntiles;
};

struct CompSten {
   const float* p0;  // ring slot of plane k-2 at this thread's point
   const float* p1;
   const float* p2;
   const float* p3;
   const float* p4;
   const float* pm;
   const float* pl;
   const float* psx;
   const float* psy;
   float muc[5], lac[5], szv[5];
   __device__ __forceinline__ float u(int c, int di, int dj, int dk) const
   {
      const float* p = dk == -2 ? p0 : dk == -1 ? p1 : dk == 0 ? p2 : dk == 1 ? p3 : p4;
      return p[c * PL + dj * HX + di];
   }
   __device__ __forceinline__ float mu(int di, int dj, int dk) const { return dk ? muc[dk + 2] : pm[dj * HX + di]; }
   __device__ __forceinline__ float la(int di, int dj, int dk) const { return dk ? lac[dk + 2] : pl[dj * HX + di]; }
   __device__ __forceinline__ float sx(int d) const { return psx[d]; }
   __device__ __forceinline__ float sy(int d) const { return psy[d]; }
   __device__ __forceinline__ float sz(int d) const { return szv[d + 2]; }
};

// CORR = false: A = cof L(U) / rho.
// CORR = true:  Dn = (D + dt2 A) + c4/rho cof L(A) - damping(D);  (U, C) += Dn.
// D of the planes k..k+2 is kept in an LDS ring for the in-plane damping
// stencil, the thread's own column of D for k-2..k+2 and rho at k-1, k+1 in
// registers (ew-cfromfort.C:985-1041 on D = U - Um).
template <bool CORR>
__global__ void __launch_bounds__(NT, 2)
k_comp_sweep(Box b, CompTiles map, int kbeg, int kend, int kc, const float* __restrict__ stg,
             const float* __restrict__ mu, const float* __restrict__ la, const float* __restrict__ rho, SgCoefF sg,
             const float* __restrict__ dold, float* __restrict__ dnew, float* __restrict__ uu,
             float* __restrict__ cc, float* __restrict__ aout, float cof, float dt2, float c4, float beta)
{
   constexpr int NU = (3 * PL + NT - 1) / NT;
   constexpr int NM = (PL + NT - 1) / NT;
   __shared__ float su[5 * 3 * PL];  // ring: plane kp in slot (kp - k0 + 2) % 5
   __shared__ float sm[PL], sl[PL];
   __shared__ float sd[CORR ? 3 * 3 * PL : 1];  // D, plane kp in slot (kp - k0) % 3
   __shared__ float sr[CORR ? PL : 1];          // rho of plane k
   __shared__ float ssx[HX], ssy[HY];
   __shared__ float sgx[CORR ? 5 * TX : 1], sgy[CORR ? 5 * TY : 1];
   // XCD-aware tile order (as k_sweep_lds): id -> (xcd = id % 8, slot = id / 8)
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const int k0 = kbeg + tkz * kc;
   const int k1 = min(kend, k0 + kc - 1);
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;

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
   float ru[NU], rd[CORR ? NU : 1], rmu[NM], rla[NM], rrh[CORR ? NM : 1];
   // out-of-tile elements have offset 0: loads are unconditional, then masked
   auto load_plane = [&](int kp) {
      const long pb = nij * (kp - b.kb);
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const float v = stg[pb + goff[p]];
         ru[p] = gok[p] ? v : 0.0f;
         if (CORR) {
            const float w = dold[pb + goff[p]];
            rd[p] = gok[p] ? w : 0.0f;
         }
      }
   };
   auto store_plane = [&](int slot, int dslot) {
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const int e = tid + p * NT;
         if (e < 3 * PL) {
            su[slot * 3 * PL + e] = ru[p];
            if (CORR) sd[dslot * 3 * PL + e] = rd[p];
         }
      }
   };
   auto load_mat = [&](int kp) {
      const long pb = nij * (kp - b.kb);
#pragma unroll
      for (int p = 0; p < NM; p++) {
         const float vm = mu[pb + moff[p]], vl = la[pb + moff[p]];
         rmu[p] = mok[p] ? vm : 0.0f;
         rla[p] = mok[p] ? vl : 0.0f;
         if (CORR) {
            const float vr = rho[pb + moff[p]];
            rrh[p] = mok[p] ? vr : 0.0f;
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
            if (CORR) sr[e] = rrh[p];
         }
      }
   };

   CompSten s;
   float dcol[5][3], rkm = 0, rkp = 0, rk0 = 0;
   if (CORR) {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const float v = dold[col + nij * (k0 - 2 + q - b.kb) + npts * c];
            dcol[q][c] = active ? v : 0.0f;
         }
      const float v = rho[col + nij * (k0 - 1 - b.kb)];
      rkm = active ? v : 0.0f;
      if (tid < TX) {
         const int ii = ia + 2 + tid - b.ib;
         const bool ok = ia + 2 + tid <= b.ie - 2;
         sgx[0 * TX + tid] = ok ? sg.dcx[ii - 1] : 0.0f;
         sgx[1 * TX + tid] = ok ? sg.dcx[ii] : 0.0f;
         sgx[2 * TX + tid] = ok ? sg.dcx[ii + 1] : 0.0f;
         sgx[3 * TX + tid] = ok ? sg.strx[ii] : 0.0f;
         sgx[4 * TX + tid] = ok ? sg.cox[ii] : 0.0f;
      } else if (tid >= 64 && tid < 64 + TY) {
         const int q = tid - 64, jj = ja + 2 + q - b.jb;
         const bool ok = ja + 2 + q <= b.je - 2;
         sgy[0 * TY + q] = ok ? sg.dcy[jj - 1] : 0.0f;
         sgy[1 * TY + q] = ok ? sg.dcy[jj] : 0.0f;
         sgy[2 * TY + q] = ok ? sg.dcy[jj + 1] : 0.0f;
         sgy[3 * TY + q] = ok ? sg.stry[jj] : 0.0f;
         sgy[4 * TY + q] = ok ? sg.coy[jj] : 0.0f;
      }
   }
   if (tid >= 128 && tid < 128 + HX) ssx[tid - 128] = ia + tid - 128 <= b.ie ? sg.strx[ia + tid - 128 - b.ib] : 1.0f;
   if (tid >= 192 && tid < 192 + HY) ssy[tid - 192] = ja + tid - 192 <= b.je ? sg.stry[ja + tid - 192 - b.jb] : 1.0f;
#pragma unroll
   for (int q = 0; q < 4; q++) {
      load_plane(k0 - 2 + q);
      store_plane(q, q >= 2 ? q - 2 : 2);  // D ring: planes k0 (slot 0), k0+1 (slot 1)
      const long pq = nij * (k0 - 2 + q - b.kb);
      const float vm = mu[col + pq], vl = la[col + pq];
      s.muc[q] = active ? vm : 0.0f;
      s.lac[q] = active ? vl : 0.0f;
      s.szv[q] = sg.strz[k0 - 2 + q - b.kb];
   }
   s.pm = sm + o;
   s.pl = sl + o;
   s.psx = ssx + tx + 2;
   s.psy = ssy + ty + 2;

   for (int k = k0; k <= k1; k++) {
      const int sl0 = (k - k0) % 5;
      const long pk = col + nij * (k - b.kb);
      // every global load of the iteration, then the LDS stores and a barrier
      load_plane(k + 2);
      load_mat(k);
      const float cmu = mu[col + nij * (k + 2 - b.kb)], cla = la[col + nij * (k + 2 - b.kb)];
      const float csz = sg.strz[k + 2 - b.kb];
      float cu[3] = {0, 0, 0}, cl[3] = {0, 0, 0}, crho = 1.0f;
      if (CORR) {
#pragma unroll
         for (int c = 0; c < 3; c++) {
            cu[c] = uu[pk + npts * c];
            cl[c] = cc[pk + npts * c];
         }
         rkp = rho[pk + nij];
      } else {
         const float v = rho[pk];
         crho = active ? v : 1.0f;
      }
      store_plane((sl0 + 4) % 5, (k + 2 - k0) % 3);
      store_mat();
      s.muc[4] = active ? cmu : 0.0f;
      s.lac[4] = active ? cla : 0.0f;
      s.szv[4] = csz;
      const int kz = k - b.kb;
      float gz[5] = {0, 0, 0, 0, 0};
      if (CORR) {
         gz[0] = sg.dcz[kz - 1]; gz[1] = sg.dcz[kz]; gz[2] = sg.dcz[kz + 1]; gz[3] = sg.strz[kz]; gz[4] = sg.coz[kz];
      }
      __syncthreads();
      if (active) {
         s.p0 = su + (sl0 * 3) * PL + o;
         s.p1 = su + (((sl0 + 1) % 5) * 3) * PL + o;
         s.p2 = su + (((sl0 + 2) % 5) * 3) * PL + o;
         s.p3 = su + (((sl0 + 3) % 5) * 3) * PL + o;
         s.p4 = su + (((sl0 + 4) % 5) * 3) * PL + o;
         float r[3];
         rhs4_interior_body(s, r);
         if (!CORR) {
            const float sc = cof / crho;
#pragma unroll
            for (int c = 0; c < 3; c++) aout[pk + npts * c] = sc * r[c];
         } else {
            const float r0 = sr[o];
            rk0 = r0;
            const float* dpl = sd + ((k - k0) % 3) * 3 * PL + o;  // D of plane k at this point
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[4][c] = sd[((k + 2 - k0) % 3) * 3 * PL + c * PL + o];
            const float birho = beta / r0;
            const float* gx = sgx + tx;
            const float* gy = sgy + ty;
            const float rxp = sr[o + 1] * gx[2 * TX], rx0 = r0 * gx[TX], rxm = sr[o - 1] * gx[0];
            const float ryp = sr[o + HX] * gy[2 * TY], ry0 = r0 * gy[TY], rym = sr[o - HX] * gy[0];
            const float rzp = rkp * gz[2], rz0 = r0 * gz[1], rzm = rkm * gz[0];
            const float wx = gx[3 * TX] * gy[4 * TY] * gz[4];
            const float wy = gy[3 * TY] * gx[4 * TX] * gz[4];
            const float wz = gz[3] * gx[4 * TX] * gy[4 * TY];
            const float sc = c4 / r0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
               const float* q = dpl + c * PL;
               const float ax = rxp * (q[2] - 2 * q[1] + q[0]) - 2 * rx0 * (q[1] - 2 * q[0] + q[-1]) +
                                rxm * (q[0] - 2 * q[-1] + q[-2]);
               const float ay = ryp * (q[2 * HX] - 2 * q[HX] + q[0]) - 2 * ry0 * (q[HX] - 2 * q[0] + q[-HX]) +
                                rym * (q[0] - 2 * q[-HX] + q[-2 * HX]);
               const float az = rzp * (dcol[4][c] - 2 * dcol[3][c] + dcol[2][c]) -
                                2 * rz0 * (dcol[3][c] - 2 * dcol[2][c] + dcol[1][c]) +
                                rzm * (dcol[2][c] - 2 * dcol[1][c] + dcol[0][c]);
               const float dmp = birho * (wx * ax + wy * ay + wz * az);
               const float dn = ((q[0] + dt2 * s.p2[c * PL]) + sc * (cof * r[c])) - dmp;
               float U = cu[c], C = cl[c];
               comp_add(U, C, dn);
               const long g = pk + npts * c;
               dnew[g] = dn;
               uu[g] = U;
               cc[g] = C;
            }
         }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++) {
         s.muc[q] = s.muc[q + 1];
         s.lac[q] = s.lac[q + 1];
         s.szv[q] = s.szv[q + 1];
      }
      if (CORR) {
#pragma unroll
         for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[q][c] = dcol[q + 1][c];
         rkm = rk0;
      }
   }
}

// The two sweeps above with their staged planes brought to LDS by
// asynchronous global-to-LDS loads (global_load_lds_dwordx4: lane l of a
// wave moves four floats to LDS at M0 + 16 l), as the FP64 sweeps of
// kernels_fast.hip: the planes of k+3 and the material of plane k+1 land
// while plane k is computed, one barrier per plane; the centre values of U,
// C (or rho) are issued after the barrier and used after the stencil, the
// column values one plane ahead.  Rings: the staged field in six slots
// (planes k-2..k+3), D in four (k..k+3), mu, lambda (and rho) in two.  Same
// arithmetic in the same order as k_comp_sweep.  Needs ni % 4 == 0 and
// 16-byte aligned arrays (every quad of floats in one row, aligned).  The
// predictor runs at 4 waves/SIMD (128 registers, 2 of them spilled: faster
// than 3 waves without the spill).
#define SW4_LDSF(p) ((__attribute__((address_space(3))) void*)(p))
template <bool CORR>
__global__ void __launch_bounds__(NT, CORR ? 2 : 4)
k_comp_dma(Box b, CompTiles map, int kbeg, int kend, int kc, const float* __restrict__ stg,
           const float* __restrict__ mu, const float* __restrict__ la, const float* __restrict__ rho, SgCoefF sg,
           const float* __restrict__ dold, float* __restrict__ dnew, float* __restrict__ uu,
           float* __restrict__ cc, float* __restrict__ aout, float cof, float dt2, float c4, float beta)
{
   constexpr int QP = 3 * PL / 4;          // quads of a staged 3-component plane
   constexpr int NQ = (QP + NT - 1) / NT;  // asynchronous loads per thread and plane
   constexpr int QM = PL / 4;              // quads of a scalar plane
   static_assert(HX % 4 == 0 && QM <= NT, "quads must tile the rows; one material load per thread");
   __shared__ __align__(16) float su[6 * 3 * PL];              // plane kp in slot (kp - k0 + 2) % 6
   __shared__ __align__(16) float sd[CORR ? 4 * 3 * PL : 4];   // D of plane kp in slot (kp - k0) % 4
   __shared__ __align__(16) float sm[2 * PL], sl[2 * PL];      // plane kp in half (kp - k0) & 1
   __shared__ __align__(16) float sr[CORR ? 2 * PL : 4];
   __shared__ float ssx[HX], ssy[HY];
   __shared__ float sgx[CORR ? 5 * TX : 1], sgy[CORR ? 5 * TY : 1];
   const int per = (int)((map.ntiles + 7) / 8);
   const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
   if (t >= map.ntiles) return;
   const int tix = t % map.ntx, tr = t / map.ntx, tjy = tr % map.nty, tkz = tr / map.nty;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const int k0 = kbeg + tkz * kc;
   const int k1 = min(kend, k0 + kc - 1);
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const long col = active ? (i - b.ib) + ni * (j - b.jb) : 0;
   const int o = (ty + 2) * HX + tx + 2;

   // quad m = tid + p NT of the staged (c, jj, ii) tile: offset of its first float
   long goff[NQ];
#pragma unroll
   for (int p = 0; p < NQ; p++) {
      const int e = 4 * (tid + p * NT);
      const int c = e / PL, r = e - c * PL, jj = r / HX, ii = r - jj * HX;
      const int gi = ia + ii, gj = ja + jj;
      const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
      goff[p] = ok ? (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0;
   }
   long moff;
   {
      const int e = 4 * tid, jj = e / HX, ii = e - jj * HX;
      const int gi = ia + ii, gj = ja + jj;
      moff = e < PL && gi <= b.ie && gj <= b.je ? (gi - b.ib) + ni * (gj - b.jb) : 0;
   }
   auto dma3 = [&](const float* src, float* ring, int kp, int slot) {
      const float* g = src + nij * (kp - b.kb);
      float* dst = ring + slot * 3 * PL + 4 * 64 * wv;
#pragma unroll
      for (int p = 0; p < NQ; p++)
         if (tid + p * NT < QP) __builtin_amdgcn_global_load_lds(g + goff[p], SW4_LDSF(dst + 4 * p * NT), 16, 0, 0);
   };
   auto dma_mat = [&](int kp, int half) {
      if (tid < QM) {
         const long pb = nij * (kp - b.kb);
         __builtin_amdgcn_global_load_lds(mu + pb + moff, SW4_LDSF(sm + half * PL + 4 * 64 * wv), 16, 0, 0);
         __builtin_amdgcn_global_load_lds(la + pb + moff, SW4_LDSF(sl + half * PL + 4 * 64 * wv), 16, 0, 0);
         if (CORR)
            __builtin_amdgcn_global_load_lds(rho + pb + moff, SW4_LDSF(sr + half * PL + 4 * 64 * wv), 16, 0, 0);
      }
   };

   CompSten s;
   float dcol[5][3], rkm = 0, rk0 = 0, nrkp = 0;
   if (CORR) {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const float v = dold[col + nij * (k0 - 2 + q - b.kb) + npts * c];
            dcol[q][c] = active ? v : 0.0f;
         }
      const float v = rho[col + nij * (k0 - 1 - b.kb)];
      rkm = active ? v : 0.0f;
      nrkp = rho[col + nij * (k0 + 1 - b.kb)];
      if (tid < TX) {
         const int ii = ia + 2 + tid - b.ib;
         const bool ok = ia + 2 + tid <= b.ie - 2;
         sgx[0 * TX + tid] = ok ? sg.dcx[ii - 1] : 0.0f;
         sgx[1 * TX + tid] = ok ? sg.dcx[ii] : 0.0f;
         sgx[2 * TX + tid] = ok ? sg.dcx[ii + 1] : 0.0f;
         sgx[3 * TX + tid] = ok ? sg.strx[ii] : 0.0f;
         sgx[4 * TX + tid] = ok ? sg.cox[ii] : 0.0f;
      } else if (tid >= 64 && tid < 64 + TY) {
         const int q = tid - 64, jj = ja + 2 + q - b.jb;
         const bool ok = ja + 2 + q <= b.je - 2;
         sgy[0 * TY + q] = ok ? sg.dcy[jj - 1] : 0.0f;
         sgy[1 * TY + q] = ok ? sg.dcy[jj] : 0.0f;
         sgy[2 * TY + q] = ok ? sg.dcy[jj + 1] : 0.0f;
         sgy[3 * TY + q] = ok ? sg.stry[jj] : 0.0f;
         sgy[4 * TY + q] = ok ? sg.coy[jj] : 0.0f;
      }
   }
   if (tid >= 128 && tid < 128 + HX) ssx[tid - 128] = ia + tid - 128 <= b.ie ? sg.strx[ia + tid - 128 - b.ib] : 1.0f;
   if (tid >= 192 && tid < 192 + HY) ssy[tid - 192] = ja + tid - 192 <= b.je ? sg.stry[ja + tid - 192 - b.jb] : 1.0f;
#pragma unroll
   for (int q = 0; q < 5; q++) dma3(stg, su, k0 - 2 + q, q);
   if (CORR) {
#pragma unroll
      for (int q = 0; q < 3; q++) dma3(dold, sd, k0 + q, q);
   }
   dma_mat(k0, 0);
#pragma unroll
   for (int q = 0; q < 4; q++) {
      const long pq = nij * (k0 - 2 + q - b.kb);
      const float vm = mu[col + pq], vl = la[col + pq];
      s.muc[q] = active ? vm : 0.0f;
      s.lac[q] = active ? vl : 0.0f;
      s.szv[q] = sg.strz[k0 - 2 + q - b.kb];
   }
   s.psx = ssx + tx + 2;
   s.psy = ssy + ty + 2;
   float cmu = mu[col + nij * (k0 + 2 - b.kb)], cla = la[col + nij * (k0 + 2 - b.kb)];
   float csz = sg.strz[k0 + 2 - b.kb];

   for (int k = k0; k <= k1; k++) {
      const int sl0 = (k - k0) % 6, h = (k - k0) & 1;
      const long pk = col + nij * (k - b.kb);
      s.muc[4] = active ? cmu : 0.0f;
      s.lac[4] = active ? cla : 0.0f;
      s.szv[4] = csz;
      const float rkp = nrkp;
      __syncthreads();  // plane k+2 (field, D) and the material of plane k have landed; plane k-1 is done
      // centre values, issued first and used after the stencil
      float cu[3] = {0, 0, 0}, cl[3] = {0, 0, 0}, crho = 1.0f;
      if (CORR) {
#pragma unroll
         for (int c = 0; c < 3; c++) {
            cu[c] = uu[pk + npts * c];
            cl[c] = cc[pk + npts * c];
         }
      } else {
         const float v = rho[pk];
         crho = active ? v : 1.0f;
      }
      if (k < k1) {
         dma3(stg, su, k + 3, (sl0 + 5) % 6);
         if (CORR) dma3(dold, sd, k + 3, (k + 3 - k0) % 4);
         dma_mat(k + 1, h ^ 1);
         cmu = mu[col + nij * (k + 3 - b.kb)];  // col = 0 where inactive: a valid address, value unused
         cla = la[col + nij * (k + 3 - b.kb)];
         csz = sg.strz[k + 3 - b.kb];
         if (CORR) nrkp = rho[col + nij * (k + 2 - b.kb)];
      }
      const int kz = k - b.kb;
      float gz[5] = {0, 0, 0, 0, 0};
      if (CORR) {
         gz[0] = sg.dcz[kz - 1]; gz[1] = sg.dcz[kz]; gz[2] = sg.dcz[kz + 1]; gz[3] = sg.strz[kz]; gz[4] = sg.coz[kz];
      }
      if (active) {
         s.p0 = su + (sl0 * 3) * PL + o;
         s.p1 = su + (((sl0 + 1) % 6) * 3) * PL + o;
         s.p2 = su + (((sl0 + 2) % 6) * 3) * PL + o;
         s.p3 = su + (((sl0 + 3) % 6) * 3) * PL + o;
         s.p4 = su + (((sl0 + 4) % 6) * 3) * PL + o;
         s.pm = sm + h * PL + o;
         s.pl = sl + h * PL + o;
         float r[3];
         rhs4_interior_body(s, r);
         if (!CORR) {
            const float sc = cof / crho;
#pragma unroll
            for (int c = 0; c < 3; c++) aout[pk + npts * c] = sc * r[c];
         } else {
            const float* srh = sr + h * PL;
            const float r0 = srh[o];
            rk0 = r0;
            const float* dpl = sd + ((k - k0) % 4) * 3 * PL + o;  // D of plane k at this point
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[4][c] = sd[((k + 2 - k0) % 4) * 3 * PL + c * PL + o];
            const float birho = beta / r0;
            const float* gx = sgx + tx;
            const float* gy = sgy + ty;
            const float rxp = srh[o + 1] * gx[2 * TX], rx0 = r0 * gx[TX], rxm = srh[o - 1] * gx[0];
            const float ryp = srh[o + HX] * gy[2 * TY], ry0 = r0 * gy[TY], rym = srh[o - HX] * gy[0];
            const float rzp = rkp * gz[2], rz0 = r0 * gz[1], rzm = rkm * gz[0];
            const float wx = gx[3 * TX] * gy[4 * TY] * gz[4];
            const float wy = gy[3 * TY] * gx[4 * TX] * gz[4];
            const float wz = gz[3] * gx[4 * TX] * gy[4 * TY];
            const float sc = c4 / r0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
               const float* q = dpl + c * PL;
               const float ax = rxp * (q[2] - 2 * q[1] + q[0]) - 2 * rx0 * (q[1] - 2 * q[0] + q[-1]) +
                                rxm * (q[0] - 2 * q[-1] + q[-2]);
               const float ay = ryp * (q[2 * HX] - 2 * q[HX] + q[0]) - 2 * ry0 * (q[HX] - 2 * q[0] + q[-HX]) +
                                rym * (q[0] - 2 * q[-HX] + q[-2 * HX]);
               const float az = rzp * (dcol[4][c] - 2 * dcol[3][c] + dcol[2][c]) -
                                2 * rz0 * (dcol[3][c] - 2 * dcol[2][c] + dcol[1][c]) +
                                rzm * (dcol[2][c] - 2 * dcol[1][c] + dcol[0][c]);
               const float dmp = birho * (wx * ax + wy * ay + wz * az);
               const float dn = ((q[0] + dt2 * s.p2[c * PL]) + sc * (cof * r[c])) - dmp;
               float U = cu[c], C = cl[c];
               comp_add(U, C, dn);
               const long g = pk + npts * c;
               dnew[g] = dn;
               uu[g] = U;
               cc[g] = C;
            }
         }
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
         s.muc[q] = s.muc[q + 1];
         s.lac[q] = s.lac[q + 1];
         s.szv[q] = s.szv[q + 1];
      }
      if (CORR) {
#pragma unroll
         for (int q = 0; q < 4; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) dcol[q][c] = dcol[q + 1][c];
         rkm = rk0;
      }
   }
}

// Damping increment of one point from the stored D (addsgd4_point on
// D = U - Um, ew-cfromfort.C:985-1041), FP32.
__device__ __forceinline__ void damp_point_f(const Box& b, int i, int j, int k, const float* __restrict__ d,
                                             const float* __restrict__ rho, const SgCoefF& sg, float beta,
                                             float out[3])
{
   const int ii = i - b.ib, jj = j - b.jb, kk = k - b.kb;
   const float r0 = rho[sidx(b, i, j, k)];
   const float birho = beta / r0;
   const float rxp = rho[sidx(b, i + 1, j, k)] * sg.dcx[ii + 1], rx0 = r0 * sg.dcx[ii],
               rxm = rho[sidx(b, i - 1, j, k)] * sg.dcx[ii - 1];
   const float ryp = rho[sidx(b, i, j + 1, k)] * sg.dcy[jj + 1], ry0 = r0 * sg.dcy[jj],
               rym = rho[sidx(b, i, j - 1, k)] * sg.dcy[jj - 1];
   const float rzp = rho[sidx(b, i, j, k + 1)] * sg.dcz[kk + 1], rz0 = r0 * sg.dcz[kk],
               rzm = rho[sidx(b, i, j, k - 1)] * sg.dcz[kk - 1];
   const float wx = sg.strx[ii] * sg.coy[jj] * sg.coz[kk];
   const float wy = sg.stry[jj] * sg.cox[ii] * sg.coz[kk];
   const float wz = sg.strz[kk] * sg.cox[ii] * sg.coy[jj];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      float dx[5], dy[5], dz[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
         dx[t] = d[vidx<1>(b, c, i + t - 2, j, k)];
         dy[t] = d[vidx<1>(b, c, i, j + t - 2, k)];
         dz[t] = d[vidx<1>(b, c, i, j, k + t - 2)];
      }
      const float ax = rxp * (dx[4] - 2 * dx[3] + dx[2]) - 2 * rx0 * (dx[3] - 2 * dx[2] + dx[1]) +
                       rxm * (dx[2] - 2 * dx[1] + dx[0]);
      const float ay = ryp * (dy[4] - 2 * dy[3] + dy[2]) - 2 * ry0 * (dy[3] - 2 * dy[2] + dy[1]) +
                       rym * (dy[2] - 2 * dy[1] + dy[0]);
      const float az = rzp * (dz[4] - 2 * dz[3] + dz[2]) - 2 * rz0 * (dz[3] - 2 * dz[2] + dz[1]) +
                       rzm * (dz[2] - 2 * dz[1] + dz[0]);
      out[c] = birho * (wx * ax + wy * ay + wz * az);
   }
}

// Row-space accessor of the staged slab (sbp_row_body.h), FP32.
struct SlabStenF {
   const float* ub;  // staged plane q = 0 at this thread's point
   const float* mb;  // mu plane q = 1
   const float* lb;
   const float* psx;
   const float* psy;
   __device__ __forceinline__ float u(int c, int di, int dj, int q) const { return ub[(q * 3 + c) * PL + dj * HX + di]; }
   __device__ __forceinline__ float mu(int di, int dj, int q) const { return mb[(q - 1) * PL + dj * HX + di]; }
   __device__ __forceinline__ float la(int di, int dj, int q) const { return lb[(q - 1) * PL + dj * HX + di]; }
   __device__ __forceinline__ float sx(int d) const { return psx[d]; }
   __device__ __forceinline__ float sy(int d) const { return psy[d]; }
};

// The six SBP rows at a free surface (top: k = 1..6, bottom: nk..nk-5), one
// 32 x 8 tile per workgroup with nine planes of U (predictor) or A
// (corrector) and eight material planes staged in LDS.
template <bool CORR>
__global__ void __launch_bounds__(NT, 1)
k_comp_rows(Box b, int top, int nk, int ntx, const SbpTablesT<float>* __restrict__ sbp, const float* __restrict__ stg,
            const float* __restrict__ mu, const float* __restrict__ la, const float* __restrict__ rho, SgCoefF sg,
            const float* __restrict__ dold, float* __restrict__ dnew, float* __restrict__ uu, float* __restrict__ cc,
            float* __restrict__ aout, float cof, float dt2, float c4, float beta)
{
   constexpr int NU = (3 * PL + NT - 1) / NT;
   constexpr int NM = (PL + NT - 1) / NT;
   __shared__ float su[9 * 3 * PL];
   __shared__ float sm[8 * PL], sl[8 * PL];
   __shared__ float ssx[HX], ssy[HY];
   const int tix = blockIdx.x % ntx, tjy = blockIdx.x / ntx;
   const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
   const int ia = b.ib + tix * TX, ja = b.jb + tjy * TY;
   const int i = ia + 2 + tx, j = ja + 2 + ty;
   const bool active = i <= b.ie - 2 && j <= b.je - 2;
   const long ni = b.ni, nij = b.nij, npts = b.npts;
   const int kbase = top ? 1 : nk;
   const int sgn = top ? 1 : -1;
   const int o = (ty + 2) * HX + tx + 2;
#pragma unroll 1
   for (int q = 0; q <= 8; q++) {
      const long pb = nij * (kbase + sgn * (q - 1) - b.kb);
#pragma unroll
      for (int p = 0; p < NU; p++) {
         const int e = tid + p * NT;
         const int c = e / PL, r = e - c * PL, y = r / HX, x = r - y * HX;
         const int gi = ia + x, gj = ja + y;
         const bool ok = e < 3 * PL && gi <= b.ie && gj <= b.je;
         const float v = stg[ok ? pb + (gi - b.ib) + ni * (gj - b.jb) + npts * c : 0];
         if (e < 3 * PL) su[(q * 3 + c) * PL + r] = ok ? v : 0.0f;
      }
      if (q >= 1) {
#pragma unroll
         for (int p = 0; p < NM; p++) {
            const int e = tid + p * NT;
            const int y = e / HX, x = e - y * HX;
            const int gi = ia + x, gj = ja + y;
            const bool ok = e < PL && gi <= b.ie && gj <= b.je;
            const long g = ok ? pb + (gi - b.ib) + ni * (gj - b.jb) : 0;
            const float vm = mu[g], vl = la[g];
            if (e < PL) {
               sm[(q - 1) * PL + e] = ok ? vm : 0.0f;
               sl[(q - 1) * PL + e] = ok ? vl : 0.0f;
            }
         }
      }
   }
   if (tid < HX) ssx[tid] = ia + tid <= b.ie ? sg.strx[ia + tid - b.ib] : 1.0f;
   if (tid >= 64 && tid < 64 + HY) ssy[tid - 64] = ja + tid - 64 <= b.je ? sg.stry[ja + tid - 64 - b.jb] : 1.0f;
   __syncthreads();
   if (!active) return;
   const SlabStenF st{su + o, sm + o, sl + o, ssx + tx + 2, ssy + ty + 2};
   const long colp = (i - b.ib) + ni * (j - b.jb);
#pragma unroll 1
   for (int kr = 1; kr <= 6; kr++) {
      const int k = kbase + sgn * (kr - 1);
      float r[3];
      sbp_row_body(st, kr, (float)sgn, sbp, r);
      const long p = colp + nij * (k - b.kb);
      if (!CORR) {
         const float sc = cof / rho[p];
#pragma unroll
         for (int c = 0; c < 3; c++) aout[p + npts * c] = sc * r[c];
      } else {
         float d[3] = {0, 0, 0};
         if (beta != 0) damp_point_f(b, i, j, k, dold, rho, sg, beta, d);
         const float sc = c4 / rho[p];
#pragma unroll
         for (int c = 0; c < 3; c++) {
            const long g = p + npts * c;
            const float dn = ((dold[g] + dt2 * st.u(c, 0, 0, kr)) + sc * (cof * r[c])) - d[c];
            float U = uu[g], C = cc[g];
            comp_add(U, C, dn);
            dnew[g] = dn;
            uu[g] = U;
            cc[g] = C;
         }
      }
   }
}

// Ghost shell (shell_point): the predictor has A = 0 there (L(u) and F
// vanish); the corrector carries Up: Dn = D + dt2 A, (U, C) += Dn.
template <bool CORR>
__global__ void k_comp_shell(Box b, CompState st, float dt2)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   int i, j, k;
   if (!shell_point(b, q, i, j, k)) return;
   const long p = sidx(b, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      if (!CORR) {
         st.A[v] = 0.0f;
      } else {
         const float dn = st.D[v] + dt2 * st.A[v];
         float U = st.U[v], C = st.C[v];
         comp_add(U, C, dn);
         st.Dn[v] = dn;
         st.U[v] = U;
         st.C[v] = C;
      }
   }
}

// Sources in FP64 (time functions and forces), the increment rounded once:
// predictor A += F(t)/rho (coef 1), corrector Dn += coef/rho F_tt(t) with the
// same increment added to (U, C).
template <bool CORR>
__global__ void k_comp_src(Box b, SourceTable tab, double t, double coef, CompState st)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= tab.nunique) return;
   const int s0 = tab.first[r], s1 = tab.first[r + 1];
   const int i = tab.ijk[3 * s0], j = tab.ijk[3 * s0 + 1], k = tab.ijk[3 * s0 + 2];
   double f[3] = {0, 0, 0};
   for (int s = s0; s < s1; s++) {
      const double g = time_function(tab.tfkind[s], CORR ? 1 : 0, tab.freq[s], t - tab.t0[s]);
      f[0] += tab.force[3 * s] * g;
      f[1] += tab.force[3 * s + 1] * g;
      f[2] += tab.force[3 * s + 2] * g;
   }
   const long p = sidx(b, i, j, k);
   const double sc = coef / (double)st.rho[p];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      const float inc = (float)(sc * f[c]);
      if (!CORR) {
         st.A[v] += inc;
      } else {
         float U = st.U[v], C = st.C[v];
         comp_add(U, C, inc);
         st.Dn[v] += inc;
         st.U[v] = U;
         st.C[v] = C;
      }
   }
}

// Views of the solution the boundary conditions act on, in FP64.
// Predictor: Up = (U + C) + (D + dt2 A); a ghost value g sets
// A = ((g - (U + C)) - D) / dt2.
struct PredView {
   const float* U;
   const float* C;
   const float* D;
   float* A;
   double dt2;
   __device__ __forceinline__ double operator[](long q) const
   {
      return ((double)U[q] + (double)C[q]) + ((double)D[q] + dt2 * (double)A[q]);
   }
   __device__ __forceinline__ void set(long q, double g) const
   {
      A[q] = (float)(((g - ((double)U[q] + (double)C[q])) - (double)D[q]) / dt2);
   }
};
// Corrector: u^{n+1} = U + C; a ghost value g replaces it, and Dn takes the
// same change.
struct CorrView {
   float* U;
   float* C;
   float* Dn;
   __device__ __forceinline__ double operator[](long q) const { return (double)U[q] + (double)C[q]; }
   __device__ __forceinline__ void set(long q, double g) const
   {
      const double u = (double)U[q] + (double)C[q];
      Dn[q] = (float)((double)Dn[q] + (g - u));
      const float hi = (float)g;
      U[q] = hi;
      C[q] = (float)(g - (double)hi);
   }
};

struct CompWindow {
   int i0, i1, j0, j1, k0, k1;
};

template <class V>
__global__ void k_cbc_fill(Box b, CompWindow w, V v, const double* __restrict__ bf)
{
   const long wi = w.i1 - w.i0 + 1, wj = w.j1 - w.j0 + 1, wk = w.k1 - w.k0 + 1;
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= wi * wj * wk) return;
   const int i = w.i0 + (int)(q % wi);
   const int j = w.j0 + (int)((q / wi) % wj);
   const int k = w.k0 + (int)(q / (wi * wj));
   for (int c = 0; c < 3; c++) v.set(vidx<1>(b, c, i, j, k), bf ? bf[3 * q + c] : 0.0);
}

// Ghost shell and, in the same pass, the zero ghost values of the sides in
// zfill (bit s: side s is Dirichlet or supergrid with zero data, its two
// ghost layers): at such a point the shell's update is followed by the
// view's set(0), leaving exactly what k_comp_shell and then k_cbc_fill
// would.  Grid-point sources never lie in the shell, so the sources that
// run in between do not see the difference.
template <bool CORR>
__global__ void k_comp_shell_fill(Box b, CompState st, float dt2, double dt2d, int zfill)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   int i, j, k;
   if (!shell_point(b, q, i, j, k)) return;
   const bool zero = ((zfill & 1) && i <= b.ib + 1) || ((zfill & 2) && i >= b.ie - 1) ||
                     ((zfill & 4) && j <= b.jb + 1) || ((zfill & 8) && j >= b.je - 1) ||
                     ((zfill & 16) && k <= b.kb + 1) || ((zfill & 32) && k >= b.ke - 1);
   const long p = sidx(b, i, j, k);
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long v = p + b.npts * c;
      if (!CORR) {
         st.A[v] = 0.0f;
         if (zero) PredView{st.U, st.C, st.D, st.A, dt2d}.set(v, 0.0);
      } else {
         const float dn = st.D[v] + dt2 * st.A[v];
         float U = st.U[v], C = st.C[v];
         comp_add(U, C, dn);
         if (zero) {
            const double u = (double)U + (double)C;
            st.Dn[v] = (float)((double)dn + (0.0 - u));
            st.U[v] = 0.0f;
            st.C[v] = 0.0f;
         } else {
            st.Dn[v] = dn;
            st.U[v] = U;
            st.C[v] = C;
         }
      }
   }
}

template <class V>
__global__ void k_cbc_periodic(Box b, CompWindow w, int pi, int pj, int pk, V v)
{
   const long wi = w.i1 - w.i0 + 1, wj = w.j1 - w.j0 + 1, wk = w.k1 - w.k0 + 1;
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= wi * wj * wk) return;
   const int i = w.i0 + (int)(q % wi);
   const int j = w.j0 + (int)((q / wi) % wj);
   const int k = w.k0 + (int)(q / (wi * wj));
   for (int c = 0; c < 3; c++) v.set(vidx<1>(b, c, i, j, k), v[vidx<1>(b, c, i + pi, j + pj, k + pk)]);
}

// Free-surface ghost values from the FP64 closure (freesurf_point) on the view.
template <class V>
__global__ void k_cbc_freesurf(Box b, int k, int kl, double h, const SbpTables* __restrict__ sbp, V v,
                               const double* __restrict__ mu, const double* __restrict__ la,
                               const double* __restrict__ strx, const double* __restrict__ stry)
{
   const int i = b.ib + 2 + blockIdx.x * blockDim.x + threadIdx.x;
   const int j = b.jb + 2 + blockIdx.y * blockDim.y + threadIdx.y;
   if (i > b.ie - 2 || j > b.je - 2) return;
   const double t[3] = {0, 0, 0};
   double g[3];
   freesurf_point<1>(b, i, j, k, kl, h, sbp->sbop, v, mu, la, t, strx, stry, g);
   for (int c = 0; c < 3; c++) v.set(vidx<1>(b, c, i, j, k - kl), g[c]);
}

__global__ void k_comp_record(Box b, int nrec, const int* __restrict__ rec, CompState st, double dt,
                              double* __restrict__ out)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= nrec) return;
   const long p = sidx(b, rec[4 * r], rec[4 * r + 1], rec[4 * r + 2]);
   const int mode = rec[4 * r + 3];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      const long q = p + b.npts * c;
      out[3 * r + c] = mode == 1 ? ((double)st.Dn[q] + (double)st.D[q]) / (2 * dt) : (double)st.U[q] + (double)st.C[q];
   }
}

__global__ void k_comp_init(long n, const double* __restrict__ u, const double* __restrict__ um,
                            float* __restrict__ U, float* __restrict__ C, float* __restrict__ D)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= n) return;
   const double a = u[q];
   const float hi = (float)a;
   U[q] = hi;
   C[q] = (float)(a - (double)hi);
   D[q] = (float)(a - um[q]);
}

__global__ void k_comp_flush(long n, const float* __restrict__ U, const float* __restrict__ C,
                             const float* __restrict__ D, double* __restrict__ u, double* __restrict__ um)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q >= n) return;
   const double a = (double)U[q] + (double)C[q];
   u[q] = a;
   if (um) um[q] = a - (double)D[q];
}

__global__ void k_to_float(long n, const double* __restrict__ src, float* __restrict__ dst)
{
   const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if (q < n) dst[q] = (float)src[q];
}

// ===========================================================================
// launchers
// ===========================================================================
int launch_comp_init(long n3, const double* u, const double* um, float* U, float* C, float* D, hipStream_t s)
{
   k_comp_init<<<cdiv(n3, 256), 256, 0, s>>>(n3, u, um, U, C, D);
   return check_hip(hipGetLastError(), "FP32 state conversion");
}

int launch_comp_flush(long n3, const float* U, const float* C, const float* D, double* u, double* um,
                      hipStream_t s)
{
   k_comp_flush<<<cdiv(n3, 256), 256, 0, s>>>(n3, U, C, D, u, um);
   return check_hip(hipGetLastError(), "FP32 state write-back");
}

int launch_to_float(long n, const double* src, float* dst, hipStream_t s)
{
   if (n > 0) k_to_float<<<cdiv(n, 256), 256, 0, s>>>(n, src, dst);
   return check_hip(hipGetLastError(), "float conversion");
}

namespace {
int comp_sweep(bool corr, const Box& b, int k1, int k2, int kc, const CompState& st, const SgCoefF& sg, float cof,
               float dt2, float c4, float beta, hipStream_t s, bool async)
{
   if (k2 < k1) return OK;
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   if (kc <= 0) kc = 64;
   CompTiles map;
   map.ntx = (int)cdiv(nx, TX);
   map.nty = (int)cdiv(ny, TY);
   map.ntiles = map.ntx * map.nty * (int)cdiv(k2 - k1 + 1, kc);
   const unsigned grd = 8 * cdiv(map.ntiles, 8);
   // asynchronous loads where every staged quad of floats lies in one row, aligned
   auto al = [](const float* p) { return ((uintptr_t)p & 15) == 0; };
   const bool fits = async && b.ni % 4 == 0 && al(st.mu) && al(st.la) && al(st.rho) && (corr ? al(st.A) && al(st.D) : al(st.U));
   if (corr && fits)
      k_comp_dma<true><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, st.A, st.mu, st.la, st.rho, sg, st.D, st.Dn, st.U, st.C,
                                          nullptr, cof, dt2, c4, beta);
   else if (corr)
      k_comp_sweep<true><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, st.A, st.mu, st.la, st.rho, sg, st.D, st.Dn, st.U,
                                            st.C, nullptr, cof, dt2, c4, beta);
   else if (fits)
      k_comp_dma<false><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, st.U, st.mu, st.la, st.rho, sg, nullptr, nullptr,
                                           nullptr, nullptr, st.A, cof, dt2, c4, beta);
   else
      k_comp_sweep<false><<<grd, NT, 0, s>>>(b, map, k1, k2, kc, st.U, st.mu, st.la, st.rho, sg, nullptr, nullptr,
                                             nullptr, nullptr, st.A, cof, dt2, c4, beta);
   return check_hip(hipGetLastError(), corr ? "FP32 corrector sweep launch" : "FP32 predictor sweep launch");
}

int comp_rows(bool corr, const Box& b, int side, int nk, const SbpTablesT<float>* d_sbpf, const CompState& st,
              const SgCoefF& sg, float cof, float dt2, float c4, float beta, hipStream_t s)
{
   const long nx = b.ie - b.ib - 3, ny = b.je - b.jb - 3;
   const int ntx = (int)cdiv(nx, TX), nty = (int)cdiv(ny, TY);
   if (corr)
      k_comp_rows<true><<<ntx * nty, NT, 0, s>>>(b, side == 4, nk, ntx, d_sbpf, st.A, st.mu, st.la, st.rho, sg, st.D,
                                                 st.Dn, st.U, st.C, nullptr, cof, dt2, c4, beta);
   else
      k_comp_rows<false><<<ntx * nty, NT, 0, s>>>(b, side == 4, nk, ntx, d_sbpf, st.U, st.mu, st.la, st.rho, sg,
                                                  nullptr, nullptr, nullptr, nullptr, st.A, cof, dt2, c4, beta);
   return check_hip(hipGetLastError(), "FP32 SBP rows launch");
}
}  // namespace

int launch_comp_pred(const Box& b, int nk, const int* onesided, const SbpTablesT<float>* d_sbpf, const CompState& st,
                     const SgCoefF& sgf, double h, hipStream_t s, const std::function<int()>& after_interior,
                     int kchunk, double dt, int zfill, bool async)
{
   const float cof = (float)(1.0 / (h * h));
   const int k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   const int k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
   int rc = comp_sweep(false, b, k1, k2, kchunk, st, sgf, cof, 0.0f, 0.0f, 0.0f, s, async);
   if (rc || (after_interior && (rc = after_interior()))) return rc;
   for (int side = 4; side <= 5; side++)
      if (onesided[side] == 1 && (rc = comp_rows(false, b, side, nk, d_sbpf, st, sgf, cof, 0.0f, 0.0f, 0.0f, s)))
         return rc;
   const long ns = shell_count(b);
   if (zfill)
      k_comp_shell_fill<false><<<cdiv(ns, 256), 256, 0, s>>>(b, st, 0.0f, dt * dt, zfill);
   else
      k_comp_shell<false><<<cdiv(ns, 256), 256, 0, s>>>(b, st, 0.0f);
   return check_hip(hipGetLastError(), "FP32 predictor shell launch");
}

int launch_comp_corr(const Box& b, int nk, const int* onesided, const SbpTablesT<float>* d_sbpf, const CompState& st,
                     const SgCoefF& sgf, double beta, double h, double dt, hipStream_t s,
                     const std::function<int()>& after_interior, int kchunk, int zfill, bool async)
{
   const float cof = (float)(1.0 / (h * h)), dt2 = (float)(dt * dt), c4 = (float)(dt * dt * dt * dt / 12);
   const int k1 = onesided[4] == 1 ? 7 : b.kb + 2;
   const int k2 = onesided[5] == 1 ? nk - 6 : b.ke - 2;
   int rc = comp_sweep(true, b, k1, k2, kchunk, st, sgf, cof, dt2, c4, (float)beta, s, async);
   if (rc || (after_interior && (rc = after_interior()))) return rc;
   for (int side = 4; side <= 5; side++)
      if (onesided[side] == 1 && (rc = comp_rows(true, b, side, nk, d_sbpf, st, sgf, cof, dt2, c4, (float)beta, s)))
         return rc;
   const long ns = shell_count(b);
   if (zfill)
      k_comp_shell_fill<true><<<cdiv(ns, 256), 256, 0, s>>>(b, st, dt2, dt * dt, zfill);
   else
      k_comp_shell<true><<<cdiv(ns, 256), 256, 0, s>>>(b, st, dt2);
   return check_hip(hipGetLastError(), "FP32 corrector shell launch");
}

int launch_comp_src(bool corr, const Box& b, const SourceTable& tab, double t, double coef, const CompState& cs,
                    hipStream_t s)
{
   if (tab.nunique <= 0) return OK;
   if (corr)
      k_comp_src<true><<<cdiv(tab.nunique, 128), 128, 0, s>>>(b, tab, t, coef, cs);
   else
      k_comp_src<false><<<cdiv(tab.nunique, 128), 128, 0, s>>>(b, tab, t, coef, cs);
   return check_hip(hipGetLastError(), "FP32 source launch");
}

namespace {
template <class V>
int comp_bc_view(const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd, const V& v, double h,
                 const SbpTables* d_sbp, const double* mu, const double* la, const double* strx, const double* stry,
                 hipStream_t s, int skip)
{
   for (int side = 0; side < 6; side++) {
      if (skip & (1 << side)) continue;  // zero ghost values already set with the shell
      const int* w = wind + 6 * side;
      const CompWindow win{w[0], w[1], w[2], w[3], w[4], w[5]};
      const long n = (long)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1);
      if (bccnd[side] == 1 || bccnd[side] == 2) {
         if (n > 0) k_cbc_fill<V><<<cdiv(n, 256), 256, 0, s>>>(b, win, v, nullptr);
      } else if (bccnd[side] == 3) {
         if (n <= 0) continue;
         const int pi = side == 0 ? nx : side == 1 ? -nx : 0;
         const int pj = side == 2 ? ny : side == 3 ? -ny : 0;
         const int pk = side == 4 ? nz : side == 5 ? -nz : 0;
         k_cbc_periodic<V><<<cdiv(n, 256), 256, 0, s>>>(b, win, pi, pj, pk, v);
      } else if (bccnd[side] == 0) {
         if (side != 4 && side != 5)
            return set_error(E_INVAL, "bcfortsg: free surface is only implemented on sides 4 and 5");
         const int k = side == 4 ? 1 : nz, kl = side == 4 ? 1 : -1;
         const int nii = b.ie - b.ib - 3, njj = b.je - b.jb - 3;
         dim3 blk(64, 4, 1), grd(cdiv(nii, 64), cdiv(njj, 4), 1);
         k_cbc_freesurf<V><<<grd, blk, 0, s>>>(b, k, kl, h, d_sbp, v, mu, la, strx, stry);
      }
   }
   return check_hip(hipGetLastError(), "FP32 boundary condition launch");
}
}  // namespace

int launch_comp_bc(bool corr, const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd,
                   const CompState& cs, double h, double dt, const SbpTables* d_sbp, const double* mu,
                   const double* la, const double* strx, const double* stry, hipStream_t s, int skip)
{
   if (corr) {
      const CorrView v{cs.U, cs.C, cs.Dn};
      return comp_bc_view(b, wind, nx, ny, nz, bccnd, v, h, d_sbp, mu, la, strx, stry, s, skip);
   }
   const PredView v{cs.U, cs.C, cs.D, cs.A, dt * dt};
   return comp_bc_view(b, wind, nx, ny, nz, bccnd, v, h, d_sbp, mu, la, strx, stry, s, skip);
}

int launch_comp_record(const Box& b, int nrec, const int* rec, const CompState& cs, double dt, double* out,
                       hipStream_t s)
{
   if (nrec <= 0) return OK;
   k_comp_record<<<cdiv(nrec, 64), 64, 0, s>>>(b, nrec, rec, cs, dt, out);
   return check_hip(hipGetLastError(), "FP32 receiver launch");
}

}  // namespace sw4amd
