// adjoint.cpp -- discrete adjoint of the FP64 time step: material gradients
// of a receiver misfit by a reverse sweep over checkpointed forward states.
//
// One step (EW::timesteploop, EW.C:2527-2842) maps (U^n, U^{n-1}) to U^{n+1}:
//   Up  = BC(2U - Um + dt^2/rho (L(U) + F))                     predictor
//   A   = (Up - 2U + Um) / dt^2                                 dpdmt
//   U^{n+1} = BC(Up + dt^4/(12 rho) (L(A) + F_tt) - D(U - Um))  corrector, damping
// with L = rhs4sg on the rows (zero on the ghost points), D the supergrid
// damping on the interior and BC = bcfortsg with zero boundary data.  The
// reverse step applies the transposes of these maps in the opposite order
// (kernels_adj.hip) and accumulates the derivatives of L, D, BC and the 1/rho
// factors with respect to the material.  The reverse sweep recomputes the
// forward states of one checkpoint segment at a time.
#include <cmath>
#include <map>
#include <vector>

#include "ew.h"
#include "kernels_adj.h"
#include "kernels_fast.h"

namespace sw4amd {

int launch_fill(double* a, long n, double v, hipStream_t s);

int EW::reset()
{
   if (!m_setup_done) return set_error(E_STATE, "EW: reset before setup");
   int rc;
   if ((rc = comp_end())) return rc;
   m_comp = false;
   if ((rc = launch_fill(mU.dev, mU.size(), 0.0, m_stream)) || (rc = launch_fill(mUm.dev, mUm.size(), 0.0, m_stream)))
      return rc;
   mT = 0;
   mCycle = 0;
   m_nsamples = 0;
   return record(mU, mUm);
}

int EW::fused_once(double t, double* acc)
{
   const GridSetup& g = m_opt.grid;
   const SgCoef c = sg();
   const SourceTable st = sources();
   const double dt = mDt;
   const int zf = zfill();
   // kernel variant 8: the acceleration stored in acc (a work field) as in
   // the time-step loop
   int ka = 0, kz = 0;
   fused_planes(m_box, g.nz, m_onesided, &ka, &kz);
   double* af = acc && m_variant >= 8 && (m_box.ni % 2 == 0 || m_variant >= 10) && kz >= ka ? acc : nullptr;
   const DampTile* dtiles = nullptr;
   int ndtiles = 0;
   int rc;
   if ((rc = damp_tile_list(ka, kz, &dtiles, &ndtiles))) return rc;
   if ((rc = launch_pred_fused(m_box, g.nz, m_onesided, m_dsbp, mU.dev, mUm.dev, mMu.dev, mLambda.dev, mRho.dev, c,
                               mUp.dev, g.h, dt, m_stream, nullptr, m_variant, m_kchunk, zf, af)) ||
       (rc = launch_src_apply(m_box, st, t, 0, dt * dt, mRho.dev, mUp.dev, m_stream)) ||
       (rc = enforceBC(mUp, mMu, mLambda, t + dt, zf)))
      return rc;
   if (af && ((rc = launch_accel_fix(m_box, ka, kz, (zf & 15) != 15, true, mUp.dev, mU.dev, mUm.dev, dt, af,
                                     m_stream)) ||
              (rc = launch_accel_points(m_box, st, mUp.dev, mU.dev, mUm.dev, dt, af, m_stream))))
      return rc;
   if ((rc = launch_corr_fused(m_box, g.nz, m_onesided, m_dsbp, mUp.dev, mU.dev, mUm.dev, mMu.dev, mLambda.dev,
                               mRho.dev, c, m_opt.beta, mUp2.dev, m_dacc, g.h, dt, m_stream, nullptr, nullptr, m_variant,
                               m_kchunk, zf, af, dtiles, ndtiles)) ||
       (rc = launch_src_apply(m_box, st, t, 1, dt * dt * dt * dt / 12, mRho.dev, mUp2.dev, m_stream)) ||
       (rc = enforceBC(mUp2, mMu, mLambda, t + dt, zf)))
      return rc;
   return OK;
}

// Reverse of step n: U = U^n, Um = U^{n-1}, U1 = U^{n+1}; lam_a (adjoint of
// U^{n+1}) is consumed and used as work space; lam_b += dJ/dU^n and
// lam_c = dJ/dU^{n-1} through this step (overwritten).  Of the forward step
// only Up is recomputed (fused predictor, sources, BC).  The 1/rho derivatives
// use state differences on the interior, where every grid-point source lies:
//   dt^2/rho (L(U) + F)        = Up - 2U + Um = dt^2 A,
//   c4/rho (L(A) + F_tt) - D   = U^{n+1} - Up,
// so d/drho [y . c4/rho (...)] + d/drho [-y . D]|_(1/rho) = -y . (U^{n+1} - Up)/rho.
// tmp: 4 work fields of 3 components on the box and one on the padded box
// (pad_box); pad: padded material and the
// transposed operator's work fields (kernels_adj.h), pad.w zero in the
// pad.
// up_st (with acc_st where the predictor stores the acceleration): Up (and
// A) of this step kept from the segment's recomputation, the same bits as
// the recomputation here would give; then nothing of the forward step is
// recomputed.
int EW::step_vjp(int n, double* u, double* um, double* u1, double* lam_a, double* lam_b, double* lam_c,
                 double* grho, double* gmu, double* gla, double* const* tmp, double* plane, const AdjPad& pad,
                 double* up_st, double* acc_st)
{
   const GridSetup& g = m_opt.grid;
   const SgCoef c = sg();
   const SourceTable st = sources();
   const double dt = mDt, t = n * dt, dt2 = dt * dt, c4 = dt2 * dt2 / 12;
   const long n3 = 3 * m_box.npts;
   const RowMap rm = make_rowmap(m_box, g.nz, m_onesided);
   double *Up = tmp[0], *Acc = tmp[1], *Db = tmp[2], *Z = tmp[3];
   // the second transposed operator's result goes to tmp[4] (on the padded
   // box), so that the first one's stays in pad.o for the final pass
   AdjPad pad2 = pad;
   pad2.o = tmp[4];
   // Y: the adjoint of the corrected field, then (in place) of Up
   double* Y = lam_a;
   Sarray UpA = mU;
   UpA.dev = Up;
   int rc;
   // forward: Up = BC(2U - Um + dt^2/rho (L(U) + F)), A = (Up - 2U + Um)/dt^2.
   // Kernel variant 8: the predictor sweep and its ghost shell store A, the
   // planes outside the sweep, the side frame (when the boundary conditions
   // change it) and the source points are completed from the final Up, in
   // the expression k_dpdmt uses (the same bits); else one k_dpdmt pass.
   const int zf = zfill();
   int ka = 0, kz = 0;
   fused_planes(m_box, g.nz, m_onesided, &ka, &kz);
   double* accf = m_variant >= 8 && (m_box.ni % 2 == 0 || m_variant >= 10) && kz >= ka ? Acc : nullptr;
   if (up_st) {
      Up = up_st;
      if (accf && acc_st)
         Acc = acc_st;
      else if ((rc = launch_dpdmt(m_box, Up, u, um, Acc, 1.0 / dt2, m_stream)))
         return rc;
   } else {
      if ((rc = launch_pred_fused(m_box, g.nz, m_onesided, m_dsbp, u, um, mMu.dev, mLambda.dev, mRho.dev, c, Up, g.h,
                                  dt, m_stream, nullptr, m_variant, m_kchunk, zf, accf)) ||
          (rc = launch_src_apply(m_box, st, t, 0, dt2, mRho.dev, Up, m_stream)) ||
          (rc = enforceBC(UpA, mMu, mLambda, t + dt, zf)))
         return rc;
      if (accf) {
         if ((rc = launch_accel_fix(m_box, ka, kz, (zf & 15) != 15, true, Up, u, um, dt, Acc, m_stream)) ||
             (rc = launch_accel_points(m_box, st, Up, u, um, dt, Acc, m_stream)))
            return rc;
      } else if ((rc = launch_dpdmt(m_box, Up, u, um, Acc, 1.0 / dt2, m_stream)))
         return rc;
   }
   // corrector BC, transposed (acting on U^{n+1})
   if ((rc = launch_bcfortsg_T(m_box, m_wind, g.nx, g.ny, g.nz, g.bc, Y, u1, g.h, m_dsbp, mMu.dev, mLambda.dev, c.strx,
                               c.stry, gmu, gla, plane, m_stream)))
      return rc;
   // damping: U^{n+1} -= D(U - Um); corrector: Up + c4/rho (L(A) + F_tt).
   // Kernel variant 9: the damping transpose runs over the supergrid layers
   // only; Db is zero elsewhere from the start of the reverse sweep.
   const DampTile* dT = nullptr;
   int ndT = 0;
   if ((rc = damp_T_tile_list(&dT, &ndT))) return rc;
   if ((rc = launch_addsgd4_T_march(m_box, Y, u, um, mRho.dev, c, m_opt.beta, Db, grho, m_stream, false, dT, ndT)) ||
       (rc = launch_vjp_corr_pw(m_box, rm, c.strx, c.stry, c.strz, c4, mRho.dev, Y, u1, Up, Z, grho, pad, m_stream)) ||
       (rc = launch_rhs4sg_T(m_box, rm, m_dsbp, Z, mMu.dev, mLambda.dev, g.h, c.strx, c.stry, c.strz, nullptr, pad,
                             m_stream, true)) ||
       (rc = launch_rhs4sg_matgrad(m_box, rm, m_dsbp, Z, Acc, mMu.dev, mLambda.dev, g.h, c.strx, c.stry, c.strz, gmu,
                                   gla, m_stream)))
      return rc;
   // A = (Up - 2U + Um)/dt^2: the Up part (L^T of the corrector stays in pad.o for the U, Um parts)
   if ((rc = launch_vjp_axpy_pad(m_box, 1.0 / dt2, pad, Y, m_stream))) return rc;
   // predictor BC, transposed (acting on Up)
   if ((rc = launch_bcfortsg_T(m_box, m_wind, g.nx, g.ny, g.nz, g.bc, Y, Up, g.h, m_dsbp, mMu.dev, mLambda.dev, c.strx,
                               c.stry, gmu, gla, plane, m_stream)))
      return rc;
   // predictor: 2U - Um + dt^2/rho (L(U) + F); then the U, Um parts of every stage
   if ((rc = launch_vjp_pred_pw(m_box, rm, c.strx, c.stry, c.strz, dt2, mRho.dev, Y, Acc, Z, grho, pad, m_stream)) ||
       (rc = launch_rhs4sg_T(m_box, rm, m_dsbp, Z, mMu.dev, mLambda.dev, g.h, c.strx, c.stry, c.strz, nullptr, pad2,
                             m_stream, true)) ||
       (rc = launch_rhs4sg_matgrad(m_box, rm, m_dsbp, Z, u, mMu.dev, mLambda.dev, g.h, c.strx, c.stry, c.strz, gmu,
                                   gla, m_stream)) ||
       (rc = launch_vjp_lam(m_box, 1.0 / dt2, Db, pad, Y, pad2, lam_b, lam_c, m_stream)))
      return rc;
   return OK;
}

void* WorkPool::take(size_t bytes)
{
   {
      std::lock_guard<std::mutex> lk(m_mtx);
      Buf* best = nullptr;
      for (Buf& b : m_bufs)
         if (!b.used && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
      if (best) {
         best->used = true;
         return best->p;
      }
   }
   // allocations and frees run outside the lock, so that the other callers'
   // takes and gives do not wait for them
   void* p = nullptr;
   if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      // make room: free the idle buffers (none fits this request), then try once more
      std::vector<void*> idle;
      {
         std::lock_guard<std::mutex> lk(m_mtx);
         std::vector<Buf> keep;
         for (const Buf& b : m_bufs) {
            if (b.used)
               keep.push_back(b);
            else
               idle.push_back(b.p);
         }
         m_bufs.swap(keep);
      }
      for (void* q : idle) (void)hipFree(q);
      if (hipMalloc(&p, bytes) != hipSuccess) {
         (void)hipGetLastError();
         return nullptr;
      }
   }
   std::lock_guard<std::mutex> lk(m_mtx);
   m_bufs.push_back(Buf{bytes, p, true});
   return p;
}

void WorkPool::give(void* p)
{
   std::lock_guard<std::mutex> lk(m_mtx);
   for (Buf& b : m_bufs)
      if (b.p == p) b.used = false;
}

double WorkPool::held() const
{
   std::lock_guard<std::mutex> lk(m_mtx);
   double t = 0;
   for (const Buf& b : m_bufs) t += (double)b.bytes;
   return t;
}

void WorkPool::release()
{
   std::lock_guard<std::mutex> lk(m_mtx);
   for (const Buf& b : m_bufs) (void)hipFree(b.p);
   m_bufs.clear();
}

void* EW::pool_take(size_t bytes)
{
   if (m_shared) {
      void* p = m_shared->take(bytes);
      if (p) m_taken.push_back(p);
      return p;
   }
   for (PoolBuf& b : m_pool)
      if (!b.used && b.bytes == bytes) {
         b.used = true;
         return b.p;
      }
   void* p = nullptr;
   if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      // make room: drop the idle buffers of other sizes, then try once more
      std::vector<PoolBuf> keep;
      for (const PoolBuf& b : m_pool) {
         if (b.used)
            keep.push_back(b);
         else
            (void)hipFree(b.p);
      }
      m_pool.swap(keep);
      if (hipMalloc(&p, bytes) != hipSuccess) {
         (void)hipGetLastError();
         return nullptr;
      }
   }
   m_pool.push_back(PoolBuf{bytes, p, true});
   return p;
}

void EW::pool_return()
{
   for (PoolBuf& b : m_pool) b.used = false;
   if (!m_taken.empty()) {
      // kernels of this stream may still read them (error paths)
      (void)hipStreamSynchronize(m_stream);
      for (void* p : m_taken) m_shared->give(p);
      m_taken.clear();
   }
}

int EW::release_workspace()
{
   int rc = check_hip(hipStreamSynchronize(m_stream), "EW: workspace release");
   for (const PoolBuf& b : m_pool) (void)hipFree(b.p);
   m_pool.clear();
   return rc;
}

double EW::workspace_bytes() const
{
   double t = 0;
   for (const PoolBuf& b : m_pool) t += (double)b.bytes;
   return t;
}

// Device doubles of the gradient's work memory besides the stored states:
// 4 work fields and one on the padded box, 3 adjoint fields, 3 gradients,
// the padded material, one plane of the boundary transposes and the
// receiver traces.
static double adjoint_fixed_doubles(const Box& b, int nsteps, int nrec)
{
   return 7.0 * 3 * b.npts + 3.0 * pad_box(b).npts + 3.0 * b.npts + (double)adj_pad_doubles(b) + 3.0 * b.nij +
          3.0 * (nsteps + 1) * nrec + 2.0 * (8 * nrec + 1);
}

// States held by a checkpoint interval c: the segment buffer (c + 2) and two
// per checkpoint (U, Um), or 1.5 for the float (U, C, D) of the FP32 form;
// with store, the predicted field and the acceleration of every step of a
// segment (2 c).  The reverse sweep recomputes every segment but the last,
// nsteps - (length of the last segment) forward steps in all.
static double adjoint_states(int nsteps, int c, bool fp32, bool store)
{
   return (c + 2) + (store ? 2.0 * c : 0.0) + (fp32 ? 1.5 : 2.0) * ((nsteps + c - 1) / c);
}

int EW::gradient_plan(int nsteps, int ckpt, int* used, double* bytes, bool fp32, int store, int* stored)
{
   if (nsteps < 1) return set_error(E_INVAL, "EW: gradient plan needs nsteps >= 1");
   const int nrec = (int)m_rec.size() / 4;
   const double n3 = 3.0 * m_box.npts, fixed = adjoint_fixed_doubles(m_box, nsteps, nrec);
   // the FP32 form recomputes float steps, which give the FP64 reverse step
   // no predicted field to keep
   const bool can_store = !fp32 && store != 0;
   int c = std::min(ckpt, nsteps);
   bool st = can_store && store == 1;
   if (ckpt < 1 || (can_store && store < 0)) {
      // free device memory (with the workspace already held), keeping a reserve
      size_t freeb = 0, totalb = 0;
      if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
         (void)hipGetLastError();
         freeb = 0;
      }
      const double reserve = 6.0 * (1 << 30) + 0.02 * (double)totalb;
      const double room = ((double)freeb + workspace_bytes() - reserve) / 8.0 - fixed;
      // the largest interval whose states fit (all of them when they fit: no recomputation)
      auto largest = [&](bool s) {
         for (int q = nsteps; q >= 1 && room > 0; q--)
            if (adjoint_states(nsteps, q, fp32, s) * n3 <= room) return q;
         return 0;
      };
      if (ckpt < 1) {
         c = largest(false);
         if (c < 1) c = std::max(1, (int)std::lround(std::sqrt((double)nsteps)));
         if (can_store) {
            // forward steps recomputed: every segment but the last; without
            // the kept fields every reverse step recomputes its predictor as
            // well, 0.4 of a forward step (256^3, MI355X)
            auto recomputed = [&](int q) { return (double)((nsteps + q - 1) / q - 1) * q; };
            const int cs = largest(true);
            st = cs >= 1 && (store == 1 || recomputed(cs) < recomputed(c) + 0.4 * nsteps);
            if (st) c = cs;
         }
      } else if (store < 0) {
         st = adjoint_states(nsteps, c, fp32, true) * n3 <= room;
      }
   }
   if (used) *used = c;
   if (stored) *stored = st ? 1 : 0;
   // the FP32 form's own state (18 floats a point) is held outside the pool
   if (bytes)
      *bytes = 8.0 * (fixed + adjoint_states(nsteps, c, fp32, st) * n3) + (fp32 && !m_cbuf ? 72.0 * m_box.npts : 0.0);
   return OK;
}

int EW::gradient(int nsteps, int ckpt, const double* gtr, double* grho, double* gmu, double* gla, bool on_device,
                 bool fp32, int store)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: gradient before setup");
   if (!m_opt.corder) return set_error(E_INVAL, "EW: the adjoint needs the component-blocked layout (corder=1)");
   if (m_curv) return set_error(E_INVAL, "EW: the adjoint is Cartesian");
   const GridSetup& g = m_opt.grid;
   for (int s = 0; s < 6; s++)
      if (g.bc[s] == BC_PERIODIC) return set_error(E_INVAL, "EW: the adjoint does not support periodic sides");
   const int nrec = (int)m_rec.size() / 4;
   if (nsteps < 1 || nsteps + 1 > m_maxsamples || nrec == 0)
      return set_error(E_INVAL, "EW: gradient needs receivers and 1 <= nsteps <= the set-up step count");
   if (!gtr || !grho || !gmu || !gla) return set_error(E_INVAL, "EW: null gradient argument");
   int keep = 0;  // the segment's predicted fields and accelerations are kept
   {
      int rc0 = gradient_plan(nsteps, ckpt, &ckpt, nullptr, fp32, store, &keep);
      if (rc0) return rc0;
   }
   const long n1 = m_box.npts, n3 = 3 * n1;
   const int ncp = (nsteps + ckpt - 1) / ckpt;
   // work memory: 5 fields, 3 adjoint fields, ckpt+2 segment states (and 2 ckpt
   // kept fields), 2 per checkpoint, 3 gradients, pad, traces
   // buffers come from the workspace pool and go back to it at the end
   auto release = [&]() { pool_return(); };
   auto grab = [&](size_t n, double** out) {
      void* p = pool_take(sizeof(double) * n);
      if (!p) return set_error(E_NOMEM, "EW: adjoint work allocation failed");
      *out = static_cast<double*>(p);
      return (int)OK;
   };
   int rc = OK;
   double* tmp[5];
   double *lam[3], *gr[3], *plane = nullptr;
   std::vector<double*> seg(ckpt + 2), cp(2 * ncp);
   for (int q = 0; q < 4 && rc == OK; q++) rc = grab(n3, &tmp[q]);
   if (rc == OK) rc = grab(3 * pad_box(m_box).npts, &tmp[4]);  // on the padded box
   double* padmem = nullptr;
   if (rc == OK) rc = grab(adj_pad_doubles(m_box), &padmem);
   for (int q = 0; q < 3 && rc == OK; q++) rc = grab(n3, &lam[q]);
   for (int q = 0; q < 3 && rc == OK; q++) rc = grab(n1, &gr[q]);
   for (int q = 0; q <= ckpt + 1 && rc == OK; q++) rc = grab(n3, &seg[q]);
   // kept[2 m], kept[2 m + 1]: Up and the acceleration of the segment's step m
   std::vector<double*> kept(keep ? 2 * ckpt : 0);
   for (size_t q = 0; q < kept.size() && rc == OK; q++) rc = grab(n3, &kept[q]);
   // checkpoints: (U^{s0}, U^{s0-1}) in FP64, or the float (U, C, D) of the
   // FP32 form (cp[2 q], 9 floats a point) at the start of segment q
   for (int q = 0; q < 2 * ncp && rc == OK; q++)
      if (!fp32) rc = grab(n3, &cp[q]);
      else if (q % 2 == 0) rc = grab((size_t)(4.5 * n1) + 1, &cp[q]);
   if (rc == OK) rc = grab(3 * m_box.nij, &plane);
   // receiver traces as adjoint sources, grouped by the grid value they act
   // on (component offsets fit an int for the grids this path is meant for)
   if (rc == OK && n3 > 0x7fffffff) rc = set_error(E_RANGE, "EW: adjoint source index beyond 32 bits");
   std::map<int, std::vector<int>> groups;
   std::vector<int> hmode(nrec);
   for (int r = 0; r < nrec; r++) {
      const long p = sidx(m_box, m_rec[4 * r], m_rec[4 * r + 1], m_rec[4 * r + 2]);
      hmode[r] = m_rec[4 * r + 3];
      for (int c = 0; c < 3; c++) groups[(int)(p + n1 * c)].push_back(3 * r + c);
   }
   std::vector<int> hint;  // uidx | off | ent | mode
   for (const auto& gq : groups) hint.push_back(gq.first);
   hint.push_back(0);
   for (const auto& gq : groups) hint.push_back(hint.back() + (int)gq.second.size());
   for (const auto& gq : groups) hint.insert(hint.end(), gq.second.begin(), gq.second.end());
   hint.insert(hint.end(), hmode.begin(), hmode.end());
   const size_t ntr = (size_t)(nsteps + 1) * 3 * nrec;
   double* dtr = nullptr;
   int* dint = nullptr;
   if (rc == OK) rc = grab(ntr, &dtr);
   if (rc == OK) {
      dint = static_cast<int*>(pool_take(sizeof(int) * hint.size()));
      if (!dint) rc = set_error(E_NOMEM, "EW: adjoint source allocation failed");
   }
   if (rc == OK)
      rc = check_hip(hipMemcpyAsync(dint, hint.data(), sizeof(int) * hint.size(), hipMemcpyHostToDevice, m_stream),
                     "adjoint source upload");
   if (rc == OK)
      rc = check_hip(hipMemcpyAsync(dtr, gtr, sizeof(double) * ntr, hipMemcpyHostToDevice, m_stream),
                     "adjoint source upload");
   if (rc == OK) rc = check_hip(hipStreamSynchronize(m_stream), "adjoint source upload");
   if (rc) {
      release();
      return rc;
   }
   ReceiverAdj R;
   R.nu = (int)groups.size();
   R.nrec = nrec;
   R.uidx = dint;
   R.off = dint + R.nu;
   R.ent = R.off + R.nu + 1;
   R.mode = R.ent + 3 * nrec;
   R.g = dtr;
   auto add_direct = [&](int s, double* lamp) {
      return launch_receiver_sources_T(R, s, nsteps, 2 * mDt, lamp, m_stream);
   };
   // forward pass: segment by segment into seg (seg[m] = U^{s0-1+m}), saving
   // the checkpoint (U^{s0}, U^{s0-1}) of each; seg ends up holding the last
   // segment, which the reverse sweep then uses without recomputing it
   double *sU = mU.dev, *sUm = mUm.dev, *sUp = mUp.dev, *sUp2 = mUp2.dev;
   auto fill_segment = [&](int cq) {
      const int s0 = cq * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      int e;
      if ((e = check_hip(hipMemcpyAsync(seg[0], cp[2 * cq + 1], sizeof(double) * n3, hipMemcpyDeviceToDevice, m_stream),
                         "checkpoint restore")) ||
          (e = check_hip(hipMemcpyAsync(seg[1], cp[2 * cq], sizeof(double) * n3, hipMemcpyDeviceToDevice, m_stream),
                         "checkpoint restore")))
         return e;
      for (int m = 1; m <= s1 - s0 && e == OK; m++) {
         mU.dev = seg[m];
         mUm.dev = seg[m - 1];
         mUp.dev = keep ? kept[2 * (m - 1)] : tmp[0];
         mUp2.dev = seg[m + 1];
         // tmp[1]: the reverse step's acceleration, free here
         e = fused_once((s0 + m - 1) * mDt, keep ? kept[2 * (m - 1) + 1] : tmp[1]);
      }
      mU.dev = sU;
      mUm.dev = sUm;
      mUp.dev = sUp;
      mUp2.dev = sUp2;
      return e;
   };
   // the FP32 form: the float state (U, C, D) is the checkpoint; each
   // recomputed step's U + C is the FP64 state the reverse step reads
   const size_t fbytes = sizeof(float) * n3;
   auto save_state32 = [&](int cq) {
      const CompState cs = comp_state();
      float* dst = reinterpret_cast<float*>(cp[2 * cq]);
      int e;
      if ((e = check_hip(hipMemcpyAsync(dst, cs.U, fbytes, hipMemcpyDeviceToDevice, m_stream), "checkpoint")) ||
          (e = check_hip(hipMemcpyAsync(dst + n3, cs.C, fbytes, hipMemcpyDeviceToDevice, m_stream), "checkpoint")))
         return e;
      return check_hip(hipMemcpyAsync(dst + 2 * n3, cs.D, fbytes, hipMemcpyDeviceToDevice, m_stream), "checkpoint");
   };
   auto fill_segment32 = [&](int cq) {
      const int s0 = cq * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      m_cd = 0;
      const CompState cs = comp_state();
      const float* src = reinterpret_cast<const float*>(cp[2 * cq]);
      int e;
      if ((e = check_hip(hipMemcpyAsync(cs.U, src, fbytes, hipMemcpyDeviceToDevice, m_stream), "checkpoint restore")) ||
          (e = check_hip(hipMemcpyAsync(cs.C, src + n3, fbytes, hipMemcpyDeviceToDevice, m_stream),
                         "checkpoint restore")) ||
          (e = check_hip(hipMemcpyAsync(cs.D, src + 2 * n3, fbytes, hipMemcpyDeviceToDevice, m_stream),
                         "checkpoint restore")) ||
          (e = launch_comp_flush(n3, cs.U, cs.C, cs.D, seg[1], seg[0], m_stream)))
         return e;
      mT = s0 * mDt;
      mCycle = s0;
      for (int m = 1; m <= s1 - s0 && e == OK; m++) {
         if ((e = step_comp(1))) break;
         const CompState c2 = comp_state();
         e = launch_comp_flush(n3, c2.U, c2.C, c2.D, seg[m + 1], nullptr, m_stream);
      }
      return e;
   };
   if ((rc = reset())) {
      release();
      return rc;
   }
   if (fp32) {
      m_norecord = true;
      rc = comp_begin();
      if (rc == OK) rc = save_state32(0);
   } else {
      rc = check_hip(hipMemcpyAsync(cp[0], mU.dev, sizeof(double) * n3, hipMemcpyDeviceToDevice, m_stream),
                     "checkpoint");
      if (rc == OK)
         rc = check_hip(hipMemcpyAsync(cp[1], mUm.dev, sizeof(double) * n3, hipMemcpyDeviceToDevice, m_stream),
                        "checkpoint");
   }
   for (int cq = 0; cq < ncp && rc == OK; cq++) {
      if (fp32) {
         if ((rc = fill_segment32(cq))) break;
         if (cq + 1 < ncp && (rc = save_state32(cq + 1))) break;
         continue;
      }
      if ((rc = fill_segment(cq))) break;
      if (cq + 1 < ncp) {
         const int len = std::min(nsteps, (cq + 1) * ckpt) - cq * ckpt;  // U^{s1} is seg[len+1]
         if ((rc = check_hip(hipMemcpyAsync(cp[2 * cq + 2], seg[len + 1], sizeof(double) * n3, hipMemcpyDeviceToDevice,
                                            m_stream), "checkpoint")) ||
             (rc = check_hip(hipMemcpyAsync(cp[2 * cq + 3], seg[len], sizeof(double) * n3, hipMemcpyDeviceToDevice,
                                            m_stream), "checkpoint")))
            break;
      }
   }
   // reverse sweep
   const SgCoef sgc = sg();
   const AdjPad pad = make_adj_pad(m_box, padmem);
   if (rc == OK) rc = launch_adj_pad_material(m_box, mMu.dev, mLambda.dev, sgc.strx, sgc.stry, sgc.strz, pad, m_stream);
   // the reverse step writes W on the box points only: its pad stays zero
   if (rc == OK) rc = launch_zero(pad.w, 3 * pad.pb.npts, m_stream);
   for (int q = 0; q < 3 && rc == OK; q++) rc = launch_zero(gr[q], n1, m_stream);
   // the damping transpose's output (tmp[2]) where its layer tiles do not reach
   if (rc == OK && m_variant >= 9) rc = launch_zero(tmp[2], n3, m_stream);
   for (int q = 0; q < 2 && rc == OK; q++) rc = launch_zero(lam[q], n3, m_stream);
   double *la = lam[0], *lb = lam[1], *lc = lam[2];
   if (rc == OK) rc = add_direct(nsteps, la);
   if (rc == OK) rc = add_direct(nsteps - 1, lb);
   for (int cq = ncp - 1; cq >= 0 && rc == OK; cq--) {
      const int s0 = cq * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      if (cq < ncp - 1 && (rc = fp32 ? fill_segment32(cq) : fill_segment(cq))) break;  // the last one is still in seg
      for (int n = s1 - 1; n >= s0 && rc == OK; n--) {
         // lc: the adjoint of U^{n-1} starts here, then takes its direct source
         if ((rc = step_vjp(n, seg[n - s0 + 1], seg[n - s0], seg[n - s0 + 2], la, lb, lc, gr[0], gr[1], gr[2], tmp,
                            plane, pad, keep ? kept[2 * (n - s0)] : nullptr, keep ? kept[2 * (n - s0) + 1] : nullptr)) ||
             (rc = add_direct(n - 1, lc)))
            break;
         // rotate: U^{n} becomes the complete adjoint; la was consumed
         double* old = la;
         la = lb;
         lb = lc;
         lc = old;
      }
   }
   const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
   double* outs[3] = {grho, gmu, gla};
   for (int q = 0; q < 3 && rc == OK; q++)
      rc = check_hip(hipMemcpyAsync(outs[q], gr[q], sizeof(double) * n1, kind, m_stream), "gradient copy");
   const int rs = check_hip(hipStreamSynchronize(m_stream), "gradient synchronize");
   if (fp32) {
      // the float state was the gradient's work: the solution stays the
      // zero data of reset() (U, Um untouched), at t = 0
      m_comp = false;
      m_norecord = false;
      m_cd = 0;
      mT = 0;
      mCycle = 0;
   }
   release();
   return rc ? rc : rs;
}

}  // namespace sw4amd
