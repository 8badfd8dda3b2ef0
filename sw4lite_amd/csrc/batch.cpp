// batch.cpp -- several shots advanced together (declarations in batch.h).
//
// The fused FP64 step of every shot, in the stage order of EW::step_fused:
// the three interior sweeps (predictor, corrector, damping) of all shots are
// one launch each on the batch's stream; the SBP rows, ghost shell, sources,
// ghost-point boundary conditions and receivers run per grid, shot q on side
// stream q mod NSIDE between a fork and a join of the batch's stream.  Sources and
// receivers read the time and the sample index from per-shot device clocks,
// so a captured three-step period (U, Um, Up' rotate with period 3) replays
// without changes.
#include "batch.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>

extern "C" const char* sw4hip_last_error(void);

namespace sw4amd {

Batch::~Batch()
{
   if (m_exec) (void)hipGraphExecDestroy(m_exec);
   for (void* p : m_allocs) (void)hipFree(p);
   for (int i = 0; i < NSIDE; i++) {
      if (m_join[i]) (void)hipEventDestroy(m_join[i]);
      if (m_side[i]) (void)hipStreamDestroy(m_side[i]);
   }
   if (m_fork) (void)hipEventDestroy(m_fork);
   if (m_s) (void)hipStreamDestroy(m_s);
}

int Batch::fork()
{
   int rc = check_hip(hipEventRecord(m_fork, m_s), "batch fork");
   for (int i = 0; i < NSIDE && rc == OK; i++) rc = check_hip(hipStreamWaitEvent(m_side[i], m_fork, 0), "batch fork");
   return rc;
}

int Batch::join()
{
   int rc = OK;
   for (int i = 0; i < NSIDE && rc == OK; i++)
      if ((rc = check_hip(hipEventRecord(m_join[i], m_side[i]), "batch join")) == OK)
         rc = check_hip(hipStreamWaitEvent(m_s, m_join[i], 0), "batch join");
   return rc;
}

int Batch::init()
{
   const int n = (int)m_shots.size();
   if (n < 1) return set_error(E_INVAL, "batch: no shots");
   for (EW* e : m_shots) {
      if (!e) return set_error(E_INVAL, "batch: null shot");
      if (!e->m_setup_done) return set_error(E_STATE, "batch: every shot must be set up");
      if (!e->m_opt.corder || e->m_curv) return set_error(E_INVAL, "batch: shots must be Cartesian with corder = 1");
      if (e->m_comp) return set_error(E_STATE, "batch: a shot holds an FP32 state; write it back first");
   }
   for (size_t a = 0; a < m_shots.size(); a++)
      for (size_t b = a + 1; b < m_shots.size(); b++)
         if (m_shots[a] == m_shots[b]) return set_error(E_INVAL, "batch: a shot is listed twice");
   int rc;
   if ((rc = check_hip(hipStreamCreate(&m_s), "batch stream")) ||
       (rc = check_hip(hipEventCreateWithFlags(&m_fork, hipEventDisableTiming), "batch event")))
      return rc;
   for (int i = 0; i < NSIDE; i++)
      if ((rc = check_hip(hipStreamCreate(&m_side[i]), "batch stream")) ||
          (rc = check_hip(hipEventCreateWithFlags(&m_join[i], hipEventDisableTiming), "batch event")))
         return rc;
   void* p = nullptr;
   const size_t tab = sizeof(SweepShot) * 9 * n, clk = (sizeof(double) * 2 + sizeof(int) * 2) * n;
   if (hipMalloc(&p, tab + clk) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(E_NOMEM, "batch: table allocation failed");
   }
   m_allocs.push_back(p);
   m_dtab = static_cast<SweepShot*>(p);
   m_clk_t = reinterpret_cast<double*>(static_cast<char*>(p) + tab);
   m_clk_dt = m_clk_t + n;
   m_clk_s = reinterpret_cast<int*>(m_clk_dt + n);
   m_clk_max = m_clk_s + n;
   // kernel variant 9: each shot's damping sweep runs over the tiles that
   // reach its supergrid layers (damp_tiles; zero damping elsewhere), in
   // chunks of the single-grid damping sweep's planes
   m_dtiles.assign(n, nullptr);
   m_ndtiles.assign(n, 0);
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      if (e->m_variant < 9) continue;
      const Box& b = e->m_box;
      const GridSetup& g = e->m_opt.grid;
      const int k1 = e->m_onesided[4] == 1 ? 7 : b.kb + 2;
      const int k2 = e->m_onesided[5] == 1 ? g.nz - 6 : b.ke - 2;
      const std::vector<DampTile> v = damp_tiles(b, k1, k2, EW::kDampChunk, e->m_sg_host.dcx.data(),
                                                 e->m_sg_host.dcy.data(), e->m_sg_host.dcz.data());
      if (hipMalloc(&p, sizeof(DampTile) * std::max<size_t>(v.size(), 1)) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "batch: damping tile list allocation failed");
      }
      m_allocs.push_back(p);
      if (!v.empty() &&
          (rc = check_hip(hipMemcpy(p, v.data(), sizeof(DampTile) * v.size(), hipMemcpyHostToDevice), "tile upload")))
         return rc;
      m_dtiles[q] = static_cast<DampTile*>(p);
      m_ndtiles[q] = (int)v.size();
   }
   // kernel variant 8 and later: the predictor stores the acceleration in the
   // shot's own field and the corrector stages it (where both sweeps take
   // the asynchronous loads), as the single-grid step does
   m_acc.assign(n, nullptr);
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      const Box& b = e->m_box;
      const GridSetup& g = e->m_opt.grid;
      const int k1 = e->m_onesided[4] == 1 ? 7 : b.kb + 2;
      const int k2 = e->m_onesided[5] == 1 ? g.nz - 6 : b.ke - 2;
      if (e->m_variant >= 8 && b.ni % 2 == 0 && k2 >= k1 && (rc = e->accel_field(&m_acc[q]))) return rc;
   }
   return refresh();
}

int Batch::refresh()
{
   const int n = (int)m_shots.size();
   m_base.assign(3 * n, nullptr);
   std::vector<double> t(n), dt(n);
   std::vector<int> smp(n), mx(n);
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      m_base[3 * q] = e->mUm.dev;
      m_base[3 * q + 1] = e->mU.dev;
      m_base[3 * q + 2] = e->mUp2.dev;
      t[q] = e->mT;
      dt[q] = e->mDt;
      smp[q] = e->m_nsamples;
      mx[q] = (e->m_rec.empty() || !e->m_drecdata) ? 0 : e->m_maxsamples;
   }
   m_phase = 0;
   int rc;
   if ((rc = check_hip(hipMemcpy(m_clk_t, t.data(), sizeof(double) * n, hipMemcpyHostToDevice), "clock upload")) ||
       (rc = check_hip(hipMemcpy(m_clk_dt, dt.data(), sizeof(double) * n, hipMemcpyHostToDevice), "clock upload")) ||
       (rc = check_hip(hipMemcpy(m_clk_s, smp.data(), sizeof(int) * n, hipMemcpyHostToDevice), "clock upload")) ||
       (rc = check_hip(hipMemcpy(m_clk_max, mx.data(), sizeof(int) * n, hipMemcpyHostToDevice), "clock upload")))
      return rc;
   return build_tables();
}

int Batch::set_kchunk(int kc)
{
   if (kc < 1) return set_error(E_INVAL, "batch: k-block must be positive");
   m_kchunk = kc;
   return build_tables();
}

int Batch::build_tables()
{
   const int n = (int)m_shots.size();
   std::vector<SweepShot> h(9 * n);
   m_useacc.assign(3 * n, 0);
   for (int ph = 0; ph < 3; ph++) {
      // kinds 0 and 2: the grids whose planes take the asynchronous loads
      // first (their own launch, block ids from 0), then the others
      int g0[3] = {0, 0, 0}, ga[3] = {0, 0, 0}, slot[3] = {0, 0, 0};
      for (int pass = 0; pass < 2; pass++)
         for (int q = 0; q < n; q++) {
            EW* e = m_shots[q];
            const GridSetup& g = e->m_opt.grid;
            const Box& b = e->m_box;
            double* um = m_base[3 * q + ph % 3];  // phase ph: (Um, U, Up2) = base rotated by ph
            double* u = m_base[3 * q + (ph + 1) % 3];
            double* up2 = m_base[3 * q + (ph + 2) % 3];
            const SgCoef c = e->sg();
            const int k1 = e->m_onesided[4] == 1 ? 7 : b.kb + 2;
            const int k2 = e->m_onesided[5] == 1 ? g.nz - 6 : b.ke - 2;
            SweepShot* t = &h[(ph * 3) * n];
            const double* pa[3] = {u, e->mMu.dev, e->mLambda.dev};
            const double* da[4] = {u, um, e->mRho.dev, up2};
            const double* ca[4] = {m_acc[q], e->mUp.dev, e->mMu.dev, e->mLambda.dev};
            const bool acc = m_acc[q] && async_loads_fit(b, pa, 3) && async_loads_fit(b, ca, 4);
            m_useacc[ph * n + q] = acc;
            const bool fa[3] = {async_loads_fit(b, pa, 3), acc, async_loads_fit(b, da, 4)};
            const bool damp = e->m_opt.beta != 0;
            for (int k = 0; k < 3; k++) {
               if (fa[k] != (pass == 0)) continue;
               int& gc = fa[k] ? ga[k] : g0[k];
               SweepShot* d = &t[k * n + slot[k]++];
               if (k == 0)
                  gc = fill_sweep_shot(false, b, k1, k2, m_kchunk, nullptr, u, um, e->mMu.dev, e->mLambda.dev,
                                       e->mRho.dev, c, e->mUp.dev, g.h, e->mDt, 0.0, false, gc, d, nullptr, 0,
                                       acc ? m_acc[q] : nullptr);
               else if (k == 1)
                  gc = fill_sweep_shot(false, b, k1, k2, m_kchunk, e->mUp.dev, u, um, e->mMu.dev, e->mLambda.dev,
                                       e->mRho.dev, c, up2, g.h, e->mDt, e->m_opt.beta, true, gc, d, nullptr, 0,
                                       acc ? m_acc[q] : nullptr);
               else
                  gc = fill_sweep_shot(true, b, k1, damp ? k2 : k1 - 1, m_kchunk, nullptr, u, um, nullptr, nullptr,
                                       e->mRho.dev, c, up2, g.h, e->mDt, e->m_opt.beta, true, gc, d, m_dtiles[q],
                                       m_ndtiles[q]);
            }
            if (pass == 0)
               for (int k = 0; k < 3; k++) m_nasync[ph][k] = slot[k];
         }
      for (int k = 0; k < 3; k++) {
         m_nblk[ph][k] = g0[k];
         m_nblk_async[ph][k] = ga[k];
      }
   }
   if (m_exec) {
      (void)hipGraphExecDestroy(m_exec);
      m_exec = nullptr;
   }
   return check_hip(hipMemcpy(m_dtab, h.data(), sizeof(SweepShot) * h.size(), hipMemcpyHostToDevice), "table upload");
}

int Batch::one_step(int ph)
{
   const int n = (int)m_shots.size();
   const SweepShot* tab = m_dtab + (ph * 3) * n;
   int rc;
   const int a0 = m_nasync[ph][0], a1 = m_nasync[ph][1], a2 = m_nasync[ph][2];
   // each grid's SBP rows and ghost shell (which write only the planes and
   // points the interior sweeps do not, from inputs the sweeps do not
   // change) run on the side streams while the batched sweeps run, instead
   // of after them (48 shots: 55.5 -> 55.1 ms per step with graph replay,
   // 57.1 -> 56.4 ms without)
   auto pred_tail = [&](int q) {
      EW* e = m_shots[q];
      const GridSetup& g = e->m_opt.grid;
      double* acc = m_useacc[ph * n + q] ? m_acc[q] : nullptr;
      return launch_pred_tail(e->m_box, g.nz, e->m_onesided, e->m_dsbp, m_base[3 * q + (ph + 1) % 3],
                              m_base[3 * q + ph % 3], e->mMu.dev, e->mLambda.dev, e->mRho.dev, e->sg(), e->mUp.dev,
                              g.h, e->mDt, m_side[q % NSIDE], e->zfill(), acc);
   };
   auto corr_tail = [&](int q) {
      EW* e = m_shots[q];
      const GridSetup& g = e->m_opt.grid;
      const double* acc = m_useacc[ph * n + q] ? m_acc[q] : nullptr;
      return launch_corr_tail(e->m_box, g.nz, e->m_onesided, e->m_dsbp, e->mUp.dev, m_base[3 * q + (ph + 1) % 3],
                              m_base[3 * q + ph % 3], e->mMu.dev, e->mLambda.dev, e->mRho.dev, e->sg(),
                              e->m_opt.beta, m_base[3 * q + (ph + 2) % 3], g.h, e->mDt, m_side[q % NSIDE], e->zfill(),
                              acc, e->m_variant >= 9);
   };
   if ((rc = fork())) return rc;
   for (int q = 0; q < n; q++)
      if ((rc = pred_tail(q))) return rc;
   if ((rc = launch_sweep_batch_async(0, tab, a0, m_nblk_async[ph][0], m_s)) ||
       (rc = launch_sweep_batch(0, tab + a0, n - a0, m_nblk[ph][0], m_s)) || (rc = join()) || (rc = fork()))
      return rc;
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      e->m_stream = m_side[q % NSIDE];
      const GridSetup& g = e->m_opt.grid;
      double* um = m_base[3 * q + ph % 3];
      double* u = m_base[3 * q + (ph + 1) % 3];
      const double dt = e->mDt;
      hipStream_t st = e->m_stream;
      const int zf = e->zfill();
      double* acc = m_useacc[ph * n + q] ? m_acc[q] : nullptr;
      if ((rc = launch_src_apply(e->m_box, e->sources(), 0.0, 0, dt * dt, e->mRho.dev, e->mUp.dev, st, m_clk_t + q)) ||
          (rc = e->enforceBC(e->mUp, e->mMu, e->mLambda, 0.0, zf)))
         return rc;
      if (acc) {
         // the acceleration outside the sweep's planes and frame, and at the
         // source points, from the final Up (as EW::step_fused)
         int ka = 0, kz = 0;
         fused_planes(e->m_box, g.nz, e->m_onesided, &ka, &kz);
         if ((rc = launch_accel_fix(e->m_box, ka, kz, (zf & 15) != 15, true, e->mUp.dev, u, um, dt, acc, st)) ||
             (rc = launch_accel_points(e->m_box, e->sources(), e->mUp.dev, u, um, dt, acc, st)))
            return rc;
      }
   }
   if ((rc = join()) || (rc = fork())) return rc;
   for (int q = 0; q < n; q++)
      if ((rc = corr_tail(q))) return rc;
   if ((rc = launch_sweep_batch_async(1, tab + n, a1, m_nblk_async[ph][1], m_s)) ||
       (rc = launch_sweep_batch(1, tab + n + a1, n - a1, m_nblk[ph][1], m_s)) ||
       (rc = launch_sweep_batch_async(2, tab + 2 * n, a2, m_nblk_async[ph][2], m_s)) ||
       (rc = launch_sweep_batch(2, tab + 2 * n + a2, n - a2, m_nblk[ph][2], m_s)) || (rc = join()) || (rc = fork()))
      return rc;
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      hipStream_t st = m_side[q % NSIDE];
      e->m_stream = st;
      double* um = m_base[3 * q + ph % 3];
      double* up2 = m_base[3 * q + (ph + 2) % 3];
      const double dt = e->mDt;
      Sarray v = e->mU;
      v.dev = up2;
      const int zf = e->zfill();
      if ((rc = launch_src_apply(e->m_box, e->sources(), 0.0, 1, dt * dt * dt * dt / 12, e->mRho.dev, up2, st,
                                 m_clk_t + q)) ||
          (rc = e->enforceBC(v, e->mMu, e->mLambda, 0.0, zf)))
         return rc;
      const int nrec = (int)e->m_rec.size() / 4;
      if (nrec > 0 && e->m_drecdata &&
          (rc = launch_record(e->m_box, nrec, e->m_drec, up2, um, dt, e->m_drecdata, st, m_clk_s + q, e->m_maxsamples)))
         return rc;
   }
   if ((rc = join())) return rc;
   return launch_clock_tick(n, m_clk_t, m_clk_s, m_clk_dt, m_clk_max, m_s);
}

int Batch::run_period_graph()
{
   int rc;
   if (!m_exec) {
      hipGraph_t graph = nullptr;
      if ((rc = check_hip(hipStreamBeginCapture(m_s, hipStreamCaptureModeGlobal), "graph capture"))) return rc;
      for (int ph = 0; ph < 3 && rc == OK; ph++) rc = one_step(ph);
      const int re = check_hip(hipStreamEndCapture(m_s, &graph), "graph capture");
      if (rc || re) {
         if (graph) (void)hipGraphDestroy(graph);
         return rc ? rc : re;
      }
      rc = check_hip(hipGraphInstantiate(&m_exec, graph, nullptr, nullptr, 0), "graph instantiate");
      (void)hipGraphDestroy(graph);
      if (rc) return rc;
   }
   return check_hip(hipGraphLaunch(m_exec, m_s), "graph launch");
}

int Batch::sync_clocks(int nsteps)
{
   const int n = (int)m_shots.size();
   std::vector<double> t(n);
   std::vector<int> smp(n);
   int rc;
   if ((rc = check_hip(hipStreamSynchronize(m_s), "batch synchronize")) ||
       (rc = check_hip(hipMemcpy(t.data(), m_clk_t, sizeof(double) * n, hipMemcpyDeviceToHost), "clock readback")) ||
       (rc = check_hip(hipMemcpy(smp.data(), m_clk_s, sizeof(int) * n, hipMemcpyDeviceToHost), "clock readback")))
      return rc;
   for (int q = 0; q < n; q++) {
      EW* e = m_shots[q];
      e->mT = t[q];
      if (!e->m_rec.empty() && e->m_drecdata) e->m_nsamples = smp[q];
      e->mCycle += nsteps;
      e->mUm.dev = m_base[3 * q + m_phase % 3];
      e->mU.dev = m_base[3 * q + (m_phase + 1) % 3];
      e->mUp2.dev = m_base[3 * q + (m_phase + 2) % 3];
   }
   return OK;
}

int Batch::step(int nsteps, bool graph)
{
   if (nsteps < 0) return set_error(E_INVAL, "batch: negative step count");
   const int n = (int)m_shots.size();
   int rc = OK;
   // shots advanced outside the batch since the last call: start afresh
   for (int q = 0; q < n && rc == OK; q++) {
      EW* e = m_shots[q];
      if (e->m_comp) return set_error(E_STATE, "batch: a shot holds an FP32 state; write it back first");
      if (e->mUm.dev != m_base[3 * q + m_phase % 3] || e->mU.dev != m_base[3 * q + (m_phase + 1) % 3] ||
          e->mUp2.dev != m_base[3 * q + (m_phase + 2) % 3])
         rc = refresh();
   }
   if (rc) return rc;
   {  // the clocks follow the shots' host state (t, samples) at entry
      std::vector<double> t(n);
      std::vector<int> smp(n);
      for (int q = 0; q < n; q++) {
         t[q] = m_shots[q]->mT;
         smp[q] = m_shots[q]->m_nsamples;
      }
      if ((rc = check_hip(hipMemcpyAsync(m_clk_t, t.data(), sizeof(double) * n, hipMemcpyHostToDevice, m_s), "clock")) ||
          (rc = check_hip(hipMemcpyAsync(m_clk_s, smp.data(), sizeof(int) * n, hipMemcpyHostToDevice, m_s), "clock")) ||
          (rc = check_hip(hipStreamSynchronize(m_s), "clock")))
         return rc;
   }
   m_saved.clear();
   for (EW* e : m_shots) {
      m_saved.push_back(e->m_stream);
      e->m_stream = m_s;
   }
   int left = nsteps;
   while (rc == OK && left > 0) {
      if (graph && m_phase == 0 && left >= 3) {
         rc = run_period_graph();
         left -= 3;
         continue;
      }
      rc = one_step(m_phase);
      m_phase = (m_phase + 1) % 3;
      left--;
   }
   const int rs = sync_clocks(nsteps - left);
   for (int q = 0; q < n; q++) m_shots[q]->m_stream = m_saved[q];
   return rc ? rc : rs;
}

int Batch::tune(int* best, double* ms, int ncand, const int* cand)
{
   static const int defaults[7] = {16, 24, 32, 40, 48, 64, 96};
   if (ncand <= 0 || !cand) {
      ncand = 7;
      cand = defaults;
   }
   hipEvent_t e0, e1;
   int rc;
   if ((rc = check_hip(hipEventCreate(&e0), "event")) || (rc = check_hip(hipEventCreate(&e1), "event"))) return rc;
   double tbest = 1e300;
   int kbest = m_kchunk;
   for (int c = 0; c < ncand && rc == OK; c++) {
      if ((rc = set_kchunk(cand[c])) || (rc = step(1, false))) break;  // warm-up
      if ((rc = check_hip(hipEventRecord(e0, m_s), "event"))) break;
      rc = step(2, false);
      if (rc || (rc = check_hip(hipEventRecord(e1, m_s), "event")) ||
          (rc = check_hip(hipEventSynchronize(e1), "event")))
         break;
      float t = 0;
      if ((rc = check_hip(hipEventElapsedTime(&t, e0, e1), "event"))) break;
      if (ms) ms[c] = t / 2;
      if (t < tbest) {
         tbest = t;
         kbest = cand[c];
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   if (rc) return rc;
   for (EW* e : m_shots)
      if ((rc = e->reset())) return rc;
   if (best) *best = kbest;
   m_kchunk = kbest;
   return refresh();
}

int Batch::gradient(int nsteps, int ckpt, const double* const* gtr, double* const* grho, double* const* gmu,
                    double* const* gla, bool on_device, int max_concurrent, bool fp32)
{
   const int n = (int)m_shots.size();
   if (!gtr || !grho || !gmu || !gla) return set_error(E_INVAL, "batch gradient: null argument");
   if (max_concurrent < 1) max_concurrent = 1;
   int dev = 0;
   int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
   if (rc) return rc;
   // the batch's own work must be done before the shots' streams take over
   if ((rc = check_hip(hipStreamSynchronize(m_s), "batch gradient"))) return rc;
   // every shot keeps ckpt + 2 segment states and two per checkpoint: with no
   // interval given, about sqrt(nsteps), which keeps that near its least
   if (ckpt < 1) ckpt = std::max(1, (int)std::lround(std::sqrt((double)nsteps)));
   // The shots do not keep their segments' predicted fields (EW::gradient,
   // store), measured slower here (48 shots of 128^3..288^3, 100 steps,
   // four at a time: 25.5 s without, 46.9 s with them): each shot's work
   // memory, allocated and freed around its gradient, is three times as
   // large with them.
   std::vector<double> need(n, 0.0);
   for (int q = 0; q < n; q++)
      if ((rc = m_shots[q]->gradient_plan(nsteps, ckpt, nullptr, &need[q], fp32, 0))) return rc;
   size_t freeb = 0, totalb = 0;
   if ((rc = check_hip(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo"))) return rc;
   // the pool's buffers (from an earlier call) are reused
   const double budget = (double)freeb + m_wpool.held() - 4.0 * (1 << 30) - 0.02 * (double)totalb;
   // largest first: the buffers of the first shots then serve the later ones
   std::vector<int> order(n);
   for (int q = 0; q < n; q++) order[q] = q;
   std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return need[a] > need[b]; });
   std::mutex mtx;
   std::condition_variable cv;
   double inflight = 0.0;
   int running = 0, next = 0, first_rc = OK;
   std::string first_msg;
   auto worker = [&]() {
      (void)hipSetDevice(dev);
      hipStream_t st = nullptr;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
         std::lock_guard<std::mutex> lk(mtx);
         if (first_rc == OK) {
            first_rc = E_HIP;
            first_msg = "batch gradient: stream creation failed";
         }
         return;
      }
      for (;;) {
         int q;
         {
            std::unique_lock<std::mutex> lk(mtx);
            // start the next shot when it fits beside the running ones (or runs alone)
            cv.wait(lk, [&]() {
               return next >= n || first_rc != OK || running == 0 || inflight + need[order[next]] <= budget;
            });
            if (next >= n || first_rc != OK) break;
            q = order[next++];
            running++;
            inflight += need[q];
         }
         EW* e = m_shots[q];
         const hipStream_t saved = e->m_stream;
         e->m_stream = st;
         e->m_shared = &m_wpool;
         const int r = e->gradient(nsteps, ckpt, gtr[q], grho[q], gmu[q], gla[q], on_device, fp32, 0);
         e->m_shared = nullptr;
         e->m_stream = saved;
         {
            std::lock_guard<std::mutex> lk(mtx);
            running--;
            inflight -= need[q];
            if (r != OK && first_rc == OK) {
               first_rc = r;
               first_msg = "batch gradient, shot " + std::to_string(q) + ": " + sw4hip_last_error();
            }
         }
         cv.notify_all();
      }
      (void)hipStreamDestroy(st);
      cv.notify_all();
   };
   const int nthreads = std::min(max_concurrent, n);
   std::vector<std::thread> pool;
   for (int w = 0; w < nthreads; w++) pool.emplace_back(worker);
   for (std::thread& t : pool) t.join();
   if (first_rc != OK) return set_error(first_rc, first_msg);
   return OK;
}

int Batch::release_workspace()
{
   m_wpool.release();
   return OK;
}

}  // namespace sw4amd
