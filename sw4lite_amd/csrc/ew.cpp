// ew.cpp -- the EW facade: set-up of one Cartesian grid and the time-step
// loop on the GPU.  Stage order and formulas follow EW::timesteploop
// (EW.C:2339-2928) and the EW step methods (EW.C:3085-3636).
#include "ew.h"
#include "kernels_fast.h"

#include <cmath>
#include <cstring>

namespace sw4amd {

int launch_fill(double* a, long n, double v, hipStream_t s);

EW::EW(const EWOptions& opt) : m_opt(opt)
{
   const GridSetup& g = opt.grid;
   m_box = make_box(-1, g.nx + 2, -1, g.ny + 2, -1, g.nz + 2);
   for (int q = 0; q < 36; q++) m_wind[q] = 0;
   for (int q = 0; q < 6; q++) m_nbcpts[q] = 0;
}

EW::~EW()
{
   for (void* p : m_allocs) (void)hipFree(p);
   for (const PoolBuf& b : m_pool) (void)hipFree(b.p);
}

int EW::alloc(Sarray& a, int nc)
{
   a.m_nc = nc;
   a.m_ib = m_box.ib; a.m_ie = m_box.ie;
   a.m_jb = m_box.jb; a.m_je = m_box.je;
   a.m_kb = m_box.kb; a.m_ke = m_box.ke;
   a.m_corder = m_opt.corder != 0;
   void* p = nullptr;
   if (hipMalloc(&p, sizeof(double) * a.size()) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(E_NOMEM, "EW: device allocation of " + std::to_string(sizeof(double) * a.size()) + " bytes failed");
   }
   m_allocs.push_back(p);
   a.dev = static_cast<double*>(p);
   return launch_fill(a.dev, a.size(), 0.0, m_stream);
}

int EW::init()
{
   int rc;
   const GridSetup& g = m_opt.grid;
   if (g.nx < 12 || g.ny < 12 || g.nz < 12 || !(g.h > 0))
      return set_error(E_INVAL, "EW: need nx,ny,nz >= 12 and h > 0");
   if ((rc = alloc(mU, 3)) || (rc = alloc(mUm, 3)) || (rc = alloc(mUp, 3)) || (rc = alloc(mUp2, 3)))
      return rc;
   if ((rc = alloc(mRho, 1)) || (rc = alloc(mMu, 1)) || (rc = alloc(mLambda, 1))) return rc;
   void* p = nullptr;
   if (hipMalloc(&p, sizeof(SbpTables)) != hipSuccess) return set_error(E_NOMEM, "EW: SBP table allocation failed");
   m_allocs.push_back(p);
   m_dsbp = static_cast<SbpTables*>(p);
   SbpTables t;
   fill_stencil_coefficients(t.acof, t.ghcof, t.bope, t.sbop);
   if ((rc = check_hip(hipMemcpy(m_dsbp, &t, sizeof(t), hipMemcpyHostToDevice), "SBP upload"))) return rc;
   const long nsg = 3 * (m_box.ni + (m_box.je - m_box.jb + 1) + (m_box.ke - m_box.kb + 1));
   if (hipMalloc(&p, sizeof(double) * nsg) != hipSuccess) return set_error(E_NOMEM, "EW: supergrid allocation failed");
   m_allocs.push_back(p);
   m_dsg = static_cast<double*>(p);
   if (hipMalloc(&p, sizeof(double) * 3 * m_nred + 64) != hipSuccess) return set_error(E_NOMEM, "EW: reduction buffer allocation failed");
   m_allocs.push_back(p);
   m_dred = static_cast<double*>(p);
   if (hipMalloc(&p, sizeof(double) * accel_slab_size(m_box)) != hipSuccess)
      return set_error(E_NOMEM, "EW: acceleration slab allocation failed");
   m_allocs.push_back(p);
   m_dacc = static_cast<double*>(p);
   return OK;
}

// The whole acceleration for kernel variant 8 (3 npts doubles, allocated at
// the first step that uses it); null for the other variants and for odd row
// lengths, whose sweeps cannot stage it.
int EW::accel_field(double** p)
{
   *p = nullptr;
   if (m_variant < 8 || (m_box.ni % 2 != 0 && m_variant < 10)) return OK;
   if (!m_accf) {
      void* q = nullptr;
      if (hipMalloc(&q, sizeof(double) * 3 * m_box.npts) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "EW: acceleration field allocation failed");
      }
      m_allocs.push_back(q);
      m_accf = static_cast<double*>(q);
   }
   *p = m_accf;
   return OK;
}

int EW::damp_tile_list(int k1, int k2, const DampTile** tiles, int* n)
{
   *tiles = nullptr;
   *n = 0;
   if (m_variant < 9 || k2 < k1) return OK;
   if (!m_ddamp || m_damp_k1 != k1 || m_damp_k2 != k2 || m_damp_kc != m_kchunk_damp) {
      const std::vector<DampTile> v = damp_tiles(m_box, k1, k2, m_kchunk_damp, m_sg_host.dcx.data(),
                                                 m_sg_host.dcy.data(), m_sg_host.dcz.data());
      // at least one element, so that an empty list is a valid pointer
      const size_t bytes = sizeof(DampTile) * std::max<size_t>(v.size(), 1);
      void* q = nullptr;
      if (hipMalloc(&q, bytes) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "EW: damping tile list allocation failed");
      }
      m_allocs.push_back(q);
      if (!v.empty()) {
         const int rc = check_hip(hipMemcpy(q, v.data(), sizeof(DampTile) * v.size(), hipMemcpyHostToDevice),
                                  "damping tile list upload");
         if (rc) return rc;
      }
      m_ddamp = static_cast<DampTile*>(q);
      m_ndamp = (int)v.size();
      m_damp_k1 = k1;
      m_damp_k2 = k2;
      m_damp_kc = m_kchunk_damp;
   }
   *tiles = m_ddamp;
   *n = m_ndamp;
   return OK;
}

int EW::damp_T_tile_list(const DampTile** tiles, int* n)
{
   *tiles = nullptr;
   *n = 0;
   if (m_variant < 9) return OK;
   if (!m_ddampT) {
      const std::vector<DampTile> v =
          damp_T_tiles(m_box, m_sg_host.dcx.data(), m_sg_host.dcy.data(), m_sg_host.dcz.data());
      void* q = nullptr;
      if (hipMalloc(&q, sizeof(DampTile) * std::max<size_t>(v.size(), 1)) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "EW: damping transpose tile list allocation failed");
      }
      m_allocs.push_back(q);
      if (!v.empty()) {
         const int rc = check_hip(hipMemcpy(q, v.data(), sizeof(DampTile) * v.size(), hipMemcpyHostToDevice),
                                  "damping transpose tile list upload");
         if (rc) return rc;
      }
      m_ddampT = static_cast<DampTile*>(q);
      m_ndampT = (int)v.size();
   }
   *tiles = m_ddampT;
   *n = m_ndampT;
   return OK;
}

int EW::set_material(const double* rho, const double* mu, const double* la, bool on_device)
{
   const size_t bytes = sizeof(double) * m_box.npts;
   const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
   int rc;
   if ((rc = check_hip(hipMemcpyAsync(mRho.dev, rho, bytes, kind, m_stream), "rho copy")) ||
       (rc = check_hip(hipMemcpyAsync(mMu.dev, mu, bytes, kind, m_stream), "mu copy")) ||
       (rc = check_hip(hipMemcpyAsync(mLambda.dev, la, bytes, kind, m_stream), "lambda copy")))
      return rc;
   if (!on_device && (rc = check_hip(hipStreamSynchronize(m_stream), "material copy"))) return rc;
   m_material_set = true;
   return OK;
}

// EW::processTestPointSource (EW.C:893-950): mu = rho cs^2, la = rho cp^2 - 2 mu
int EW::set_material_const(double rho, double cs, double cp)
{
   const double mu = rho * cs * cs, la = rho * cp * cp - 2 * mu;
   int rc;
   if ((rc = launch_fill(mRho.dev, m_box.npts, rho, m_stream)) || (rc = launch_fill(mMu.dev, m_box.npts, mu, m_stream)) ||
       (rc = launch_fill(mLambda.dev, m_box.npts, la, m_stream)))
      return rc;
   m_material_set = true;
   return OK;
}

int EW::add_moment_source(const MomentSource& s)
{
   if (m_setup_done) return set_error(E_STATE, "EW: sources must be added before setup");
   if (s.tfkind < TF_RICKER || s.tfkind > TF_C6SMOOTHBUMP) return set_error(E_INVAL, "EW: unknown time function");
   m_sources.push_back(s);
   return OK;
}

int EW::set_topography(const double* z, bool on_device)
{
   if (m_setup_done) return set_error(E_STATE, "EW: topography must be set before setup");
   if (!z) return set_error(E_INVAL, "EW: null topography");
   const GridSetup& g = m_opt.grid;
   for (int s = 0; s < 4; s++)
      if (g.bc[s] == BC_PERIODIC) return set_error(E_INVAL, "EW: periodic sides are not supported on a curvilinear grid");
   if (g.bc[5] == BC_FREE) return set_error(E_INVAL, "EW: the curvilinear grid has no free surface at the bottom");
   const long n = m_box.npts;
   int rc;
   m_zhost.resize(n);
   if ((rc = check_hip(hipMemcpy(m_zhost.data(), z, sizeof(double) * n,
                                 on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost), "topography copy")))
      return rc;
   if (!mMetric.dev && ((rc = alloc(mMetric, 4)) || (rc = alloc(mJ, 1)))) return rc;
   // the metric kernel reads z from the device: stage it in the Lu array
   if (!mLu.dev && (rc = alloc(mLu, 3))) return rc;
   if ((rc = check_hip(hipMemcpyAsync(mLu.dev, m_zhost.data(), sizeof(double) * n, hipMemcpyHostToDevice, m_stream),
                       "topography upload")))
      return rc;
   int* flag = reinterpret_cast<int*>(m_dred);
   if ((rc = check_hip(hipMemsetAsync(flag, 0, sizeof(int), m_stream), "metric flag")) ||
       (rc = launch_metric(m_opt.corder, m_box, g.h, mLu.dev, mMetric.dev, mJ.dev, flag, m_stream)))
      return rc;
   int bad = 0;
   if ((rc = check_hip(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, m_stream), "metric flag")) ||
       (rc = check_hip(hipStreamSynchronize(m_stream), "metric flag")))
      return rc;
   if (bad) return set_error(E_INVAL, "EW: topography grid is folded (z not increasing in k)");
   // Lu is read at every point by the pointwise updates: clear the staging
   if ((rc = launch_fill(mLu.dev, mLu.size(), 0.0, m_stream))) return rc;
   m_curv = true;
   return OK;
}

SgCoef EW::sg() const
{
   const long ni = m_box.ni, nj = m_box.je - m_box.jb + 1, nk = m_box.ke - m_box.kb + 1;
   const double* d = m_dsg;
   SgCoef c;
   c.dcx = d; c.strx = d + ni; c.cox = d + 2 * ni;
   d += 3 * ni;
   c.dcy = d; c.stry = d + nj; c.coy = d + 2 * nj;
   d += 3 * nj;
   c.dcz = d; c.strz = d + nk; c.coz = d + 2 * nk;
   return c;
}

int EW::upload_sources()
{
   // packed: first[nu+1] | ijk[3n] | tfkind[n] | force[3n] | freq[n] | t0[n]
   const int n = (int)m_gps.size(), nu = (int)m_first.size() - 1;
   const size_t ibytes = sizeof(int) * (nu + 1 + 4 * n);
   const size_t ioff = (ibytes + 15) / 16 * 16;
   const size_t total = ioff + sizeof(double) * 5 * n + 16;
   std::vector<char> host(total, 0);
   int* ip = reinterpret_cast<int*>(host.data());
   std::memcpy(ip, m_first.data(), sizeof(int) * (nu + 1));
   int* ijk = ip + nu + 1;
   int* tf = ijk + 3 * n;
   double* dp = reinterpret_cast<double*>(host.data() + ioff);
   double *force = dp, *freq = dp + 3 * n, *t0 = dp + 4 * n;
   for (int s = 0; s < n; s++) {
      ijk[3 * s] = m_gps[s].i; ijk[3 * s + 1] = m_gps[s].j; ijk[3 * s + 2] = m_gps[s].k;
      tf[s] = m_gps[s].tfkind;
      for (int c = 0; c < 3; c++) force[3 * s + c] = m_gps[s].f[c];
      freq[s] = m_gps[s].freq;
      t0[s] = m_gps[s].t0;
   }
   void* p = nullptr;
   if (hipMalloc(&p, total) != hipSuccess) return set_error(E_NOMEM, "EW: source table allocation failed");
   m_allocs.push_back(p);
   m_dsrc = p;
   return check_hip(hipMemcpy(p, host.data(), total, hipMemcpyHostToDevice), "source upload");
}

SourceTable EW::sources() const
{
   SourceTable st;
   const int n = (int)m_gps.size(), nu = (int)m_first.size() - 1;
   st.nunique = m_dsrc ? nu : 0;
   const size_t ibytes = sizeof(int) * (nu + 1 + 4 * n);
   const size_t ioff = (ibytes + 15) / 16 * 16;
   const int* ip = static_cast<const int*>(m_dsrc);
   const double* dp = reinterpret_cast<const double*>(static_cast<const char*>(m_dsrc) + ioff);
   st.first = ip;
   st.ijk = ip + nu + 1;
   st.tfkind = ip + nu + 1 + 3 * n;
   st.force = dp;
   st.freq = dp + 3 * n;
   st.t0 = dp + 4 * n;
   return st;
}

int EW::stable_dt(double* dt)
{
   if (!m_material_set) return set_error(E_STATE, "EW: set the material before computing dt");
   // EW::computeDT (EW.C:5041-5146): Cartesian grid, or the curvilinear
   // grid's principal-part eigenvalues (EW.C:5063-5136, dt = CFL/sqrt(max))
   int rc;
   if (m_curv)
      rc = launch_maxeig_curv(m_opt.corder, m_box, mMetric.dev, mJ.dev, mMu.dev, mLambda.dev, mRho.dev, m_dred, m_nred,
                              m_dred + 3 * m_nred, m_stream);
   else
      rc = launch_maxeig(m_box.npts, mMu.dev, mLambda.dev, mRho.dev, m_dred, m_nred, m_dred + 3 * m_nred, m_stream);
   if (rc) return rc;
   double eigmax = 0;
   if ((rc = check_hip(hipMemcpyAsync(&eigmax, m_dred + 3 * m_nred, sizeof(double), hipMemcpyDeviceToHost, m_stream), "dt readback")) ||
       (rc = check_hip(hipStreamSynchronize(m_stream), "dt readback")))
      return rc;
   if (!(eigmax > 0)) return set_error(E_INVAL, "EW: material gives a non-positive (4mu+lambda)/rho");
   const double ieigmax = 1 / std::sqrt(eigmax);
   *dt = std::min(1e10, m_opt.cfl * (m_curv ? 1.0 : m_opt.grid.h) * ieigmax);
   return OK;
}

int EW::add_grid_point_source(const GridPointSource& p)
{
   if (m_setup_done) return set_error(E_STATE, "EW: sources must be added before setup");
   const GridSetup& g = m_opt.grid;
   if (p.i < -1 || p.i > g.nx + 2 || p.j < -1 || p.j > g.ny + 2 || p.k < 1 || p.k > g.nz)
      return set_error(E_INVAL, "EW: grid-point source outside the grid");
   if (p.tfkind < TF_RICKER || p.tfkind > TF_C6SMOOTHBUMP) return set_error(E_INVAL, "EW: unknown time function");
   m_gps_extra.push_back(p);
   return OK;
}

int EW::setup(double tmax, int nsteps, double dt0)
{
   if (!m_material_set) return set_error(E_STATE, "EW: set the material before setup");
   if (m_setup_done) return set_error(E_STATE, "EW: setup called twice");
   const GridSetup& g = m_opt.grid;
   int rc;
   // one-sided SBP closures where the top/bottom is a free surface
   // (EW::assign_local_bcs, EW.C:4863-4867)
   for (int s = 0; s < 6; s++) m_onesided[s] = 0;
   m_onesided[4] = g.bc[4] == BC_FREE;
   m_onesided[5] = g.bc[5] == BC_FREE;
   for (int s = 0; s < 4; s++)
      if (g.bc[s] == BC_FREE) return set_error(E_INVAL, "EW: free surface is only supported on sides 4 and 5");
   // supergrid arrays
   make_supergrid_arrays(g, m_sg_host);
   {
      std::vector<double> all;
      for (auto* v : {&m_sg_host.dcx, &m_sg_host.strx, &m_sg_host.cox, &m_sg_host.dcy, &m_sg_host.stry,
                      &m_sg_host.coy, &m_sg_host.dcz, &m_sg_host.strz, &m_sg_host.coz})
         all.insert(all.end(), v->begin(), v->end());
      if ((rc = check_hip(hipMemcpy(m_dsg, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice), "supergrid upload")))
         return rc;
   }
   make_boundary_windows(g, m_wind, m_nbcpts);
   if (dt0 > 0)
      mDt = dt0;
   else if ((rc = stable_dt(&mDt)))
      return rc;
   if (tmax > 0) {
      mNumberOfTimeSteps = static_cast<int>(tmax / mDt + 0.5);
      if (mNumberOfTimeSteps == 0) mNumberOfTimeSteps = 1;
      mDt = tmax / mNumberOfTimeSteps;
   } else {
      mNumberOfTimeSteps = nsteps;
   }
   // grid-point sources
   m_gps.clear();
   for (const MomentSource& s : m_sources) {
      const int q = m_curv ? moment_source_points_curv(g, m_zhost.data(), s, m_gps) : moment_source_points(g, s, m_gps);
      if (q == -1) return set_error(E_INVAL, "EW: grid too small for the source stencil");
      if (q == -2) return set_error(E_INVAL, "EW: source outside the curvilinear grid");
      if (q) return set_error(E_INVAL, "EW: folded grid at the source");
   }
   m_gps.insert(m_gps.end(), m_gps_extra.begin(), m_gps_extra.end());
   sort_grid_point_sources(g, m_gps, m_first);
   if (!m_gps.empty() && (rc = upload_sources())) return rc;
   mT = 0;
   mCycle = 0;
   m_setup_done = true;
   // receivers: one sample for the initial data and one per step
   // (EW.C:2463-2480 records U at t = 0)
   const int nrec = (int)m_rec.size() / 4;
   if (nrec > 0) {
      m_maxsamples = std::max(mNumberOfTimeSteps, 0) + 1;
      void* p = nullptr;
      if (hipMalloc(&p, sizeof(int) * m_rec.size()) != hipSuccess) return set_error(E_NOMEM, "EW: receiver table");
      m_allocs.push_back(p);
      m_drec = static_cast<int*>(p);
      if ((rc = check_hip(hipMemcpy(m_drec, m_rec.data(), sizeof(int) * m_rec.size(), hipMemcpyHostToDevice),
                          "receiver upload")))
         return rc;
      if (hipMalloc(&p, sizeof(double) * 3 * nrec * (size_t)m_maxsamples) != hipSuccess)
         return set_error(E_NOMEM, "EW: receiver buffer");
      m_allocs.push_back(p);
      m_drecdata = static_cast<double*>(p);
      m_nsamples = 0;
      if ((rc = record(mU, mUm))) return rc;
   }
   return OK;
}

int EW::host_solution(double* u, double* um)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: solution before setup");
   int rc;
   if ((rc = comp_end())) return rc;
   const long npts = mU.npts();
   std::vector<double> tmp;
   // component-blocked copies whatever the device layout
   for (int w = 0; w < 2; w++) {
      double* dst = w == 0 ? u : um;
      const Sarray& a = w == 0 ? mU : mUm;
      if (!dst) continue;
      double* buf = dst;
      if (!m_opt.corder) {
         tmp.resize(3 * npts);
         buf = tmp.data();
      }
      if ((rc = check_hip(hipMemcpyAsync(buf, a.dev, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, m_stream),
                          "solution readback")) ||
          (rc = check_hip(hipStreamSynchronize(m_stream), "solution readback")))
         return rc;
      if (!m_opt.corder)
         for (long p = 0; p < npts; p++)
            for (int c = 0; c < 3; c++) dst[p + npts * c] = tmp[3 * p + c];
   }
   return OK;
}

int EW::restart_from(const double* u, const double* um, double t, int cycle)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: restart before setup");
   if (cycle < 0 || cycle > mNumberOfTimeSteps)
      return set_error(E_INVAL, "EW: restart cycle outside the run's " + std::to_string(mNumberOfTimeSteps) + " steps");
   int rc;
   if ((rc = comp_end())) return rc;
   m_comp = false;
   const long npts = mU.npts();
   std::vector<double> tmp;
   for (int w = 0; w < 2; w++) {
      const double* src = w == 0 ? u : um;
      Sarray& a = w == 0 ? mU : mUm;
      if (!m_opt.corder) {
         tmp.resize(3 * npts);
         for (long p = 0; p < npts; p++)
            for (int c = 0; c < 3; c++) tmp[3 * p + c] = src[p + npts * c];
         src = tmp.data();
      }
      if ((rc = check_hip(hipMemcpyAsync(a.dev, src, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, m_stream),
                          "restart upload")) ||
          (rc = check_hip(hipStreamSynchronize(m_stream), "restart upload")))
         return rc;
   }
   if ((rc = enforceBC(mU, mMu, mLambda, t)) || (rc = enforceBC(mUm, mMu, mLambda, t - mDt))) return rc;
   mT = t;
   mCycle = cycle;
   m_nsamples = 0;
   return record(mU, mUm);
}

int EW::add_receiver(double x, double y, double z, int mode, int* index)
{
   if (m_setup_done) return set_error(E_STATE, "EW: receivers must be added before setup");
   if (mode != 0 && mode != 1) return set_error(E_INVAL, "EW: receiver mode must be 0 (displacement) or 1 (velocity)");
   const GridSetup& g = m_opt.grid;
   // EW::computeNearestGridPoint (EW.C:5149-5180), one Cartesian grid
   int i = (int)std::floor(x / g.h) + 1;
   if (x - (i - 0.5) * g.h > 0.) i++;
   int j = (int)std::floor(y / g.h) + 1;
   if (y - (j - 0.5) * g.h > 0.) j++;
   int k = (int)std::floor((z - g.zmin) / g.h) + 1;
   if (z - (g.zmin + (k - 0.5) * g.h) > 0.) k++;
   if (m_curv && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
      // nearest grid point along the column (i, j) of the curvilinear grid
      const long ni = g.nx + 4, nij = ni * (g.ny + 4);
      const double* col = m_zhost.data() + (i + 1) + ni * (j + 1);
      k = 1;
      for (int kk = 2; kk <= g.nz; kk++)
         if (std::fabs(col[nij * (kk + 1)] - z) < std::fabs(col[nij * (k + 1)] - z)) k = kk;
   }
   if (i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz)
      return set_error(E_INVAL, "EW: receiver outside the grid");
   if (index) *index = (int)m_rec.size() / 4;
   m_rec.insert(m_rec.end(), {i, j, k, mode});
   return OK;
}

int EW::add_receiver_ijk(int i, int j, int k, int mode, int* index)
{
   if (m_setup_done) return set_error(E_STATE, "EW: receivers must be added before setup");
   if (mode != 0 && mode != 1) return set_error(E_INVAL, "EW: receiver mode must be 0 (displacement) or 1 (velocity)");
   const GridSetup& g = m_opt.grid;
   if (i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz)
      return set_error(E_INVAL, "EW: receiver outside the grid");
   if (index) *index = (int)m_rec.size() / 4;
   m_rec.insert(m_rec.end(), {i, j, k, mode});
   return OK;
}

int EW::record(const Sarray& Up, const Sarray& Um)
{
   const int nrec = (int)m_rec.size() / 4;
   if (nrec == 0 || !m_drecdata) return OK;
   if (m_nsamples >= m_maxsamples) return OK;  // steps beyond the set-up count are not recorded
   const int rc = launch_record(m_box, nrec, m_drec, Up.dev, Um.dev, mDt,
                                m_drecdata + (size_t)3 * nrec * m_nsamples, m_stream);
   if (rc == OK) m_nsamples++;
   return rc;
}

int EW::get_receivers(int* nsamples, double* out)
{
   const int nrec = (int)m_rec.size() / 4;
   if (nsamples) *nsamples = m_nsamples;
   if (!out || nrec == 0 || m_nsamples == 0) return OK;
   int rc;
   if ((rc = check_hip(hipMemcpyAsync(out, m_drecdata, sizeof(double) * 3 * nrec * (size_t)m_nsamples,
                                      hipMemcpyDeviceToHost, m_stream), "receiver readback")))
      return rc;
   return check_hip(hipStreamSynchronize(m_stream), "receiver readback");
}

// ---------------------------------------------------------------------------
// reference step methods
// ---------------------------------------------------------------------------
int EW::Force(double t, Sarray& F, bool tt)
{
   return launch_force(m_opt.corder, m_box, F.dev, sources(), t, tt ? 1 : 0, true, m_stream);
}

int EW::evalRHS(Sarray& U, Sarray& Mu, Sarray& Lambda, Sarray& Lu)
{
   const SgCoef c = sg();
   if (m_curv)  // EW.C:3218-3243
      return launch_rhs4sgcurv(m_opt.corder, m_box, m_onesided, m_dsbp, U.dev, Mu.dev, Lambda.dev, mMetric.dev, mJ.dev,
                               c.strx, c.stry, Lu.dev, m_stream);
   return launch_rhs4sg(m_opt.corder, m_box, m_opt.grid.nz, m_onesided, m_dsbp, Lu.dev, U.dev, Mu.dev,
                        Lambda.dev, m_opt.grid.h, c.strx, c.stry, c.strz, m_stream);
}

int EW::evalPredictor(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Rho, Sarray& Lu, Sarray& F)
{
   return launch_pred(m_opt.corder, m_box, Up.dev, U.dev, Um.dev, Lu.dev, F.dev, Rho.dev, mDt * mDt, m_stream);
}

int EW::evalCorrector(Sarray& Up, Sarray& Rho, Sarray& Lu, Sarray& F)
{
   return launch_corr(m_opt.corder, m_box, Up.dev, Lu.dev, F.dev, Rho.dev, mDt * mDt * mDt * mDt, m_stream);
}

int EW::evalDpDmInTime(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Uacc)
{
   return launch_dpdmt(m_box, Up.dev, U.dev, Um.dev, Uacc.dev, 1. / (mDt * mDt), m_stream);
}

int EW::addSuperGridDamping(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Rho)
{
   if (m_curv)  // no damping along k on the curvilinear grid (EW.C:3560-3582, 4745)
      return launch_addsgd4c(m_opt.corder, m_box, Up.dev, U.dev, Um.dev, Rho.dev, mJ.dev, sg(), m_opt.beta, m_stream);
   return launch_addsgd4(m_opt.corder, m_box, Up.dev, U.dev, Um.dev, Rho.dev, sg(), m_opt.beta, m_stream);
}

// EW::enforceBC with cartesian_bc_forcing giving zero boundary data
// (EW.C:3320-3344, 3446-3501)
int EW::zfill() const
{
   if (m_curv) return 0;
   const GridSetup& g = m_opt.grid;
   int z = 0;
   for (int q = 0; q < 6; q++) {
      if (g.bc[q] == BC_PERIODIC) return 0;
      if (g.bc[q] == BC_DIRICHLET || g.bc[q] == BC_SUPERGRID) z |= 1 << q;
   }
   return z;
}

int EW::enforceBC(Sarray& U, Sarray& Mu, Sarray& Lambda, double t, int skip)
{
   (void)t;
   const GridSetup& g = m_opt.grid;
   const double* bforce[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
   const SgCoef c = sg();
   if (!m_curv) {
      int bc[6];
      for (int q = 0; q < 6; q++) bc[q] = (skip >> q) & 1 ? BC_NONE : g.bc[q];
      return launch_bcfortsg(m_opt.corder, m_box, m_wind, g.nx, g.ny, g.nz, U.dev, g.h, bc, m_dsbp, Mu.dev,
                             Lambda.dev, bforce, c.strx, c.stry, m_stream);
   }
   // curvilinear top: the Cartesian closure of bcfortsg is overwritten by
   // freesurfcurvisg at the same points (EW.C:3479-3497), so it is skipped
   int bc[6];
   for (int s = 0; s < 6; s++) bc[s] = g.bc[s] == BC_FREE ? BC_NONE : g.bc[s];
   int rc = launch_bcfortsg(m_opt.corder, m_box, m_wind, g.nx, g.ny, g.nz, U.dev, g.h, bc, m_dsbp, Mu.dev, Lambda.dev,
                            bforce, c.strx, c.stry, m_stream);
   if (rc || g.bc[4] != BC_FREE) return rc;
   return launch_freesurf_curv(m_opt.corder, m_box, g.nz, 4, U.dev, Mu.dev, Lambda.dev, mMetric.dev, m_dsbp, nullptr,
                               c.strx, c.stry, m_stream);
}

// EW::cycleSolutionArrays (EW.C:3060-3082): pointer rotation, no copy
void EW::cycleSolutionArrays()
{
   double* tmp = mUm.dev;
   mUm.dev = mU.dev;
   mU.dev = mUp.dev;
   mUp.dev = tmp;
}

int EW::ensure_reference_arrays()
{
   int rc;
   if (!mLu.dev && (rc = alloc(mLu, 3))) return rc;
   if (!mF.dev && (rc = alloc(mF, 3))) return rc;
   if (!mUacc.dev && (rc = alloc(mUacc, 3))) return rc;
   return OK;
}

int EW::step_reference(int n)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: step before setup");
   int rc;
   if ((rc = comp_end())) return rc;
   if ((rc = ensure_reference_arrays())) return rc;
   for (int s = 0; s < n; s++) {
      const double t = mT;
      if ((rc = Force(t, mF, false)) || (rc = evalRHS(mU, mMu, mLambda, mLu)) ||
          (rc = evalPredictor(mUp, mU, mUm, mRho, mLu, mF)) || (rc = enforceBC(mUp, mMu, mLambda, t + mDt)) ||
          (rc = Force(t, mF, true)) || (rc = evalDpDmInTime(mUp, mU, mUm, mUacc)) ||
          (rc = evalRHS(mUacc, mMu, mLambda, mLu)) || (rc = evalCorrector(mUp, mRho, mLu, mF)) ||
          (rc = addSuperGridDamping(mUp, mU, mUm, mRho)) || (rc = enforceBC(mUp, mMu, mLambda, t + mDt)) ||
          (rc = record(mUp, mUm)))
         return rc;
      mT += mDt;
      mCycle++;
      cycleSolutionArrays();
   }
   return OK;
}

int EW::step_fused(int n, double* stage_ms)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: step before setup");
   if (int rc = comp_end()) return rc;
   if (!m_opt.corder) return set_error(E_INVAL, "EW: the fused step needs the component-blocked layout (corder=1)");
   if (m_curv) return set_error(E_INVAL, "EW: the fused step is Cartesian; step a curvilinear grid in the reference order");
   const GridSetup& g = m_opt.grid;
   const SgCoef c = sg();
   const SourceTable st = sources();
   int rc = OK;
   // optional per-stage timing: one event before the step and after each stage
   std::vector<hipEvent_t> ev;
   if (stage_ms) {
      ev.resize((size_t)n * (kStages + 1));
      for (auto& e : ev)
         if ((rc = check_hip(hipEventCreate(&e), "hipEventCreate"))) return rc;
   }
   auto mark = [&](int s, int q) {
      return stage_ms ? check_hip(hipEventRecord(ev[(size_t)s * (kStages + 1) + q], m_stream), "hipEventRecord") : OK;
   };
   const int zf = zfill();
   int ka = 0, kz = 0;
   fused_planes(m_box, g.nz, m_onesided, &ka, &kz);
   double* accf = nullptr;  // variant 8: the acceleration stored by the predictor, staged by the corrector
   // (a grid whose SBP rows leave the sweep no plane has nothing to store it)
   if (rc == OK && kz >= ka) rc = accel_field(&accf);
   const DampTile* dtiles = nullptr;
   int ndtiles = 0;
   if (rc == OK) rc = damp_tile_list(ka, kz, &dtiles, &ndtiles);
   // completes the stored acceleration from the final Up (after sources and BC)
   auto accel_rest = [&](double dt) -> int {
      if (!accf) return OK;
      // the shell's values of the side frame are final when every side is zero-filled
      const bool frame = (zf & 15) != 15;
      // every plane outside the sweep: the corrector's SBP rows stage it too
      const int r = launch_accel_fix(m_box, ka, kz, frame, true, mUp.dev, mU.dev, mUm.dev, dt, accf, m_stream);
      return r ? r : launch_accel_points(m_box, st, mUp.dev, mU.dev, mUm.dev, dt, accf, m_stream);
   };
   for (int s = 0; s < n && rc == OK; s++) {
      const double t = mT, dt = mDt;
      if ((rc = mark(s, 0)) ||
          (rc = launch_pred_fused(m_box, g.nz, m_onesided, m_dsbp, mU.dev, mUm.dev, mMu.dev, mLambda.dev, mRho.dev,
                                  c, mUp.dev, g.h, dt, m_stream, [&]() { return mark(s, 1); }, m_variant, m_kchunk,
                                  zf, accf)) ||
          (rc = mark(s, 2)) || (rc = launch_src_apply(m_box, st, t, 0, dt * dt, mRho.dev, mUp.dev, m_stream)) ||
          (rc = enforceBC(mUp, mMu, mLambda, t + dt, zf)) || (rc = accel_rest(dt)) || (rc = mark(s, 3)) ||
          (rc = launch_corr_fused(m_box, g.nz, m_onesided, m_dsbp, mUp.dev, mU.dev, mUm.dev, mMu.dev, mLambda.dev,
                                  mRho.dev, c, m_opt.beta, mUp2.dev, m_dacc, g.h, dt, m_stream, [&]() { return mark(s, 4); },
                                  [&]() { return mark(s, 5); }, m_variant, m_kchunk, zf, accf, dtiles, ndtiles)) ||
          (rc = mark(s, 6)) ||
          (rc = launch_src_apply(m_box, st, t, 1, dt * dt * dt * dt / 12, mRho.dev, mUp2.dev, m_stream)) ||
          (rc = enforceBC(mUp2, mMu, mLambda, t + dt, zf)) || (rc = mark(s, 7)) || (rc = record(mUp2, mUm)))
         break;
      mT += dt;
      mCycle++;
      // Um <- U, U <- corrected, the predictor buffer and old Um are scratch
      double* old_um = mUm.dev;
      mUm.dev = mU.dev;
      mU.dev = mUp2.dev;
      mUp2.dev = old_um;
   }
   if (stage_ms) {
      if (rc == OK) rc = check_hip(hipStreamSynchronize(m_stream), "timed step synchronize");
      for (int q = 0; q < kStages; q++) stage_ms[q] = 0;
      for (int s = 0; s < n && rc == OK; s++)
         for (int q = 0; q < kStages; q++) {
            float ms = 0;
            rc = check_hip(hipEventElapsedTime(&ms, ev[(size_t)s * (kStages + 1) + q], ev[(size_t)s * (kStages + 1) + q + 1]),
                           "hipEventElapsedTime");
            if (rc) break;
            stage_ms[q] += ms;
         }
      for (auto& e : ev) (void)hipEventDestroy(e);
   }
   return rc;
}

// ---------------------------------------------------------------------------
// FP32 form with the compensated time update
// ---------------------------------------------------------------------------
CompState EW::comp_state() const
{
   const long n = m_box.npts, n3 = 3 * n;
   CompState c;
   c.U = m_cbuf;
   c.C = m_cbuf + n3;
   c.D = m_cbuf + (2 + m_cd) * n3;
   c.Dn = m_cbuf + (3 - m_cd) * n3;
   c.A = m_cbuf + 4 * n3;
   c.mu = m_cbuf + 5 * n3;
   c.la = m_cbuf + 5 * n3 + n;
   c.rho = m_cbuf + 5 * n3 + 2 * n;
   return c;
}

SgCoefF EW::sgf() const
{
   const long ni = m_box.ni, nj = m_box.je - m_box.jb + 1, nk = m_box.ke - m_box.kb + 1;
   const float* d = m_dsgf;
   SgCoefF c;
   c.dcx = d; c.strx = d + ni; c.cox = d + 2 * ni;
   d += 3 * ni;
   c.dcy = d; c.stry = d + nj; c.coy = d + 2 * nj;
   d += 3 * nj;
   c.dcz = d; c.strz = d + nk; c.coz = d + 2 * nk;
   return c;
}

int EW::comp_begin()
{
   if (m_comp) return OK;
   if (!m_opt.corder) return set_error(E_INVAL, "EW: the FP32 step needs the component-blocked layout (corder=1)");
   if (m_curv) return set_error(E_INVAL, "EW: the FP32 step is Cartesian");
   const long n = m_box.npts;
   int rc;
   if (!m_cbuf) {
      const long nsg = 3 * (m_box.ni + (m_box.je - m_box.jb + 1) + (m_box.ke - m_box.kb + 1));
      void* p = nullptr;
      if (hipMalloc(&p, sizeof(float) * 18 * n) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "EW: FP32 state allocation of " + std::to_string(sizeof(float) * 18 * n) +
                                       " bytes failed");
      }
      m_allocs.push_back(p);
      m_cbuf = static_cast<float*>(p);
      if (hipMalloc(&p, sizeof(float) * nsg + sizeof(SbpTablesT<float>)) != hipSuccess) {
         (void)hipGetLastError();
         return set_error(E_NOMEM, "EW: FP32 coefficient allocation failed");
      }
      m_allocs.push_back(p);
      m_dsbpf = static_cast<SbpTablesT<float>*>(p);
      m_dsgf = reinterpret_cast<float*>(static_cast<char*>(p) + sizeof(SbpTablesT<float>));
      SbpTables t;
      fill_stencil_coefficients(t.acof, t.ghcof, t.bope, t.sbop);
      SbpTablesT<float> tf;
      for (int q = 0; q < 384; q++) tf.acof[q] = (float)t.acof[q];
      for (int q = 0; q < 48; q++) tf.bope[q] = (float)t.bope[q];
      for (int q = 0; q < 6; q++) tf.ghcof[q] = (float)t.ghcof[q];
      for (int q = 0; q < 5; q++) tf.sbop[q] = (float)t.sbop[q];
      if ((rc = check_hip(hipMemcpy(m_dsbpf, &tf, sizeof(tf), hipMemcpyHostToDevice), "FP32 SBP upload"))) return rc;
      if ((rc = launch_to_float(nsg, m_dsg, m_dsgf, m_stream))) return rc;
   }
   float* mat = m_cbuf + 15 * n;
   if ((rc = launch_to_float(n, mMu.dev, mat, m_stream)) || (rc = launch_to_float(n, mLambda.dev, mat + n, m_stream)) ||
       (rc = launch_to_float(n, mRho.dev, mat + 2 * n, m_stream)))
      return rc;
   const CompState cs = comp_state();
   if ((rc = launch_comp_init(3 * n, mU.dev, mUm.dev, cs.U, cs.C, cs.D, m_stream))) return rc;
   m_comp = true;
   return OK;
}

int EW::comp_end()
{
   if (!m_comp) return OK;
   m_comp = false;
   const CompState cs = comp_state();
   return launch_comp_flush(3 * m_box.npts, cs.U, cs.C, cs.D, mU.dev, mUm.dev, m_stream);
}

int EW::step_comp(int n, double* stage_ms)
{
   if (!m_setup_done) return set_error(E_STATE, "EW: step before setup");
   int rc = comp_begin();
   if (rc) return rc;
   const GridSetup& g = m_opt.grid;
   const SgCoef c = sg();
   const SgCoefF cf = sgf();
   const SourceTable st = sources();
   const int nrec = (int)m_rec.size() / 4;
   std::vector<hipEvent_t> ev;
   if (stage_ms) {
      ev.resize((size_t)n * (kStages + 1));
      for (auto& e : ev)
         if ((rc = check_hip(hipEventCreate(&e), "hipEventCreate"))) return rc;
   }
   auto mark = [&](int s, int q) {
      return stage_ms ? check_hip(hipEventRecord(ev[(size_t)s * (kStages + 1) + q], m_stream), "hipEventRecord") : OK;
   };
   // zero ghost values of Dirichlet/supergrid sides set in the shell's pass
   // (the shell is exactly the union of those sides' two ghost layers; with a
   // periodic side the copies have to see the sides in their order)
   int zfill = 0;
   bool periodic = false;
   for (int q = 0; q < 6; q++) {
      if (g.bc[q] == BC_DIRICHLET || g.bc[q] == BC_SUPERGRID) zfill |= 1 << q;
      periodic = periodic || g.bc[q] == BC_PERIODIC;
   }
   if (periodic) zfill = 0;
   for (int s = 0; s < n && rc == OK; s++) {
      const double t = mT, dt = mDt;
      const CompState cs = comp_state();
      // the damping is part of the corrector sweep: stages 4 and 5 coincide
      if ((rc = mark(s, 0)) ||
          (rc = launch_comp_pred(m_box, g.nz, m_onesided, m_dsbpf, cs, cf, g.h, m_stream,
                                 [&]() { return mark(s, 1); }, m_kchunk, dt, zfill, m_variant >= 7)) ||
          (rc = mark(s, 2)) || (rc = launch_comp_src(false, m_box, st, t, 1.0, cs, m_stream)) ||
          (rc = launch_comp_bc(false, m_box, m_wind, g.nx, g.ny, g.nz, g.bc, cs, g.h, dt, m_dsbp, mMu.dev,
                               mLambda.dev, c.strx, c.stry, m_stream, zfill)) ||
          (rc = mark(s, 3)) ||
          (rc = launch_comp_corr(m_box, g.nz, m_onesided, m_dsbpf, cs, cf, m_opt.beta, g.h, dt, m_stream,
                                 [&]() { const int r = mark(s, 4); return r ? r : mark(s, 5); }, m_kchunk, zfill,
                                 m_variant >= 7)) ||
          (rc = mark(s, 6)) ||
          (rc = launch_comp_src(true, m_box, st, t, dt * dt * dt * dt / 12, cs, m_stream)) ||
          (rc = launch_comp_bc(true, m_box, m_wind, g.nx, g.ny, g.nz, g.bc, cs, g.h, dt, m_dsbp, mMu.dev,
                               mLambda.dev, c.strx, c.stry, m_stream, zfill)) ||
          (rc = mark(s, 7)))
         break;
      if (!m_norecord && nrec > 0 && m_drecdata && m_nsamples < m_maxsamples) {
         if ((rc = launch_comp_record(m_box, nrec, m_drec, cs, dt, m_drecdata + (size_t)3 * nrec * m_nsamples,
                                      m_stream)))
            break;
         m_nsamples++;
      }
      mT += dt;
      mCycle++;
      m_cd ^= 1;  // Dn becomes D
   }
   if (stage_ms) {
      if (rc == OK) rc = check_hip(hipStreamSynchronize(m_stream), "timed step synchronize");
      for (int q = 0; q < kStages; q++) stage_ms[q] = 0;
      for (int s = 0; s < n && rc == OK; s++)
         for (int q = 0; q < kStages; q++) {
            float ms = 0;
            rc = check_hip(hipEventElapsedTime(&ms, ev[(size_t)s * (kStages + 1) + q], ev[(size_t)s * (kStages + 1) + q + 1]),
                           "hipEventElapsedTime");
            if (rc) break;
            stage_ms[q] += ms;
         }
      for (auto& e : ev) (void)hipEventDestroy(e);
   }
   return rc;
}

// EW::exactSol + normOfDifference for the first source (EW.C:3650-3669,
// 3899-4523, 4526-4610), written to the scratch array mUp.
int EW::point_source_error(double& linf, double& l2, double& solinf)
{
   if (!m_setup_done || m_sources.empty()) return set_error(E_STATE, "EW: point-source error needs setup and a source");
   if (m_curv) return set_error(E_INVAL, "EW: the analytic point-source solution is for a Cartesian grid");
   const GridSetup& g = m_opt.grid;
   const MomentSource& s0 = m_sources[0];
   if (s0.tfkind != TF_GAUSSIAN && s0.tfkind != TF_C6SMOOTHBUMP)
      return set_error(E_INVAL, "EW: analytic solution needs a Gaussian or C6SmoothBump source");
   int rc;
   if ((rc = comp_end())) return rc;
   // material at the grid midpoint
   const int im = (m_box.ib + m_box.ie) / 2, jm = (m_box.jb + m_box.je) / 2, km = (m_box.kb + m_box.ke) / 2;
   const long pm = sidx(m_box, im, jm, km);
   double rmu[3];
   if ((rc = check_hip(hipMemcpyAsync(&rmu[0], mRho.dev + pm, sizeof(double), hipMemcpyDeviceToHost, m_stream), "readback")) ||
       (rc = check_hip(hipMemcpyAsync(&rmu[1], mMu.dev + pm, sizeof(double), hipMemcpyDeviceToHost, m_stream), "readback")) ||
       (rc = check_hip(hipMemcpyAsync(&rmu[2], mLambda.dev + pm, sizeof(double), hipMemcpyDeviceToHost, m_stream), "readback")) ||
       (rc = check_hip(hipStreamSynchronize(m_stream), "readback")))
      return rc;
   ExactParams P;
   P.x0 = s0.x; P.y0 = s0.y; P.z0 = s0.z;
   P.rho = rmu[0];
   P.beta = std::sqrt(rmu[1] / P.rho);
   P.alpha = std::sqrt((2 * rmu[1] + rmu[2]) / P.rho);
   P.fr = s0.freq;
   P.time = (mT - s0.t0) * s0.freq;
   if (s0.tfkind == TF_GAUSSIAN) {
      P.fr = 1 / P.fr;
      P.time = P.time * P.fr;
   }
   for (int q = 0; q < 6; q++) P.m[q] = s0.m[q];
   P.tfkind = s0.tfkind;
   P.h = g.h;
   P.zmin = g.zmin;
   P.eps = 1e-3 * g.h;
   if ((rc = launch_exact(m_box, P, mUp.dev, m_stream))) return rc;
   // window without the supergrid layers (EW.C:4537-4570)
   const int nsg = (int)(0.5 + g.sg_gp_thickness * g.h / g.h);
   NormWindow w;
   w.imin = g.bc[0] == BC_SUPERGRID ? std::max(1, nsg + 1) : 1;
   w.imax = g.bc[1] == BC_SUPERGRID ? std::min(g.nx, g.nx - nsg) : g.nx;
   w.jmin = g.bc[2] == BC_SUPERGRID ? std::max(1, nsg + 1) : 1;
   w.jmax = g.bc[3] == BC_SUPERGRID ? std::min(g.ny, g.ny - nsg) : g.ny;
   w.kmin = g.bc[4] == BC_SUPERGRID ? std::max(1, nsg + 1) : 1;
   w.kmax = g.bc[5] == BC_SUPERGRID ? std::min(g.nz, g.nz - nsg) : g.nz;
   w.x0 = s0.x; w.y0 = s0.y; w.z0 = s0.z;
   w.r2 = (4 * g.h) * (4 * g.h);
   w.h = g.h;
   w.zmin = g.zmin;
   if ((rc = launch_solerr(m_box, w, mUp.dev, mU.dev, m_dred, m_nred, m_dred + 3 * m_nred, m_stream))) return rc;
   double out[3];
   if ((rc = check_hip(hipMemcpyAsync(out, m_dred + 3 * m_nred, sizeof(out), hipMemcpyDeviceToHost, m_stream), "norm readback")) ||
       (rc = check_hip(hipStreamSynchronize(m_stream), "norm readback")))
      return rc;
   linf = out[0];
   l2 = std::sqrt(out[1]);
   solinf = out[2];
   return OK;
}

}  // namespace sw4amd
