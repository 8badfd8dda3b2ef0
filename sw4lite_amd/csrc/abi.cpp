// abi.cpp -- extern "C" entry points declared in include/sw4hip.h.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <new>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sw4hip.h"
#include "batch.h"
#include "deck.h"
#include "topo.h"
#include "ew.h"
#include "kernels_adj.h"
#include "kernels_curv.h"

namespace sw4amd {

namespace {
thread_local std::string g_last_error;

// One immutable device copy per distinct SBP table set and device: a buffer
// is written once, when it is created, and never again, so kernels still
// running on a caller's stream keep reading the tables they were launched
// with whatever tables later calls bring.
struct SbpEntry {
   SbpTables host;
   SbpTables* dev;
};
constexpr int kMaxSbpSets = 32;  // distinct table sets per device
std::mutex g_sbp_mutex;
std::vector<SbpEntry> g_sbp_cache[64];

// Device copy of an SBP table set on the current device.
int device_sbp(const SbpTables& t, const SbpTables** out)
{
   int dev = 0;
   int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
   if (rc) return rc;
   if (dev < 0 || dev >= 64) return set_error(E_RANGE, "device ordinal beyond 63");
   std::lock_guard<std::mutex> lock(g_sbp_mutex);
   std::vector<SbpEntry>& c = g_sbp_cache[dev];
   for (const SbpEntry& e : c)
      if (std::memcmp(&e.host, &t, sizeof(t)) == 0) {
         *out = e.dev;
         return OK;
      }
   if ((int)c.size() >= kMaxSbpSets)
      return set_error(E_RANGE, "more than 32 distinct SBP coefficient table sets on one device");
   void* p = nullptr;
   if (hipMalloc(&p, sizeof(SbpTables)) != hipSuccess) return set_error(E_NOMEM, "SBP table allocation failed");
   if ((rc = check_hip(hipMemcpy(p, &t, sizeof(t), hipMemcpyHostToDevice), "SBP table upload"))) {
      hipFree(p);
      return rc;
   }
   c.push_back(SbpEntry{t, static_cast<SbpTables*>(p)});
   *out = c.back().dev;
   return OK;
}

// Device copy of a grid-point source table, sorted by grid point as
// EW::sort_grid_point_sources does (stable, so sources at one point keep
// their order), packed as launch_force reads it.  One immutable buffer per
// distinct table and device (up to 16; then the device is synchronised and
// the set starts over).
struct SrcEntry {
   std::vector<char> key;
   void* dev;
   int nunique, n;
};
std::mutex g_src_mutex;
std::vector<SrcEntry> g_src_cache[64];

int device_source_table(const Box& b, int n, const int* ijk, const double* force, const double* freq,
                        const double* t0, const int* tfkind, SourceTable* out)
{
   int dev = 0;
   int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
   if (rc) return rc;
   if (dev < 0 || dev >= 64) return set_error(E_RANGE, "device ordinal beyond 63");
   *out = SourceTable{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
   if (n == 0) return OK;
   std::vector<int> order(n);
   for (int q = 0; q < n; q++) order[q] = q;
   auto key = [&](int q) { return sidx(b, ijk[3 * q], ijk[3 * q + 1], ijk[3 * q + 2]); };
   std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return key(a) < key(c); });
   std::vector<int> first;
   for (int s = 0; s < n; s++)
      if (s == 0 || key(order[s]) != key(order[s - 1])) first.push_back(s);
   const int nu = (int)first.size();
   first.push_back(n);
   // packed: first[nu+1] | ijk[3n] | tfkind[n] | force[3n] | freq[n] | t0[n]
   const size_t ibytes = sizeof(int) * (nu + 1 + 4 * n), ioff = (ibytes + 15) / 16 * 16;
   std::vector<char> host(ioff + sizeof(double) * 5 * n, 0);
   int* ip = reinterpret_cast<int*>(host.data());
   std::memcpy(ip, first.data(), sizeof(int) * (nu + 1));
   int* pijk = ip + nu + 1;
   int* ptf = pijk + 3 * n;
   double* dp = reinterpret_cast<double*>(host.data() + ioff);
   for (int s = 0; s < n; s++) {
      const int q = order[s];
      for (int c = 0; c < 3; c++) {
         pijk[3 * s + c] = ijk[3 * q + c];
         dp[3 * s + c] = force[3 * q + c];
      }
      ptf[s] = tfkind[q];
      dp[3 * n + s] = freq[q];
      dp[4 * n + s] = t0[q];
   }
   std::lock_guard<std::mutex> lock(g_src_mutex);
   std::vector<SrcEntry>& c = g_src_cache[dev];
   const SrcEntry* hit = nullptr;
   for (const SrcEntry& e : c)
      if (e.key == host) hit = &e;
   if (!hit) {
      if (c.size() >= 16) {
         if ((rc = check_hip(hipDeviceSynchronize(), "source table cache"))) return rc;
         for (SrcEntry& e : c) (void)hipFree(e.dev);
         c.clear();
      }
      void* p = nullptr;
      if (hipMalloc(&p, host.size()) != hipSuccess) return set_error(E_NOMEM, "source table allocation failed");
      if ((rc = check_hip(hipMemcpy(p, host.data(), host.size(), hipMemcpyHostToDevice), "source table upload"))) {
         (void)hipFree(p);
         return rc;
      }
      c.push_back(SrcEntry{host, p, nu, n});
      hit = &c.back();
   }
   const int* di = static_cast<const int*>(hit->dev);
   const double* dd = reinterpret_cast<const double*>(static_cast<const char*>(hit->dev) + ioff);
   out->nunique = nu;
   out->first = di;
   out->ijk = di + nu + 1;
   out->tfkind = di + nu + 1 + 3 * n;
   out->force = dd;
   out->freq = dd + 3 * n;
   out->t0 = dd + 4 * n;
   return OK;
}

SbpTables default_tables()
{
   SbpTables t;
   fill_stencil_coefficients(t.acof, t.ghcof, t.bope, t.sbop);
   return t;
}

bool box_ok(int ib, int ie, int jb, int je, int kb, int ke)
{
   return ie - ib >= 4 && je - jb >= 4 && ke - kb >= 4;
}

GridSetup to_grid(const sw4hip_grid_desc* d)
{
   GridSetup g;
   g.nx = d->nx; g.ny = d->ny; g.nz = d->nz;
   g.h = d->h;
   g.zmin = d->zmin;
   for (int s = 0; s < 6; s++) g.bc[s] = d->bc[s];
   g.sg_gp_thickness = d->sg_gp_thickness > 0 ? d->sg_gp_thickness : 20;
   return g;
}

int check_desc(const sw4hip_grid_desc* d)
{
   if (!d) return set_error(E_INVAL, "null grid descriptor");
   if (d->nx < 12 || d->ny < 12 || d->nz < 12) return set_error(E_INVAL, "grid needs at least 12 points per axis");
   if (!(d->h > 0)) return set_error(E_INVAL, "grid spacing must be positive");
   if (d->corder != 0 && d->corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   for (int s = 0; s < 6; s++)
      if (d->bc[s] < 0 || d->bc[s] > 3) return set_error(E_INVAL, "bc must be one of SW4HIP_BC_*");
   return OK;
}
}  // namespace

int set_error(int code, const std::string& msg)
{
   g_last_error = msg;
   return code;
}

int check_hip(hipError_t e, const char* what)
{
   if (e == hipSuccess) return OK;
   return set_error(e == hipErrorOutOfMemory ? E_NOMEM : E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace sw4amd

namespace sw4amd {
// Arguments of sw4hip_bcfortsg_adjoint and its allocation-free variant.
int check_bcfortsg_adjoint(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny, int nz,
                           const double* y, const double* fwd, const int* bccnd, const double* sbop, const double* mu,
                           const double* lambda, const double* strx, const double* stry, const double* gmu,
                           const double* glambda)
{
   if (!box_ok(ib, ie, jb, je, kb, ke)) return set_error(E_INVAL, "box too small");
   if (ib != -1 || jb != -1 || kb != -1 || ke != nz + 2)
      return set_error(E_INVAL, "transposed kernels need the index box [-1, n+2] of one grid");
   if (ie != nx + 2 || je != ny + 2) return set_error(E_INVAL, "bcfortsg_adjoint: box must be [-1, n+2]");
   if (!wind || !y || !fwd || !bccnd || !sbop || !mu || !lambda || !strx || !stry || !gmu || !glambda)
      return set_error(E_INVAL, "bcfortsg_adjoint: null argument");
   for (int s = 0; s < 6; s++) {
      const int* w = wind + 6 * s;
      if (bccnd[s] < 0 || bccnd[s] > 2 || w[1] < w[0] || w[3] < w[2] || w[5] < w[4]) continue;
      if (w[0] < ib || w[1] > ie || w[2] < jb || w[3] > je || w[4] < kb || w[5] > ke)
         return set_error(E_INVAL, "bcfortsg_adjoint: window outside the array");
   }
   return OK;
}
}  // namespace sw4amd

using namespace sw4amd;

struct sw4hip_ew {
   EW* ew;
};

extern "C" {

const char* sw4hip_last_error(void) { return g_last_error.c_str(); }
int sw4hip_version(void) { return 1; }

int sw4hip_get_stencil_coefficients(double* acof, double* ghcof, double* bope, double* sbop)
{
   if (!acof || !ghcof || !bope || !sbop) return set_error(E_INVAL, "null coefficient buffer");
   fill_stencil_coefficients(acof, ghcof, bope, sbop);
   return OK;
}

int sw4hip_host_supergrid(const sw4hip_grid_desc* d, double* dcx, double* dcy, double* dcz, double* strx,
                          double* stry, double* strz, double* cox, double* coy, double* coz)
{
   int rc;
   if ((rc = check_desc(d))) return rc;
   SupergridArrays a;
   make_supergrid_arrays(to_grid(d), a);
   double* dst[9] = {dcx, dcy, dcz, strx, stry, strz, cox, coy, coz};
   const std::vector<double>* src[9] = {&a.dcx, &a.dcy, &a.dcz, &a.strx, &a.stry, &a.strz, &a.cox, &a.coy, &a.coz};
   for (int q = 0; q < 9; q++) {
      if (!dst[q]) return set_error(E_INVAL, "null supergrid buffer");
      std::memcpy(dst[q], src[q]->data(), sizeof(double) * src[q]->size());
   }
   return OK;
}

int sw4hip_host_boundary_windows(const sw4hip_grid_desc* d, int wind[36], int npts[6])
{
   int rc;
   if ((rc = check_desc(d))) return rc;
   if (!wind || !npts) return set_error(E_INVAL, "null window buffer");
   make_boundary_windows(to_grid(d), wind, npts);
   return OK;
}

int sw4hip_host_moment_source(const sw4hip_grid_desc* d, double x, double y, double z, const double m[6], int maxn,
                              int* n, int* ijk, double* force)
{
   int rc;
   if ((rc = check_desc(d))) return rc;
   if (!m || !n || !ijk || !force) return set_error(E_INVAL, "null source buffer");
   MomentSource s{x, y, z, {m[0], m[1], m[2], m[3], m[4], m[5]}, 0.0, 1.0, TF_GAUSSIAN};
   std::vector<GridPointSource> pts;
   if (moment_source_points(to_grid(d), s, pts)) return set_error(E_INVAL, "grid too small for the source stencil");
   if ((int)pts.size() > maxn) return set_error(E_RANGE, "more grid-point sources than maxn");
   *n = (int)pts.size();
   for (size_t q = 0; q < pts.size(); q++) {
      ijk[3 * q] = pts[q].i; ijk[3 * q + 1] = pts[q].j; ijk[3 * q + 2] = pts[q].k;
      for (int c = 0; c < 3; c++) force[3 * q + c] = pts[q].f[c];
   }
   return OK;
}

int sw4hip_host_moment_source_curv(const sw4hip_grid_desc* d, const double* zgrid, double x, double y, double z,
                                   const double m[6], int maxn, int* n, int* ijk, double* force)
{
   int rc;
   if ((rc = check_desc(d))) return rc;
   if (!zgrid || !m || !n || !ijk || !force) return set_error(E_INVAL, "null source buffer");
   MomentSource s{x, y, z, {m[0], m[1], m[2], m[3], m[4], m[5]}, 0.0, 1.0, TF_GAUSSIAN};
   std::vector<GridPointSource> pts;
   const int q = moment_source_points_curv(to_grid(d), zgrid, s, pts);
   if (q == -1) return set_error(E_INVAL, "grid too small for the source stencil");
   if (q == -2) return set_error(E_INVAL, "source outside the curvilinear grid");
   if (q) return set_error(E_INVAL, "folded grid at the source");
   if ((int)pts.size() > maxn) return set_error(E_RANGE, "more grid-point sources than maxn");
   *n = (int)pts.size();
   for (size_t k = 0; k < pts.size(); k++) {
      ijk[3 * k] = pts[k].i; ijk[3 * k + 1] = pts[k].j; ijk[3 * k + 2] = pts[k].k;
      for (int c = 0; c < 3; c++) force[3 * k + c] = pts[k].f[c];
   }
   return OK;
}

namespace {
int to_topo(const sw4hip_topography* t, int nx, int ny, double h, double zdomain, Topography& tp, StackedGrids& sg)
{
   if (!t) return set_error(E_INVAL, "null topography");
   tp.zmax = t->zmax;
   tp.order = t->order;
   tp.zeta_break = t->zeta_break;
   tp.amp = t->amp;
   tp.xc = t->xc;
   tp.yc = t->yc;
   tp.lx = t->lx;
   tp.ly = t->ly;
   return stacked_grid_sizes(tp, nx, ny, h, zdomain, sg);
}
}  // namespace

int sw4hip_host_topography_grid(const sw4hip_topography* t, int nx, int ny, double h, double zdomain, int* nz_cart,
                                int* nz_curv, double* zgrid)
{
   Topography tp;
   StackedGrids sg;
   int rc;
   if ((rc = to_topo(t, nx, ny, h, zdomain, tp, sg))) return rc;
   if (nz_cart) *nz_cart = sg.nz_cart;
   if (nz_curv) *nz_curv = sg.nz_curv;
   if (zgrid) {
      std::vector<double> z;
      curvilinear_grid_z(tp, sg, z);
      std::memcpy(zgrid, z.data(), sizeof(double) * z.size());
   }
   return OK;
}

int sw4hip_host_topography_receiver(const sw4hip_topography* t, int nx, int ny, double h, double zdomain, double x,
                                    double y, double z, int topodepth, int* grid, int ijk[3])
{
   Topography tp;
   StackedGrids sg;
   int rc;
   if ((rc = to_topo(t, nx, ny, h, zdomain, tp, sg))) return rc;
   if (!grid || !ijk) return set_error(E_INVAL, "topography_receiver: null argument");
   rc = stacked_receiver_point(tp, sg, x, y, z, topodepth != 0, *grid, ijk[0], ijk[1], ijk[2]);
   if (rc == 1) *grid = -1;
   return rc == 1 ? OK : rc;
}

int sw4hip_host_topography_source(const sw4hip_topography* t, int nx, int ny, double h, double zdomain,
                                  const double* jac, double x, double y, double z, int topodepth, const double m[6],
                                  int maxn, int* n, int* gijk, double* force)
{
   Topography tp;
   StackedGrids sg;
   int rc;
   if ((rc = to_topo(t, nx, ny, h, zdomain, tp, sg))) return rc;
   if (!jac || !m || !n || !gijk || !force) return set_error(E_INVAL, "topography_source: null argument");
   MomentSource s{x, y, z, {m[0], m[1], m[2], m[3], m[4], m[5]}, 0.0, 1.0, TF_GAUSSIAN};
   if (topodepth) {
      double zt = 0;
      if (!topo_surface_z(tp, sg, x, y, zt)) return set_error(E_INVAL, "topography_source: off the surface");
      s.z += zt;
   }
   std::vector<GridPointSource> pc, pv;
   rc = stacked_moment_source(tp, sg, jac, s, pc, pv);
   *n = 0;
   if (rc == 1) return OK;
   if (rc) return rc;
   if ((int)(pc.size() + pv.size()) > maxn) return set_error(E_RANGE, "more grid-point sources than maxn");
   int q = 0;
   for (int g = 0; g < 2; g++)
      for (const GridPointSource& p : g == 0 ? pc : pv) {
         gijk[4 * q] = g;
         gijk[4 * q + 1] = p.i;
         gijk[4 * q + 2] = p.j;
         gijk[4 * q + 3] = p.k;
         for (int c = 0; c < 3; c++) force[3 * q + c] = p.f[c];
         q++;
      }
   *n = q;
   return OK;
}

// ---------------------------------------------------------------------------
// kernel level
// ---------------------------------------------------------------------------
int sw4hip_rhs4sg(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                  const int* onesided, const double* acof, const double* bope, const double* ghcof, double* lu,
                  const double* u, const double* mu, const double* lambda, double h, const double* strx,
                  const double* stry, const double* strz, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "rhs4sg: box too small");
   if (!onesided || !acof || !bope || !ghcof || !lu || !u || !mu || !lambda || !strx || !stry || !strz)
      return set_error(E_INVAL, "rhs4sg: null argument");
   if ((onesided[4] == 1 || onesided[5] == 1) && (nk < 12 || kfirst > -1 || klast < nk + 2))
      return set_error(E_INVAL, "rhs4sg: one-sided closures need nk >= 12 and ghost points at both ends");
   if (!(h > 0)) return set_error(E_INVAL, "rhs4sg: h must be positive");
   SbpTables t = default_tables();
   std::memcpy(t.acof, acof, sizeof(t.acof));
   std::memcpy(t.bope, bope, sizeof(t.bope));
   std::memcpy(t.ghcof, ghcof, sizeof(t.ghcof));
   const SbpTables* d = nullptr;
   int rc;
   if ((rc = device_sbp(t, &d))) return rc;
   return launch_rhs4sg(corder, make_box(ifirst, ilast, jfirst, jlast, kfirst, klast), nk, onesided, d, lu, u, mu,
                        lambda, h, strx, stry, strz, (hipStream_t)stream);
}

int sw4hip_force(int corder, int ib, int ie, int jb, int je, int kb, int ke, double t, int tt, int nsrc,
                 const int* ijk, const double* force, const double* freq, const double* t0, const int* tfkind,
                 double* F, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (ie < ib || je < jb || ke < kb || !F) return set_error(E_INVAL, "force: bad box or null F");
   if (nsrc < 0 || (nsrc > 0 && (!ijk || !force || !freq || !t0 || !tfkind)))
      return set_error(E_INVAL, "force: null source table");
   for (int q = 0; q < nsrc; q++) {
      if (ijk[3 * q] < ib || ijk[3 * q] > ie || ijk[3 * q + 1] < jb || ijk[3 * q + 1] > je || ijk[3 * q + 2] < kb ||
          ijk[3 * q + 2] > ke)
         return set_error(E_INVAL, "force: grid-point source outside the box");
      if (tfkind[q] < 0 || tfkind[q] > 2) return set_error(E_INVAL, "force: time function must be 0, 1 or 2");
   }
   const Box b = make_box(ib, ie, jb, je, kb, ke);
   SourceTable st;
   int rc;
   if ((rc = device_source_table(b, nsrc, ijk, force, freq, t0, tfkind, &st))) return rc;
   return launch_force(corder, b, F, st, t, tt ? 1 : 0, true, (hipStream_t)stream);
}

int sw4hip_predfort(int corder, int ib, int ie, int jb, int je, int kb, int ke, double* up, const double* u,
                    const double* um, const double* lu, const double* fo, const double* rho, double dt2, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (ie < ib || je < jb || ke < kb || !up || !u || !um || !lu || !fo || !rho) return set_error(E_INVAL, "predfort: bad argument");
   return launch_pred(corder, make_box(ib, ie, jb, je, kb, ke), up, u, um, lu, fo, rho, dt2, (hipStream_t)stream);
}

int sw4hip_corrfort(int corder, int ib, int ie, int jb, int je, int kb, int ke, double* up, const double* lu,
                    const double* fo, const double* rho, double dt4, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (ie < ib || je < jb || ke < kb || !up || !lu || !fo || !rho) return set_error(E_INVAL, "corrfort: bad argument");
   return launch_corr(corder, make_box(ib, ie, jb, je, kb, ke), up, lu, fo, rho, dt4, (hipStream_t)stream);
}

int sw4hip_dpdmtfort(int ib, int ie, int jb, int je, int kb, int ke, const double* up, const double* u,
                     const double* um, double* uacc, double dt2i, void* stream)
{
   if (ie < ib || je < jb || ke < kb || !up || !u || !um || !uacc) return set_error(E_INVAL, "dpdmtfort: bad argument");
   return launch_dpdmt(make_box(ib, ie, jb, je, kb, ke), up, u, um, uacc, dt2i, (hipStream_t)stream);
}

int sw4hip_addsgd4(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, double* up,
                   const double* u, const double* um, const double* rho, const double* dcx, const double* dcy,
                   const double* dcz, const double* strx, const double* stry, const double* strz, const double* cox,
                   const double* coy, const double* coz, double beta, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "addsgd4: box too small");
   if (!up || !u || !um || !rho || !dcx || !dcy || !dcz || !strx || !stry || !strz || !cox || !coy || !coz)
      return set_error(E_INVAL, "addsgd4: null argument");
   SgCoef c{dcx, dcy, dcz, strx, stry, strz, cox, coy, coz};
   return launch_addsgd4(corder, make_box(ifirst, ilast, jfirst, jlast, kfirst, klast), up, u, um, rho, c, beta,
                         (hipStream_t)stream);
}

int sw4hip_bcfortsg(int corder, int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny,
                    int nz, double* u, double h, const int* bccnd, const double* sbop, const double* mu,
                    const double* lambda, const double* const* bforce, const double* strx, const double* stry,
                    void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ib, ie, jb, je, kb, ke)) return set_error(E_INVAL, "bcfortsg: box too small");
   if (!wind || !u || !bccnd || !sbop || !mu || !lambda || !bforce || !strx || !stry)
      return set_error(E_INVAL, "bcfortsg: null argument");
   for (int s = 0; s < 6; s++) {
      const int* w = wind + 6 * s;
      const bool active = bccnd[s] >= 0 && bccnd[s] <= 3;
      if (!active || w[1] < w[0] || w[3] < w[2] || w[5] < w[4]) continue;
      if (w[0] < ib || w[1] > ie || w[2] < jb || w[3] > je || w[4] < kb || w[5] > ke)
         return set_error(E_INVAL, "bcfortsg: window outside the array");
      if (bccnd[s] == 3) {
         const int pi = s == 0 ? nx : s == 1 ? -nx : 0, pj = s == 2 ? ny : s == 3 ? -ny : 0,
                   pk = s == 4 ? nz : s == 5 ? -nz : 0;
         if (w[0] + pi < ib || w[1] + pi > ie || w[2] + pj < jb || w[3] + pj > je || w[4] + pk < kb || w[5] + pk > ke)
            return set_error(E_INVAL, "bcfortsg: periodic source outside the array");
      }
      if (bccnd[s] == 0) {
         if (s != 4 && s != 5) return set_error(E_INVAL, "bcfortsg: free surface only on sides 4 and 5");
         // top: reads planes 1..4 and writes 0; bottom: reads nz-3..nz, writes nz+1
         const bool fits = s == 4 ? (kb <= 0 && ke >= 4) : (nz - 3 >= kb && nz + 1 <= ke);
         if (!fits) return set_error(E_INVAL, "bcfortsg: free-surface plane too close to the array end");
      }
   }
   SbpTables t = default_tables();
   std::memcpy(t.sbop, sbop, sizeof(t.sbop));
   const SbpTables* d = nullptr;
   int rc;
   if ((rc = device_sbp(t, &d))) return rc;
   return launch_bcfortsg(corder, make_box(ib, ie, jb, je, kb, ke), wind, nx, ny, nz, u, h, bccnd, d, mu, lambda,
                          bforce, strx, stry, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Transposed kernels (discrete adjoint)
// ---------------------------------------------------------------------------
namespace {
int sbp_from(const double* acof, const double* bope, const double* ghcof, const SbpTables** d)
{
   SbpTables t = default_tables();
   if (acof) std::memcpy(t.acof, acof, sizeof(t.acof));
   if (bope) std::memcpy(t.bope, bope, sizeof(t.bope));
   if (ghcof) std::memcpy(t.ghcof, ghcof, sizeof(t.ghcof));
   return device_sbp(t, d);
}
int rows_ok(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk, const int* onesided)
{
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "box too small");
   if (ifirst != -1 || jfirst != -1 || kfirst != -1 || klast != nk + 2)
      return set_error(E_INVAL, "transposed kernels need the index box [-1, n+2] of one grid");
   if (onesided && (onesided[4] == 1 || onesided[5] == 1) && nk < 12)
      return set_error(E_INVAL, "one-sided closures need nk >= 12");
   return OK;
}
}  // namespace


int sw4hip_rhs4sg_adjoint(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                          const int* onesided, const double* acof, const double* bope, const double* ghcof, double* out,
                          const double* z, const double* mu, const double* lambda, double h, const double* strx,
                          const double* stry, const double* strz, void* stream)
{
   int rc;
   if ((rc = rows_ok(ifirst, ilast, jfirst, jlast, kfirst, klast, nk, onesided))) return rc;
   if (!onesided || !acof || !bope || !ghcof || !out || !z || !mu || !lambda || !strx || !stry || !strz)
      return set_error(E_INVAL, "rhs4sg_adjoint: null argument");
   if (!(h > 0)) return set_error(E_INVAL, "rhs4sg_adjoint: h must be positive");
   const SbpTables* d = nullptr;
   if ((rc = sbp_from(acof, bope, ghcof, &d))) return rc;
   const Box b = make_box(ifirst, ilast, jfirst, jlast, kfirst, klast);
   void* work = nullptr;
   if (hipMalloc(&work, sizeof(double) * adj_pad_doubles(b)) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(E_NOMEM, "rhs4sg_adjoint: scratch allocation failed");
   }
   const AdjPad pad = make_adj_pad(b, static_cast<double*>(work));
   rc = launch_adj_pad_material(b, mu, lambda, strx, stry, strz, pad, (hipStream_t)stream);
   if (rc == OK)
      rc = launch_rhs4sg_T(b, make_rowmap(b, nk, onesided), d, z, mu, lambda, h, strx, stry, strz, out, pad,
                           (hipStream_t)stream);
   const int rs = check_hip(hipStreamSynchronize((hipStream_t)stream), "rhs4sg_adjoint");
   (void)hipFree(work);
   return rc ? rc : rs;
}

int sw4hip_rhs4sg_matgrad(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                          const int* onesided, const double* acof, const double* bope, const double* ghcof,
                          const double* z, const double* u, const double* mu, const double* lambda, double h,
                          const double* strx, const double* stry, const double* strz, double* gmu, double* glambda,
                          void* stream)
{
   int rc;
   if ((rc = rows_ok(ifirst, ilast, jfirst, jlast, kfirst, klast, nk, onesided))) return rc;
   if (!onesided || !acof || !bope || !ghcof || !z || !u || !mu || !lambda || !strx || !stry || !strz || !gmu ||
       !glambda)
      return set_error(E_INVAL, "rhs4sg_matgrad: null argument");
   if (!(h > 0)) return set_error(E_INVAL, "rhs4sg_matgrad: h must be positive");
   const SbpTables* d = nullptr;
   if ((rc = sbp_from(acof, bope, ghcof, &d))) return rc;
   const Box b = make_box(ifirst, ilast, jfirst, jlast, kfirst, klast);
   rc = launch_rhs4sg_matgrad(b, make_rowmap(b, nk, onesided), d, z, u, mu, lambda, h, strx, stry, strz, gmu, glambda,
                              (hipStream_t)stream);
   const int rs = check_hip(hipStreamSynchronize((hipStream_t)stream), "rhs4sg_matgrad");
   return rc ? rc : rs;
}

int sw4hip_addsgd4_adjoint(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, const double* y,
                           const double* u, const double* um, const double* rho, const double* dcx, const double* dcy,
                           const double* dcz, const double* strx, const double* stry, const double* strz,
                           const double* cox, const double* coy, const double* coz, double beta, double* dbar,
                           double* grho, void* stream)
{
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "addsgd4_adjoint: box too small");
   if (!y || !u || !um || !rho || !dcx || !dcy || !dcz || !strx || !stry || !strz || !cox || !coy || !coz || !dbar ||
       !grho)
      return set_error(E_INVAL, "addsgd4_adjoint: null argument");
   SgCoef c{dcx, dcy, dcz, strx, stry, strz, cox, coy, coz};
   return launch_addsgd4_T_march(make_box(ifirst, ilast, jfirst, jlast, kfirst, klast), y, u, um, rho, c, beta, dbar, grho,
                           (hipStream_t)stream);
}

int sw4hip_bcfortsg_adjoint(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny, int nz,
                            double* y, const double* fwd, double h, const int* bccnd, const double* sbop,
                            const double* mu, const double* lambda, const double* strx, const double* stry,
                            double* gmu, double* glambda, void* stream)
{
   int rc;
   if ((rc = sw4amd::check_bcfortsg_adjoint(ib, ie, jb, je, kb, ke, wind, nx, ny, nz, y, fwd, bccnd, sbop, mu, lambda, strx,
                                    stry, gmu, glambda)))
      return rc;
   SbpTables t = default_tables();
   std::memcpy(t.sbop, sbop, sizeof(t.sbop));
   const SbpTables* d = nullptr;
   if ((rc = device_sbp(t, &d))) return rc;
   const Box b = make_box(ib, ie, jb, je, kb, ke);
   void* plane = nullptr;
   if (hipMalloc(&plane, sizeof(double) * 3 * b.nij) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(E_NOMEM, "bcfortsg_adjoint: scratch allocation failed");
   }
   rc = launch_bcfortsg_T(b, wind, nx, ny, nz, bccnd, y, fwd, h, d, mu, lambda, strx, stry, gmu, glambda,
                          static_cast<double*>(plane), (hipStream_t)stream);
   const int rs = check_hip(hipStreamSynchronize((hipStream_t)stream), "bcfortsg_adjoint");
   (void)hipFree(plane);
   return rc ? rc : rs;
}

// ---------------------------------------------------------------------------
// Curvilinear grid
// ---------------------------------------------------------------------------
namespace {
std::mutex g_flag_mutex;
int* g_flag[64] = {nullptr};

// A device int on the current device for kernels that report a condition.
int device_flag(int** out)
{
   int dev = 0;
   int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
   if (rc) return rc;
   if (dev < 0 || dev >= 64) return set_error(E_RANGE, "device ordinal beyond 63");
   std::lock_guard<std::mutex> lock(g_flag_mutex);
   if (!g_flag[dev]) {
      void* p = nullptr;
      if (hipMalloc(&p, sizeof(int)) != hipSuccess) return set_error(E_NOMEM, "flag allocation failed");
      g_flag[dev] = static_cast<int*>(p);
   }
   *out = g_flag[dev];
   return OK;
}
}  // namespace

int sw4hip_metric(int corder, int ib, int ie, int jb, int je, int kb, int ke, const double* x, const double* y,
                  const double* z, double* met, double* jac, void* stream)
{
   (void)y;
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (ie - ib < 5 || je - jb < 5 || ke - kb < 5) return set_error(E_INVAL, "metric: need 6 points per direction");
   if (!x || !z || !met || !jac) return set_error(E_INVAL, "metric: null argument");
   hipStream_t s = (hipStream_t)stream;
   int rc;
   double x01[2];
   int* flag = nullptr;
   if ((rc = device_flag(&flag))) return rc;
   if ((rc = check_hip(hipMemcpyAsync(x01, x, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "metric: x readback")) ||
       (rc = check_hip(hipStreamSynchronize(s), "metric: x readback")))
      return rc;
   const double h = x01[1] - x01[0];  // h = x(ib+1,jb,kb) - x(ib,jb,kb)
   if ((rc = check_hip(hipMemsetAsync(flag, 0, sizeof(int), s), "metric: flag clear")) ||
       (rc = launch_metric(corder, make_box(ib, ie, jb, je, kb, ke), h, z, met, jac, flag, s)))
      return rc;
   int bad = 0;
   if ((rc = check_hip(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s), "metric: flag readback")) ||
       (rc = check_hip(hipStreamSynchronize(s), "metric: flag readback")))
      return rc;
   if (bad) return set_error(E_INVAL, "metric: z is not increasing in k somewhere (z_r <= 0)");
   return OK;
}

int sw4hip_rhs4sgcurv(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                      const double* u, const double* mu, const double* lambda, const double* met, const double* jac,
                      double* lu, const int* onesided, const double* acof, const double* bope, const double* ghcof,
                      const double* strx, const double* stry, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "rhs4sgcurv: box too small");
   if (!u || !mu || !lambda || !met || !jac || !lu || !onesided || !acof || !bope || !ghcof || !strx || !stry)
      return set_error(E_INVAL, "rhs4sgcurv: null argument");
   // the top closure reads planes 0..8 and writes rows 1..6
   if (onesided[4] == 1 && (kfirst > 0 || klast < 8))
      return set_error(E_INVAL, "rhs4sgcurv: the top closure needs planes 0..8 in the array");
   SbpTables t = default_tables();
   std::memcpy(t.acof, acof, sizeof(t.acof));
   std::memcpy(t.bope, bope, sizeof(t.bope));
   std::memcpy(t.ghcof, ghcof, sizeof(t.ghcof));
   const SbpTables* d = nullptr;
   int rc;
   if ((rc = device_sbp(t, &d))) return rc;
   return launch_rhs4sgcurv(corder, make_box(ifirst, ilast, jfirst, jlast, kfirst, klast), onesided, d, u, mu,
                            lambda, met, jac, strx, stry, lu, (hipStream_t)stream);
}

int sw4hip_addsgd4c(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, double* up,
                    const double* u, const double* um, const double* rho, const double* dcx, const double* dcy,
                    const double* strx, const double* stry, const double* jac, const double* cox, const double* coy,
                    double beta, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ifirst, ilast, jfirst, jlast, kfirst, klast)) return set_error(E_INVAL, "addsgd4c: box too small");
   if (!up || !u || !um || !rho || !dcx || !dcy || !strx || !stry || !jac || !cox || !coy)
      return set_error(E_INVAL, "addsgd4c: null argument");
   SgCoef c{dcx, dcy, nullptr, strx, stry, nullptr, cox, coy, nullptr};
   return launch_addsgd4c(corder, make_box(ifirst, ilast, jfirst, jlast, kfirst, klast), up, u, um, rho, jac, c, beta,
                          (hipStream_t)stream);
}

int sw4hip_freesurfcurvisg(int corder, int ib, int ie, int jb, int je, int kb, int ke, int nz, int side, double* u,
                           const double* mu, const double* lambda, const double* met, const double* sbop,
                           const double* forcing, const double* strx, const double* stry, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!box_ok(ib, ie, jb, je, kb, ke)) return set_error(E_INVAL, "freesurfcurvisg: box too small");
   if (!u || !mu || !lambda || !met || !sbop || !strx || !stry) return set_error(E_INVAL, "freesurfcurvisg: null argument");
   if (side != 5 && side != 6) return set_error(E_INVAL, "freesurfcurvisg: side must be 5 (k = 1) or 6 (k = nz)");
   const bool fits = side == 5 ? (kb <= 0 && ke >= 4) : (nz - 3 >= kb && nz + 1 <= ke);
   if (!fits) return set_error(E_INVAL, "freesurfcurvisg: boundary plane too close to the array end");
   SbpTables t = default_tables();
   std::memcpy(t.sbop, sbop, sizeof(t.sbop));
   const SbpTables* d = nullptr;
   int rc;
   if ((rc = device_sbp(t, &d))) return rc;
   return launch_freesurf_curv(corder, make_box(ib, ie, jb, je, kb, ke), nz, side - 1, u, mu, lambda, met, d, forcing,
                               strx, stry, (hipStream_t)stream);
}

int sw4hip_enforce_cart_topo(int corder, int ib, int ie, int jb, int je, int kb_cart, int ke_cart, double* ucart,
                             int kb_curv, int ke_curv, double* ucurv, int ghost_points, void* stream)
{
   if (corder != 0 && corder != 1) return set_error(E_INVAL, "corder must be 0 or 1");
   if (!ucart || !ucurv) return set_error(E_INVAL, "enforce_cart_topo: null argument");
   if (ib > ie || jb > je) return set_error(E_INVAL, "enforce_cart_topo: empty (i, j) range");
   if (ghost_points < 1) return set_error(E_INVAL, "enforce_cart_topo: ghost_points must be at least 1");
   if (kb_cart + 2 * ghost_points > ke_cart || ke_curv - 2 * ghost_points < kb_curv)
      return set_error(E_INVAL, "enforce_cart_topo: a grid is too thin for its ghost planes");
   return launch_enforce_cart_topo(corder, make_box(ib, ie, jb, je, kb_cart, ke_cart), ucart,
                                   make_box(ib, ie, jb, je, kb_curv, ke_curv), ucurv, ghost_points, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// EW level
// ---------------------------------------------------------------------------
int sw4hip_ew_create(const sw4hip_grid_desc* d, sw4hip_ew** out)
{
   int rc;
   if (!out) return set_error(E_INVAL, "null output handle");
   *out = nullptr;
   if ((rc = check_desc(d))) return rc;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      (void)hipGetLastError();
      return set_error(E_NODEV, "no HIP device available");
   }
   EWOptions o;
   o.grid = to_grid(d);
   o.corder = d->corder;
   o.beta = d->sg_damping >= 0 ? d->sg_damping : 0.02;
   o.cfl = d->cfl > 0 ? d->cfl : 1.3;
   EW* ew = new EW(o);
   if ((rc = ew->init())) {
      delete ew;
      return rc;
   }
   *out = new sw4hip_ew{ew};
   return OK;
}

int sw4hip_ew_destroy(sw4hip_ew* h)
{
   if (!h) return OK;
   delete h->ew;
   delete h;
   return OK;
}

#define EWCHK(h)                                                  \
   if (!(h) || !(h)->ew) return set_error(E_INVAL, "null EW handle")

int sw4hip_ew_set_stream(sw4hip_ew* h, void* stream)
{
   EWCHK(h);
   h->ew->set_stream((hipStream_t)stream);
   return OK;
}

int sw4hip_ew_set_kernel_variant(sw4hip_ew* h, int variant, int kchunk)
{
   EWCHK(h);
   if (variant < 0 || variant > 10) return set_error(E_INVAL, "variant must be 0..10");
   if (kchunk > 4096) return set_error(E_INVAL, "kchunk must be at most 4096");
   h->ew->m_variant = variant;
   h->ew->m_kchunk = kchunk > 0 ? kchunk : EW::kSweepChunk;
   h->ew->m_kchunk_damp = kchunk > 0 ? kchunk : EW::kDampChunk;
   return OK;
}

int sw4hip_ew_set_material(sw4hip_ew* h, const double* rho, const double* mu, const double* la, int on_device)
{
   EWCHK(h);
   if (!rho || !mu || !la) return set_error(E_INVAL, "null material array");
   return h->ew->set_material(rho, mu, la, on_device != 0);
}

int sw4hip_ew_set_material_const(sw4hip_ew* h, double rho, double cs, double cp)
{
   EWCHK(h);
   if (!(rho > 0) || !(cs > 0) || !(cp > cs)) return set_error(E_INVAL, "need rho > 0 and cp > cs > 0");
   return h->ew->set_material_const(rho, cs, cp);
}

int sw4hip_ew_add_moment_source(sw4hip_ew* h, double x, double y, double z, const double m[6], double t0, double freq,
                                int tfkind)
{
   EWCHK(h);
   if (!m) return set_error(E_INVAL, "null moment tensor");
   if (!(freq > 0)) return set_error(E_INVAL, "source frequency must be positive");
   const GridSetup& g = h->ew->m_opt.grid;
   if (x < 0 || x > (g.nx - 1) * g.h || y < 0 || y > (g.ny - 1) * g.h)
      return set_error(E_INVAL, "source positioned outside the grid");
   // on a curvilinear grid the depth is checked against the grid at setup
   if (!h->ew->curvilinear() && (z < g.zmin || z > g.zmin + (g.nz - 1) * g.h))
      return set_error(E_INVAL, "source positioned outside the grid");
   MomentSource s{x, y, z, {m[0], m[1], m[2], m[3], m[4], m[5]}, t0, freq, tfkind};
   return h->ew->add_moment_source(s);
}

int sw4hip_ew_set_topography(sw4hip_ew* h, const double* z, int on_device)
{
   EWCHK(h);
   return h->ew->set_topography(z, on_device != 0);
}

int sw4hip_ew_setup(sw4hip_ew* h, double tmax, int nsteps)
{
   EWCHK(h);
   if (!(tmax > 0) && nsteps <= 0) return set_error(E_INVAL, "give tmax > 0 or nsteps > 0");
   return h->ew->setup(tmax, nsteps);
}

int sw4hip_ew_get_info(sw4hip_ew* h, sw4hip_ew_info* info)
{
   EWCHK(h);
   if (!info) return set_error(E_INVAL, "null info");
   EW& e = *h->ew;
   info->dt = e.mDt;
   info->nsteps = e.mNumberOfTimeSteps;
   info->t = e.mT;
   info->cycle = e.mCycle;
   for (int s = 0; s < 6; s++) info->onesided[s] = e.m_onesided[s];
   for (int q = 0; q < 36; q++) info->wind[q] = e.m_wind[q];
   info->npts = e.m_box.npts;
   info->nsources = (int)e.m_gps.size();
   info->nunique = e.m_first.empty() ? 0 : (int)e.m_first.size() - 1;
   return OK;
}

int sw4hip_ew_get_supergrid(sw4hip_ew* h, double* dcx, double* dcy, double* dcz, double* strx, double* stry,
                            double* strz, double* cox, double* coy, double* coz)
{
   EWCHK(h);
   const SupergridArrays& a = h->ew->m_sg_host;
   if (a.dcx.empty()) return set_error(E_STATE, "supergrid arrays exist after setup");
   double* dst[9] = {dcx, dcy, dcz, strx, stry, strz, cox, coy, coz};
   const std::vector<double>* src[9] = {&a.dcx, &a.dcy, &a.dcz, &a.strx, &a.stry, &a.strz, &a.cox, &a.coy, &a.coz};
   for (int q = 0; q < 9; q++) {
      if (!dst[q]) return set_error(E_INVAL, "null supergrid buffer");
      std::memcpy(dst[q], src[q]->data(), sizeof(double) * src[q]->size());
   }
   return OK;
}

int sw4hip_ew_get_sources(sw4hip_ew* h, int* ijk, double* force, double* freq, double* t0, int* tfkind, int* first)
{
   EWCHK(h);
   const EW& e = *h->ew;
   for (size_t s = 0; s < e.m_gps.size(); s++) {
      if (ijk) { ijk[3 * s] = e.m_gps[s].i; ijk[3 * s + 1] = e.m_gps[s].j; ijk[3 * s + 2] = e.m_gps[s].k; }
      if (force) for (int c = 0; c < 3; c++) force[3 * s + c] = e.m_gps[s].f[c];
      if (freq) freq[s] = e.m_gps[s].freq;
      if (t0) t0[s] = e.m_gps[s].t0;
      if (tfkind) tfkind[s] = e.m_gps[s].tfkind;
   }
   if (first) for (size_t r = 0; r < e.m_first.size(); r++) first[r] = e.m_first[r];
   return OK;
}

int sw4hip_ew_set_solution(sw4hip_ew* h, const double* u, const double* um, int on_device)
{
   EWCHK(h);
   EW& e = *h->ew;
   const size_t bytes = sizeof(double) * e.mU.size();
   const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
   int rc;
   if ((rc = e.comp_end())) return rc;
   if (u && (rc = check_hip(hipMemcpyAsync(e.mU.dev, u, bytes, kind, e.m_stream), "U copy"))) return rc;
   if (um && (rc = check_hip(hipMemcpyAsync(e.mUm.dev, um, bytes, kind, e.m_stream), "Um copy"))) return rc;
   return check_hip(hipStreamSynchronize(e.m_stream), "solution copy");
}

int sw4hip_ew_get_solution(sw4hip_ew* h, double* u, double* um, int on_device)
{
   EWCHK(h);
   EW& e = *h->ew;
   const size_t bytes = sizeof(double) * e.mU.size();
   const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
   int rc;
   if ((rc = e.comp_end())) return rc;
   if (u && (rc = check_hip(hipMemcpyAsync(u, e.mU.dev, bytes, kind, e.m_stream), "U copy"))) return rc;
   if (um && (rc = check_hip(hipMemcpyAsync(um, e.mUm.dev, bytes, kind, e.m_stream), "Um copy"))) return rc;
   return check_hip(hipStreamSynchronize(e.m_stream), "solution copy");
}

int sw4hip_ew_step(sw4hip_ew* h, int nsteps, int mode)
{
   EWCHK(h);
   if (nsteps < 0) return set_error(E_INVAL, "negative step count");
   if (mode == 0) return h->ew->step_fused(nsteps);
   if (mode == 1) return h->ew->step_reference(nsteps);
   if (mode == 2) return h->ew->step_comp(nsteps);
   return set_error(E_INVAL, "step mode must be 0 (fused), 1 (reference order) or 2 (fused FP32, compensated)");
}

int sw4hip_ew_step_timed(sw4hip_ew* h, int nsteps, double* stage_ms)
{
   EWCHK(h);
   if (nsteps < 0) return set_error(E_INVAL, "negative step count");
   if (!stage_ms) return set_error(E_INVAL, "null stage_ms");
   return h->ew->step_fused(nsteps, stage_ms);
}

int sw4hip_ew_step_comp_timed(sw4hip_ew* h, int nsteps, double* stage_ms)
{
   EWCHK(h);
   if (nsteps < 0) return set_error(E_INVAL, "negative step count");
   if (!stage_ms) return set_error(E_INVAL, "null stage_ms");
   return h->ew->step_comp(nsteps, stage_ms);
}

int sw4hip_ew_reset(sw4hip_ew* h)
{
   EWCHK(h);
   return h->ew->reset();
}

int sw4hip_ew_gradient(sw4hip_ew* h, int nsteps, int ckpt, const double* grad_traces, double* grho, double* gmu,
                       double* glambda, int on_device)
{
   EWCHK(h);
   return h->ew->gradient(nsteps, ckpt, grad_traces, grho, gmu, glambda, on_device != 0, false,
                          h->ew->m_adj_store);
}

int sw4hip_ew_gradient_comp(sw4hip_ew* h, int nsteps, int ckpt, const double* grad_traces, double* grho,
                            double* gmu, double* glambda, int on_device)
{
   EWCHK(h);
   return h->ew->gradient(nsteps, ckpt, grad_traces, grho, gmu, glambda, on_device != 0, true);
}

int sw4hip_ew_gradient_plan(sw4hip_ew* h, int nsteps, int ckpt, int* used_ckpt, double* device_bytes)
{
   EWCHK(h);
   if (!used_ckpt) return set_error(E_INVAL, "null output");
   return h->ew->gradient_plan(nsteps, ckpt, used_ckpt, device_bytes, false, h->ew->m_adj_store);
}

int sw4hip_ew_set_gradient_store(sw4hip_ew* h, int store)
{
   EWCHK(h);
   if (store < -1 || store > 1) return set_error(E_INVAL, "gradient store must be -1, 0 or 1");
   h->ew->m_adj_store = store;
   return OK;
}

int sw4hip_ew_release_workspace(sw4hip_ew* h)
{
   EWCHK(h);
   return h->ew->release_workspace();
}

int sw4hip_ew_point_source_error(sw4hip_ew* h, double* linf, double* l2, double* solinf)
{
   EWCHK(h);
   if (!linf || !l2 || !solinf) return set_error(E_INVAL, "null output");
   return h->ew->point_source_error(*linf, *l2, *solinf);
}

int sw4hip_ew_add_receiver(sw4hip_ew* h, double x, double y, double z, int mode, int* index)
{
   EWCHK(h);
   return h->ew->add_receiver(x, y, z, mode, index);
}

int sw4hip_ew_get_receivers(sw4hip_ew* h, int* nsamples, double* out)
{
   EWCHK(h);
   return h->ew->get_receivers(nsamples, out);
}

int sw4hip_ew_device_array(sw4hip_ew* h, int which, double** ptr)
{
   EWCHK(h);
   if (!ptr) return set_error(E_INVAL, "null output");
   EW& e = *h->ew;
   if (which <= 1) {
      const int rc = e.comp_end();
      if (rc) return rc;
   }
   switch (which) {
   case 0: *ptr = e.mU.dev; break;
   case 1: *ptr = e.mUm.dev; break;
   case 2: *ptr = e.mRho.dev; break;
   case 3: *ptr = e.mMu.dev; break;
   case 4: *ptr = e.mLambda.dev; break;
   case 5:
   case 6:
      if (!e.curvilinear()) return set_error(E_STATE, "no topography set: the grid has no metric");
      *ptr = which == 5 ? e.mMetric.dev : e.mJ.dev;
      break;
   default: return set_error(E_INVAL, "which must be 0..6");
   }
   return OK;
}

struct sw4hip_batch {
   Batch* b;
};

int sw4hip_batch_create(sw4hip_ew* const* shots, int nshots, sw4hip_batch** out)
{
   if (!shots || nshots < 1 || !out) return set_error(E_INVAL, "batch: need shots and an output handle");
   std::vector<EW*> v;
   for (int q = 0; q < nshots; q++) {
      if (!shots[q] || !shots[q]->ew) return set_error(E_INVAL, "batch: null shot");
      v.push_back(shots[q]->ew);
   }
   Batch* b = new (std::nothrow) Batch(v);
   if (!b) return set_error(E_NOMEM, "batch: allocation failed");
   const int rc = b->init();
   if (rc) {
      delete b;
      return rc;
   }
   *out = new (std::nothrow) sw4hip_batch{b};
   if (!*out) {
      delete b;
      return set_error(E_NOMEM, "batch: allocation failed");
   }
   return OK;
}

int sw4hip_batch_destroy(sw4hip_batch* h)
{
   if (!h) return OK;
   delete h->b;
   delete h;
   return OK;
}

int sw4hip_batch_set_kchunk(sw4hip_batch* h, int kchunk)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->set_kchunk(kchunk);
}

int sw4hip_batch_tune(sw4hip_batch* h, int ncand, const int* cand, int* best, double* ms)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->tune(best, ms, ncand, cand);
}

int sw4hip_batch_gradient(sw4hip_batch* h, int nsteps, int ckpt, const double* const* grad_traces,
                          double* const* grho, double* const* gmu, double* const* glambda, int on_device,
                          int max_concurrent)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->gradient(nsteps, ckpt, grad_traces, grho, gmu, glambda, on_device != 0, max_concurrent);
}

int sw4hip_batch_gradient_comp(sw4hip_batch* h, int nsteps, int ckpt, const double* const* grad_traces,
                               double* const* grho, double* const* gmu, double* const* glambda, int on_device,
                               int max_concurrent)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->gradient(nsteps, ckpt, grad_traces, grho, gmu, glambda, on_device != 0, max_concurrent, true);
}

int sw4hip_batch_release_workspace(sw4hip_batch* h)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->release_workspace();
}

int sw4hip_batch_step(sw4hip_batch* h, int nsteps, int use_graph)
{
   if (!h || !h->b) return set_error(E_INVAL, "null batch");
   return h->b->step(nsteps, use_graph != 0);
}

int sw4hip_ew_sync(sw4hip_ew* h)
{
   EWCHK(h);
   return check_hip(hipStreamSynchronize(h->ew->m_stream), "stream synchronize");
}

}  // extern "C"

namespace {
void fill_report(const sw4amd::DeckResult& r, sw4hip_deck_report* rep)
{
   rep->nx = r.nx;
   rep->ny = r.ny;
   rep->nz = r.nz;
   rep->h = r.h;
   rep->dt = r.dt;
   rep->t = r.t;
   rep->nsteps = r.nsteps;
   rep->cycle = r.cycle;
   rep->nstations = r.nstations;
   rep->ncheckpoints = r.ncheckpoints;
   rep->point_source_test = r.point_source_test;
   rep->linf = r.linf;
   rep->l2 = r.l2;
   rep->solinf = r.solinf;
   rep->ngrids = r.ngrids;
   rep->nz_curv = r.nz_curv;
}
}  // namespace

extern "C" int sw4hip_parse_deck(const char* text, sw4hip_deck_report* rep)
{
   if (!text || !rep) return set_error(E_INVAL, "parse_deck: null argument");
   Deck d;
   int rc;
   if ((rc = parse_deck(text, d))) return rc;
   DeckResult r;
   r.nx = d.nx;
   r.ny = d.ny;
   r.nz = d.nz;
   r.h = d.h;
   r.t = d.tmax;
   r.nsteps = d.steps;
   r.nstations = (int)d.stations.size();
   r.ncheckpoints = (int)d.checkpoints.size();
   r.point_source_test = d.point_source_test;
   fill_report(r, rep);
   return OK;
}

extern "C" int sw4hip_run_deck(const char* file, const char* path, int mode, sw4hip_deck_report* rep)
{
   if (!file || !rep) return set_error(E_INVAL, "run_deck: null argument");
   std::ifstream in(file);
   if (!in) return set_error(E_INVAL, std::string("run_deck: cannot open ") + file);
   std::stringstream ss;
   ss << in.rdbuf();
   Deck d;
   int rc;
   if ((rc = parse_deck(ss.str(), d))) return rc;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      (void)hipGetLastError();
      return set_error(E_NODEV, "no HIP device available");
   }
   DeckResult r;
   if ((rc = run_deck(d, path ? path : "", mode, r))) return rc;
   fill_report(r, rep);
   return OK;
}

extern "C" int sw4hip_write_station(const sw4hip_station* s, const char* path, double t0, double dt, int nsamples,
                                    const double* samples)
{
   if (!s || !s->file || !path) return set_error(E_INVAL, "write_station: null argument");
   Station st;
   st.file = s->file;
   st.sta = s->sta ? s->sta : s->file;
   st.mode = s->mode;
   st.xyz = s->nsew != 1;
   st.usgs = s->usgs != 0;
   st.sac = s->sac != 0;
   st.x = s->x;
   st.y = s->y;
   st.z = s->z;
   st.gx = s->gx;
   st.gy = s->gy;
   st.gz = s->gz;
   for (int q = 0; q < 7; q++) st.utc[q] = s->utc[q];
   std::string p = path;
   if (p != "." && !p.empty() && p.back() != '/') p += '/';
   return write_station(p, st, GeoFrame(), t0, dt, nsamples, samples);
}

extern "C" int sw4hip_write_checkpoint(const char* file, double time, int cycle, int nx, int ny, int nz,
                                       const double* um, const double* u)
{
   if (!file || !um || !u) return set_error(E_INVAL, "write_checkpoint: null argument");
   if (nx < 1 || ny < 1 || nz < 1) return set_error(E_INVAL, "write_checkpoint: empty grid");
   return write_checkpoint(file, time, cycle, nx, ny, nz, um, u);
}

extern "C" int sw4hip_read_checkpoint(const char* file, int nx, int ny, int nz, double* time, int* cycle, double* um,
                                      double* u)
{
   if (!file || !time || !cycle) return set_error(E_INVAL, "read_checkpoint: null argument");
   if (nx < 1 || ny < 1 || nz < 1) return set_error(E_INVAL, "read_checkpoint: empty grid");
   return read_checkpoint(file, nx, ny, nz, time, cycle, um, u);
}

