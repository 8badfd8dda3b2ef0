// ref_ew.cpp -- C-linkage access to the reference's EW class, built with the
// rest of the reference program from its own sources by oracle/Makefile into
// oracle/_ref/libsw4ew.so (linked against the image's MPICH and MKL).
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and tests/golden/ scripts to
// check the oracle, never by the sw4lite_amd package.
//
// An EW object is made the reference's way, from an input deck
// (EW::EW, EW.C:162-224: parseInputFile, setupRun, timesteploop), after
// which its public member kernels are called on caller arrays:
//   predfort / corrfort / dpdmtfort       ew-cfromfort.C:40-141
//   solerr3fort                           ew-cfromfort.C:144-202
//   bcfortsg / bcfortsg_indrev            ew-cfromfort.C:205-747
//   addsgd4fort(_indrev)                  ew-cfromfort.C:748-849, 952-1056
//   addsgd4cfort(_indrev)                 ew-cfromfort.C:1160-1240, 1323-1404
//   GetStencilCoefficients                ew-cfromfort.C:1487-1944
//   metricexgh(_rev)                      curvilinear-c.C:298-462
//   freesurfcurvisg(_rev)                 curvilinear-c.C:465-764
//   get_exact_point_source                EW.C:3899-4523
// plus the free time functions of time_functions.C and the SuperGrid taper
// (SuperGrid.C:54-194).  The deck run itself gives the reference's dt,
// grid-point sources and final fields.
#include <mpi.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "EW.h"
#include "GridPointSource.h"
#include "Source.h"
#include "SuperGrid.h"
#include "time_functions.h"

typedef double R;

extern "C" {

// MPI_Init once per process (the reference calls MPI in its constructor).
int refew_init(void)
{
   int flag = 0;
   MPI_Initialized(&flag);
   if (!flag) {
      int argc = 0;
      char** argv = nullptr;
      MPI_Init(&argc, &argv);
   }
   return 0;
}

// Runs the deck (parse, set-up, time loop) and keeps the object.
void* refew_create(const char* deck)
{
   refew_init();
   return new EW(std::string(deck));
}

void refew_destroy(void* ew) { delete static_cast<EW*>(ew); }

// layout switch of the member kernels (EW::m_corder, EW.h:475)
void refew_set_corder(void* ew, int corder) { static_cast<EW*>(ew)->m_corder = corder != 0; }

void refew_info(void* ew, R* dt, int* nsteps, int* ngrids, int* box)
{
   EW* e = static_cast<EW*>(ew);
   *dt = e->mDt;
   *nsteps = e->mNumberOfTimeSteps;
   *ngrids = e->mNumberOfGrids;
   for (int g = 0; g < e->mNumberOfGrids && g < 4; g++) {
      box[6 * g + 0] = e->m_iStart[g];
      box[6 * g + 1] = e->m_iEnd[g];
      box[6 * g + 2] = e->m_jStart[g];
      box[6 * g + 3] = e->m_jEnd[g];
      box[6 * g + 4] = e->m_kStart[g];
      box[6 * g + 5] = e->m_kEnd[g];
   }
}

// grid spacing, zmin and the global nz of grid g
void refew_grid(void* ew, int g, R* h, R* zmin, int* nz)
{
   EW* e = static_cast<EW*>(ew);
   *h = e->mGridSize[g];
   *zmin = e->m_zmin[g];
   *nz = e->m_global_nz[g];
}

// Copies a member array of grid g (which: 0 U, 1 Um, 2 rho, 3 mu, 4 lambda,
// 5 x, 6 y, 7 z, 8 metric, 9 Jacobian) into out, in the array's own layout;
// returns the number of doubles (out may be NULL to query).
long refew_array(void* ew, int g, int which, R* out)
{
   EW* e = static_cast<EW*>(ew);
   Sarray* a = nullptr;
   switch (which) {
      case 0: a = &e->mU[g]; break;
      case 1: a = &e->mUm[g]; break;
      case 2: a = &e->mRho[g]; break;
      case 3: a = &e->mMu[g]; break;
      case 4: a = &e->mLambda[g]; break;
      case 5: a = &e->mX; break;
      case 6: a = &e->mY; break;
      case 7: a = &e->mZ; break;
      case 8: a = &e->mMetric; break;
      case 9: a = &e->mJ; break;
      default: return -1;
   }
   const long n = (long)a->m_npts;  // m_npts counts the components too (Sarray.C:755)
   if (out) std::memcpy(out, a->c_ptr(), sizeof(R) * n);
   return n;
}

// Boundary set-up of grid g (EW::setup_boundary_arrays, assign_local_bcs):
// wind[36], npts[6], bctype[6], onesided[6].
void refew_boundary(void* ew, int g, int* wind, int* npts, int* bctype, int* onesided)
{
   EW* e = static_cast<EW*>(ew);
   for (int q = 0; q < 36; q++) wind[q] = e->m_BndryWindow[g][q];
   for (int s = 0; s < 6; s++) {
      npts[s] = e->m_NumberOfBCPoints[g][s];
      bctype[s] = (int)e->m_bcType[g][s];
      onesided[s] = e->m_onesided[g][s];
   }
}

// 1-D supergrid arrays of grid g: 9 arrays (dcx, dcy, dcz, strx, stry, strz,
// cox, coy, coz) of the box lengths ni, nj, nk, concatenated per array.
void refew_supergrid(void* ew, int g, R* out)
{
   EW* e = static_cast<EW*>(ew);
   const int ni = e->m_iEnd[g] - e->m_iStart[g] + 1, nj = e->m_jEnd[g] - e->m_jStart[g] + 1,
             nk = e->m_kEnd[g] - e->m_kStart[g] + 1;
   R* p = out;
   std::memcpy(p, e->m_sg_dc_x[g], sizeof(R) * ni), p += ni;
   std::memcpy(p, e->m_sg_dc_y[g], sizeof(R) * nj), p += nj;
   std::memcpy(p, e->m_sg_dc_z[g], sizeof(R) * nk), p += nk;
   std::memcpy(p, e->m_sg_str_x[g], sizeof(R) * ni), p += ni;
   std::memcpy(p, e->m_sg_str_y[g], sizeof(R) * nj), p += nj;
   std::memcpy(p, e->m_sg_str_z[g], sizeof(R) * nk), p += nk;
   std::memcpy(p, e->m_sg_corner_x[g], sizeof(R) * ni), p += ni;
   std::memcpy(p, e->m_sg_corner_y[g], sizeof(R) * nj), p += nj;
   std::memcpy(p, e->m_sg_corner_z[g], sizeof(R) * nk);
}

// The grid-point sources of the run (EW::m_point_sources after
// sort_grid_point_sources): per source i, j, k, grid; the forces without the
// time function (getFxyz_notime) and the forces and their second time
// derivatives at time t (getFxyz, getFxyztt; GridPointSource.C:330-586).
// Returns their number (outputs may be NULL to query).
int refew_point_sources(void* ew, R t, int* ijkg, R* force, R* ft, R* ftt)
{
   EW* e = static_cast<EW*>(ew);
   const int n = (int)e->m_point_sources.size();
   if (ijkg)
      for (int s = 0; s < n; s++) {
         GridPointSource* p = e->m_point_sources[s];
         ijkg[4 * s + 0] = p->m_i0;
         ijkg[4 * s + 1] = p->m_j0;
         ijkg[4 * s + 2] = p->m_k0;
         ijkg[4 * s + 3] = p->m_grid;
         p->getFxyz_notime(force + 3 * s);
         p->getFxyz(t, ft + 3 * s);
         p->getFxyztt(t, ftt + 3 * s);
      }
   return n;
}

void refew_predfort(void* ew, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* u, R* um, R* lu, R* fo,
                    R* rho, R dt2)
{
   static_cast<EW*>(ew)->predfort(ib, ie, jb, je, kb, ke, up, u, um, lu, fo, rho, dt2);
}

void refew_corrfort(void* ew, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* lu, R* fo, R* rho, R dt4)
{
   static_cast<EW*>(ew)->corrfort(ib, ie, jb, je, kb, ke, up, lu, fo, rho, dt4);
}

void refew_dpdmtfort(void* ew, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* u, R* um, R* u2, R dt2i)
{
   static_cast<EW*>(ew)->dpdmtfort(ib, ie, jb, je, kb, ke, up, u, um, u2, dt2i);
}

void refew_solerr3(void* ew, int ib, int ie, int jb, int je, int kb, int ke, R h, R* uex, R* u, R* li, R* l2,
                   R* xli, R zmin, R x0, R y0, R z0, R radius, int imin, int imax, int jmin, int jmax, int kmin,
                   int kmax)
{
   static_cast<EW*>(ew)->solerr3fort(ib, ie, jb, je, kb, ke, h, uex, u, *li, *l2, *xli, zmin, x0, y0, z0, radius,
                                     imin, imax, jmin, jmax, kmin, kmax);
}

// bccnd: boundaryConditionType values (sw4.h:37-38); bforce: six arrays or NULL
void refew_bcfortsg(void* ew, int corder, int ib, int ie, int jb, int je, int kb, int ke, int* wind, int nx, int ny,
                    int nz, R* u, R h, const int* bccnd, R* sbop, R* mu, R* la, R** bforce, R* strx, R* stry)
{
   boundaryConditionType bc[6];
   for (int s = 0; s < 6; s++) bc[s] = (boundaryConditionType)bccnd[s];
   EW* e = static_cast<EW*>(ew);
   if (corder)
      e->bcfortsg_indrev(ib, ie, jb, je, kb, ke, wind, nx, ny, nz, u, h, bc, sbop, mu, la, 0.0, bforce[0], bforce[1],
                         bforce[2], bforce[3], bforce[4], bforce[5], 0.0, 0.0, 0.0, strx, stry);
   else
      e->bcfortsg(ib, ie, jb, je, kb, ke, wind, nx, ny, nz, u, h, bc, sbop, mu, la, 0.0, bforce[0], bforce[1],
                  bforce[2], bforce[3], bforce[4], bforce[5], 0.0, 0.0, 0.0, strx, stry);
}

void refew_addsgd4(void* ew, int corder, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* u, R* um, R* rho,
                   R* dcx, R* dcy, R* dcz, R* strx, R* stry, R* strz, R* cox, R* coy, R* coz, R beta)
{
   EW* e = static_cast<EW*>(ew);
   if (corder)
      e->addsgd4fort_indrev(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz,
                            beta);
   else
      e->addsgd4fort(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta);
}

void refew_addsgd4c(void* ew, int corder, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* u, R* um,
                    R* rho, R* dcx, R* dcy, R* strx, R* stry, R* jac, R* cox, R* coy, R beta)
{
   EW* e = static_cast<EW*>(ew);
   if (corder)
      e->addsgd4cfort_indrev(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, strx, stry, jac, cox, coy, beta);
   else
      e->addsgd4cfort(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, strx, stry, jac, cox, coy, beta);
}

void refew_stencil_coefficients(void* ew, R* acof, R* ghcof, R* bope, R* sbop)
{
   static_cast<EW*>(ew)->GetStencilCoefficients(acof, ghcof, bope, sbop);
}

void refew_freesurfcurvisg(void* ew, int corder, int ib, int ie, int jb, int je, int kb, int ke, int nz, int side,
                           R* u, R* mu, R* la, R* met, R* sbop, R* forcing, R* strx, R* stry)
{
   EW* e = static_cast<EW*>(ew);
   if (corder)
      e->freesurfcurvisg_rev(ib, ie, jb, je, kb, ke, nz, side, u, mu, la, met, sbop, forcing, strx, stry);
   else
      e->freesurfcurvisg(ib, ie, jb, je, kb, ke, nz, side, u, mu, la, met, sbop, forcing, strx, stry);
}

void refew_metricexgh(void* ew, int corder, int ib, int ie, int jb, int je, int kb, int ke, int nz, R* x, R* y,
                      R* z, R* met, R* jac, int order, R sb, R zmax, R amp, R xc, R yc, R xl, R yl)
{
   EW* e = static_cast<EW*>(ew);
   if (corder)
      e->metricexgh_rev(ib, ie, jb, je, kb, ke, nz, x, y, z, met, jac, order, sb, zmax, amp, xc, yc, xl, yl);
   else
      e->metricexgh(ib, ie, jb, je, kb, ke, nz, x, y, z, met, jac, order, sb, zmax, amp, xc, yc, xl, yl);
}

// The analytic solution of the run's first source on grid g at time t, over
// the grid's whole box, in the object's layout (EW::get_exact_point_source).
void refew_exact_point_source(void* ew, int g, R t, R* up)
{
   EW* e = static_cast<EW*>(ew);
   e->get_exact_point_source(up, t, g, *e->m_globalUniqueSources[0]);
}

// The point-source error of the object's current U at time t, as the end of
// EW::timesteploop computes it (EW.C:2895-2905): exactSol, normOfDifference.
void refew_point_source_error(void* ew, R t, R* li, R* l2, R* xinf)
{
   EW* e = static_cast<EW*>(ew);
   std::vector<Sarray> uex(e->mNumberOfGrids);
   for (int g = 0; g < e->mNumberOfGrids; g++) uex[g].define(e->mU[g]);
   e->exactSol(t, uex, e->m_globalUniqueSources);
   e->normOfDifference(uex, e->mU, *li, *l2, *xinf, e->m_globalUniqueSources);
}

// Free time functions (time_functions.C); kind 0 Ricker, 1 Gaussian,
// 2 C6SmoothBump; deriv2 selects the second derivative.  par[0] is the
// exponent cut-off the reference sets (Source.C:2183-2187).
R refew_timefunc(int kind, int deriv2, R freq, R t, R par0)
{
   R par[2] = {par0, 0.0};
   int ipar[1] = {0};
   switch (kind) {
      case 0: return deriv2 ? RickerWavelet_tt(freq, t, par, 1, ipar, 0) : RickerWavelet(freq, t, par, 1, ipar, 0);
      case 1: return deriv2 ? Gaussian_tt(freq, t, par, 1, ipar, 0) : Gaussian(freq, t, par, 1, ipar, 0);
      case 2: return deriv2 ? C6SmoothBump_tt(freq, t, par, 1, ipar, 0) : C6SmoothBump(freq, t, par, 1, ipar, 0);
   }
   return 0.0;
}

// SuperGrid taper of one axis (SuperGrid.C:54-194): stretching, damping
// coefficient and corner taper at n coordinates x.
void refew_supergrid_taper(int left, R x0, int right, R x1, R width, int n, const R* x, R* str, R* dc, R* corner)
{
   SuperGrid s;
   s.define_taper(left != 0, x0, right != 0, x1, width);
   for (int q = 0; q < n; q++) {
      str[q] = s.stretching(x[q]);
      dc[q] = s.dampingCoeff(x[q]);
      corner[q] = s.cornerTaper(x[q]);
   }
}

}  // extern "C"
