// setup.cpp -- host-side set-up of one Cartesian grid for the time step:
// supergrid taper profiles, ghost-point windows, one-sided flags and the
// grid-point discretisation of moment-tensor sources.  These are computed
// once on the host (as in the reference) and uploaded; the per-step work is
// all on the GPU.
#include "setup.h"

#include <algorithm>
#include <cmath>

namespace sw4amd {

// ---------------------------------------------------------------------------
// Supergrid taper, SuperGrid.C:54-194
// ---------------------------------------------------------------------------
void Taper::define(bool l, double start, bool r, double end, double w)
{
   left = l;
   x0 = start;
   right = r;
   x1 = end;
   width = w;
   trans = 0.5 * w;  // full transition, no constant-damping core (SuperGrid.C:62-67)
}

// C5 step 0 -> 1 on [0,1] (SuperGrid.C:165-187).  The products are formed
// left to right as the reference forms them: near psi = 1 the stretching
// 1 - (1 - eps) psi cancels to ~1e-4, so a different grouping of the
// polynomial changes the profiles in their ninth digit.
double Taper::psi0(double xi)
{
   if (xi <= 0.) return 0.;
   if (xi >= 1.) return 1.;
   const double p6 = xi * xi * xi * xi * xi * xi;
   return p6 * (462 - 1980 * xi + 3465 * xi * xi - 3080 * xi * xi * xi + 1386 * xi * xi * xi * xi -
                252 * xi * xi * xi * xi * xi);
}

double Taper::ramp(double x, double w) const
{
   if (left && x < x0 + width) return psi0((x0 + width - x) / w);
   if (right && x > x1 - width) return psi0((x - (x1 - width)) / w);
   return 0.;
}

double Taper::lin(double x) const
{
   if (left && x < x0 + width) return (x0 + width - x) / width;
   if (right && x > x1 - width) return (x - (x1 - width)) / width;
   return 0.;
}

double Taper::stretching(double x) const { return 1 - (1 - eps) * ramp(x, width); }
double Taper::damping(double x) const { return ramp(x, trans) / stretching(x); }
double Taper::corner(double x) const { return 1.0 - (1.0 - 0.33) * lin(x); }

// EW::setup_supergrid + assign_supergrid_damping_arrays (EW.C:4636-4801),
// one grid.  A curvilinear grid has neither supergrid side in z
// (EW.C:4728-4737): its z arrays stay 0, 1, 1.
void make_supergrid_arrays(const GridSetup& g, SupergridArrays& a)
{
   const bool use = g.bc[0] == BC_SUPERGRID || g.bc[1] == BC_SUPERGRID || g.bc[2] == BC_SUPERGRID ||
                    g.bc[3] == BC_SUPERGRID || g.bc[4] == BC_SUPERGRID || g.bc[5] == BC_SUPERGRID;
   // the z taper ends at the bottom of this grid, zmin + (nz - 1) h (a
   // Cartesian grid under a curvilinear one starts below z = 0)
   const double xmax = (g.nx - 1) * g.h, ymax = (g.ny - 1) * g.h, zmax = g.zmin + (g.nz - 1) * g.h;
   const double w = g.sg_gp_thickness * g.h;
   Taper tx, ty, tz;
   tx.define(g.bc[0] == BC_SUPERGRID, 0.0, g.bc[1] == BC_SUPERGRID, xmax, w);
   ty.define(g.bc[2] == BC_SUPERGRID, 0.0, g.bc[3] == BC_SUPERGRID, ymax, w);
   tz.define(g.bc[4] == BC_SUPERGRID, 0.0, g.bc[5] == BC_SUPERGRID, zmax, w);
   const int ni = g.nx + 4, nj = g.ny + 4, nk = g.nz + 4;
   a.dcx.assign(ni, 0.0); a.strx.assign(ni, 1.0); a.cox.assign(ni, 1.0);
   a.dcy.assign(nj, 0.0); a.stry.assign(nj, 1.0); a.coy.assign(nj, 1.0);
   a.dcz.assign(nk, 0.0); a.strz.assign(nk, 1.0); a.coz.assign(nk, 1.0);
   if (!use) return;
   for (int i = -1; i <= g.nx + 2; i++) {
      const double x = (i - 1) * g.h;
      a.dcx[i + 1] = tx.damping(x);
      a.strx[i + 1] = tx.stretching(x);
      a.cox[i + 1] = tx.corner(x);
   }
   for (int j = -1; j <= g.ny + 2; j++) {
      const double y = (j - 1) * g.h;
      a.dcy[j + 1] = ty.damping(y);
      a.stry[j + 1] = ty.stretching(y);
      a.coy[j + 1] = ty.corner(y);
   }
   for (int k = -1; k <= g.nz + 2; k++) {
      const double z = g.zmin + (k - 1) * g.h;
      a.dcz[k + 1] = tz.damping(z);
      a.strz[k + 1] = tz.stretching(z);
      a.coz[k + 1] = tz.corner(z);
   }
}

// ---------------------------------------------------------------------------
// Ghost-point windows, EW::setup_boundary_arrays / side_plane (EW.C:3347-3443)
// ---------------------------------------------------------------------------
void make_boundary_windows(const GridSetup& g, int wind[36], int npts[6])
{
   const int ib = -1, ie = g.nx + 2, jb = -1, je = g.ny + 2, kb = -1, ke = g.nz + 2;
   for (int s = 0; s < 6; s++) {
      int* w = wind + 6 * s;
      w[0] = w[2] = w[4] = 999;
      w[1] = w[3] = w[5] = -999;
      npts[s] = 0;
      const int bc = g.bc[s];
      if (!(bc == BC_FREE || bc == BC_DIRICHLET || bc == BC_SUPERGRID || bc == BC_PERIODIC)) continue;
      int v[6] = {ib, ie, jb, je, kb, ke};
      const int ng = bc == BC_FREE ? 1 : 2;
      const int lo = 2 * (s / 2), hi = lo + 1;  // index pair of the side's axis
      if (s % 2 == 0)
         v[hi] = v[lo] + (ng - 1);
      else
         v[lo] = v[hi] - (ng - 1);
      if (bc == BC_FREE) {
         // forcing lives on the boundary plane itself, just inside the ghosts
         if (s % 2 == 0) {
            v[lo] += 2;
            v[hi] = v[lo];
         } else {
            v[lo] -= 2;
            v[hi] = v[lo];
         }
      }
      for (int q = 0; q < 6; q++) w[q] = v[q];
      npts[s] = (v[5] - v[4] + 1) * (v[3] - v[2] + 1) * (v[1] - v[0] + 1);
   }
}

// ---------------------------------------------------------------------------
// Source discretisation, Source::set_grid_point_sources4 (Source.C:1074-2026)
// for a single Cartesian grid.
// ---------------------------------------------------------------------------

// Regularised delta, moments 0..4 exact (Source.C:553-584): values only.
void source_delta_weights(double a, double w[6])
{
   const double a2 = a * a, a3 = a2 * a, a4 = a3 * a, a5 = a4 * a;
   const double p5 = a5 * (5.0 / 3 - 7.0 / 24 * a - 17 / 12.0 * a2 + 1.125 * a3 - 0.25 * a4);
   w[0] = 1.0 / 24 * (2 * a - a2 - 2 * a3 - 19 * a4) + p5;
   w[1] = 1.0 / 6 * (-4 * a + 4 * a2 + a3) + 4 * a4 - 5 * p5;
   w[2] = 1 - 1.25 * a2 - 97.0 / 12 * a4 + 10 * p5;
   w[3] = 1.0 / 6 * (4 * a + 4 * a2 - a3 + 49 * a4) - 10 * p5;
   w[4] = 1.0 / 24 * (-2 * a - a2 + 2 * a3) - 4.125 * a4 + 5 * p5;
   w[5] = 5.0 / 6 * a4 - p5;
}

// Regularised delta' (Source.C:587-618): values only.
void source_ddelta_weights(double a, double w[6])
{
   const double a2 = a * a, a3 = a2 * a, a4 = a3 * a;
   const double p5 = a4 * (-25.0 / 12 - 0.75 * a + 59.0 / 12 * a2 - 4 * a3 + a4);
   w[0] = 1.0 / 12 * (-1 + a + 3 * a2 + 8 * a3) + p5;
   w[1] = 2.0 / 3 * (1 - 2 * a) - 0.5 * a2 - 3.5 * a3 - 5 * p5;
   w[2] = 2.5 * a + 22.0 / 3 * a3 + 10 * p5;
   w[3] = 2.0 / 3 * (-1 - 2 * a) + 0.5 * a2 - 23.0 / 3 * a3 - 10 * p5;
   w[4] = (1 + a) / 12 - 0.25 * a2 + 4 * a3 + 5 * p5;
   w[5] = -5.0 / 6 * a3 - p5;
}

// Lower-boundary stencils used when kc==3 and ci<=0 (Source.C:621-680).
void source_delta_weights_low(double a, double w[6])
{
   const double a2 = a * a, a3 = a2 * a, a4 = a3 * a;
   w[0] = (2 * a - a2 - 2 * a3 + a4) / 24;
   w[1] = (-4 * a + 4 * a2 + a3 - a4) / 6;
   w[2] = 1 - 1.25 * a2 + 0.25 * a4;
   w[3] = (4 * a + 4 * a2 - a3 - a4) / 6;
   w[4] = (-2 * a - a2 + 2 * a3 + a4) / 24;
   w[5] = 0;
}

void source_ddelta_weights_low(double a, double w[6])
{
   const double a2 = a * a, a3 = a2 * a;
   w[0] = (-1 + a + 3 * a2 - 2 * a3) / 12;
   w[1] = (2 - 4 * a - 1.5 * a2 + 2 * a3) / 3;
   w[2] = 2.5 * a - a3;
   w[3] = (-2 - 4 * a + 1.5 * a2 + 2 * a3) / 3;
   w[4] = (1 + a - 3 * a2 - 2 * a3) / 12;
   w[5] = 0;
}

int moment_source_points(const GridSetup& g, const MomentSource& src, std::vector<GridPointSource>& out)
{
   const double h = g.h;
   const int Ni = g.nx, Nj = g.ny, Nz = g.nz;
   if (Ni < 6 || Nj < 6 || Nz < 6) return -1;
   // mapped coordinates, EW::invert_grid_mapping (EW.C:5722-5733)
   const double q0 = src.x / h + 1, r0 = src.y / h + 1, s0 = (src.z - g.zmin) / h + 1;
   int ic = (int)std::floor(q0), jc = (int)std::floor(r0), kc = (int)std::floor(s0);
   // keep the 6-point stencil away from ghost points (Source.C:1162-1204)
   if (ic <= 2) ic = 3;
   if (ic >= Ni - 2) ic = Ni - 3;
   if (jc <= 2) jc = 3;
   if (jc >= Nj - 2) jc = Nj - 3;
   if (kc >= Nz) kc = Nz - 1;
   if (kc < 1) kc = 1;
   if (kc <= 2) kc = 3;
   if (kc >= Nz - 2) kc = Nz - 3;
   const double ai = q0 - ic, bi = r0 - jc, ci = s0 - kc;
   double wi[6], wj[6], wk[6], dwi[6], dwj[6], dwk[6];
   source_delta_weights(ai, wi);
   source_delta_weights(bi, wj);
   source_delta_weights(ci, wk);
   source_ddelta_weights(ai, dwi);
   source_ddelta_weights(bi, dwj);
   source_ddelta_weights(ci, dwk);
   if (kc == 3 && ci <= 0) {
      source_delta_weights_low(ci, wk);
      source_ddelta_weights_low(ci, dwk);
   }
   // SBP norm weights near a one-sided boundary (Source.C:1085, 1411-1442)
   static const double normwgh[4] = {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0};
   if (g.bc[4] == BC_FREE) {
      for (int m = 0; m <= 5; m++) {
         const int kk = m + kc - 2;
         if (kk >= 1 && kk <= 4) {
            wk[m] /= normwgh[kk - 1];
            dwk[m] /= normwgh[kk - 1];
         }
      }
   }
   if (g.bc[5] == BC_FREE) {
      for (int m = 0; m <= 5; m++) {
         const int kk = m + kc - 2;
         if (kk >= Nz - 3 && kk <= Nz) {
            wk[m] /= normwgh[Nz - kk];
            dwk[m] /= normwgh[Nz - kk];
         }
      }
   }
   // forces from the divergence of the moment tensor against delta'
   // (Source.C:1724-1733, 1798-1800)
   const double hi = 1 / h, jaci = 1.0 / (h * h * h);
   const double* M = src.m;  // xx, xy, xz, yy, yz, zz
   for (int k = kc - 2; k <= kc + 3; k++)
      for (int j = jc - 2; j <= jc + 3; j++)
         for (int i = ic - 2; i <= ic + 3; i++) {
            const int a = i - ic + 2, b = j - jc + 2, c = k - kc + 2;
            const double wFx = hi * dwi[a] * wj[b] * wk[c];
            const double wFy = wi[a] * hi * dwj[b] * wk[c];
            const double wFz = wi[a] * wj[b] * hi * dwk[c];
            GridPointSource p;
            p.i = i;
            p.j = j;
            p.k = k;
            p.f[0] = -(M[0] * wFx + M[1] * wFy + M[2] * wFz) * jaci;
            p.f[1] = -(M[1] * wFx + M[3] * wFy + M[4] * wFz) * jaci;
            p.f[2] = -(M[2] * wFx + M[4] * wFy + M[5] * wFz) * jaci;
            p.freq = src.freq;
            p.t0 = src.t0;
            p.tfkind = src.tfkind;
            if (k >= 1 && k <= Nz) out.push_back(p);
         }
   return 0;
}

// ---------------------------------------------------------------------------
// Moment source on a curvilinear grid (Source.C:1605-1700 with the forcing
// of :1724-1800 and jaci = 1/J(i,j,k)).  The grid is given by its depth z
// over the box; the reference evaluates z and its derivatives at the source
// from its analytic grid mapping (Source.C:2191-2330), here they come from
// the grid itself: the 4th-order grid derivatives (as in the metric)
// interpolated to the source with the 6-point delta weights, which are exact
// for polynomials of degree <= 4.
// ---------------------------------------------------------------------------
namespace {

// derivative along a grid line, same closures as the metric
// (curvilinear-c.C:78-146); f(t) = f[t * st], position s of n
double line_deriv(const double* f, long st, int s, int n)
{
   const double c1 = 2.0 / 3, c2 = -1.0 / 12;
   auto F = [&](int t) { return f[(long)t * st]; };
   if (s >= 2 && s <= n - 3) return c2 * (F(s + 2) - F(s - 2)) + c1 * (F(s + 1) - F(s - 1));
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
}

struct CurvGrid {
   const GridSetup& g;
   const double* z;
   int ni, nj, nk;
   long nij;
   long p(int i, int j, int k) const { return (i + 1) + (long)ni * (j + 1) + nij * (k + 1); }
   double Z(int i, int j, int k) const { return z[p(i, j, k)]; }
   double zp(int i, int j, int k) const { return line_deriv(z + p(-1, j, k), 1, i + 1, ni); }
   double zq(int i, int j, int k) const { return line_deriv(z + p(i, -1, k), ni, j + 1, nj); }
   double zr(int i, int j, int k) const { return line_deriv(z + p(i, j, -1), nij, k + 1, nk); }
};

// Interpolated grid depth and k-derivative at (q0, r0, s) with the delta
// weights centred at (ic, jc, kc).
void column_at(const CurvGrid& G, const double wi[6], const double wj[6], int ic, int jc, double s, double& z,
               double& zs)
{
   const int kc = std::min(std::max((int)std::floor(s), 1), G.g.nz - 1);
   double wk[6];
   source_delta_weights(s - kc, wk);
   z = zs = 0;
   for (int c = 0; c < 6; c++) {
      const int k = kc - 2 + c;
      if (k < -1 || k > G.g.nz + 2) continue;
      for (int b = 0; b < 6; b++)
         for (int a = 0; a < 6; a++) {
            const double w = wi[a] * wj[b] * wk[c];
            z += w * G.Z(ic - 2 + a, jc - 2 + b, k);
            zs += w * G.zr(ic - 2 + a, jc - 2 + b, k);
         }
   }
}

}  // namespace

int moment_source_points_curv(const GridSetup& g, const double* zgrid, const MomentSource& src,
                              std::vector<GridPointSource>& out)
{
   const double h = g.h;
   const int Ni = g.nx, Nj = g.ny, Nz = g.nz;
   if (Ni < 6 || Nj < 6 || Nz < 6) return -1;
   CurvGrid G{g, zgrid, Ni + 4, Nj + 4, Nz + 4, (long)(Ni + 4) * (Nj + 4)};
   const double q0 = src.x / h + 1, r0 = src.y / h + 1;
   int ic = (int)std::floor(q0), jc = (int)std::floor(r0);
   if (ic <= 2) ic = 3;
   if (ic >= Ni - 2) ic = Ni - 3;
   if (jc <= 2) jc = 3;
   if (jc >= Nj - 2) jc = Nj - 3;
   const double ai = q0 - ic, bi = r0 - jc;
   double wi[6], wj[6], dwi[6], dwj[6];
   source_delta_weights(ai, wi);
   source_delta_weights(bi, wj);
   source_ddelta_weights(ai, dwi);
   source_ddelta_weights(bi, dwj);
   // s0 with z(q0, r0, s0) = src.z: bracket on the interpolated column, then
   // Newton on the delta-weight interpolant (EW::invert_curvilinear_grid_mapping)
   double s0 = -1;
   {
      double zlo = 0, dum = 0;
      column_at(G, wi, wj, ic, jc, 1.0, zlo, dum);
      if (src.z < zlo) return -2;  // above the surface
      for (int k = 1; k < Nz; k++) {
         double za, zb;
         column_at(G, wi, wj, ic, jc, (double)k, za, dum);
         column_at(G, wi, wj, ic, jc, (double)(k + 1), zb, dum);
         if (src.z >= za && src.z <= zb) {
            s0 = k + (src.z - za) / (zb - za);
            break;
         }
      }
      if (s0 < 0) return -2;  // below the grid
      for (int it = 0; it < 8; it++) {
         double zz, zs;
         column_at(G, wi, wj, ic, jc, s0, zz, zs);
         if (!(zs > 0)) return -3;
         const double ds = (zz - src.z) / zs;
         s0 = std::min(std::max(s0 - ds, 1.0), (double)Nz);
         if (std::fabs(ds) < 1e-14) break;
      }
   }
   int kc = (int)std::floor(s0);
   if (kc >= Nz) kc = Nz - 1;
   if (kc < 1) kc = 1;
   if (kc <= 2) kc = 3;
   if (kc >= Nz - 2) kc = Nz - 3;
   const double ci = s0 - kc;
   double wk[6], dwk[6];
   source_delta_weights(ci, wk);
   source_ddelta_weights(ci, dwk);
   if (kc == 3 && ci <= 0) {
      source_delta_weights_low(ci, wk);
      source_ddelta_weights_low(ci, dwk);
   }
   // grid derivatives at the source (index units), interpolated before the
   // SBP norm corrections are applied to the k weights
   double zq = 0, zr = 0, zs = 0;
   {
      double wk0[6];
      source_delta_weights(ci, wk0);
      if (kc == 3 && ci <= 0) source_delta_weights_low(ci, wk0);
      for (int c = 0; c < 6; c++)
         for (int b = 0; b < 6; b++)
            for (int a = 0; a < 6; a++) {
               const int i = ic - 2 + a, j = jc - 2 + b, k = kc - 2 + c;
               const double w = wi[a] * wj[b] * wk0[c];
               zq += w * G.zp(i, j, k);
               zr += w * G.zq(i, j, k);
               zs += w * G.zr(i, j, k);
            }
   }
   if (!(zs > 0)) return -3;
   static const double normwgh[4] = {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0};
   if (g.bc[4] == BC_FREE)
      for (int m = 0; m <= 5; m++) {
         const int kk = m + kc - 2;
         if (kk >= 1 && kk <= 4) {
            wk[m] /= normwgh[kk - 1];
            dwk[m] /= normwgh[kk - 1];
         }
      }
   // gradient of s with respect to (x, y, z) at the source (Source.C:1646-1651)
   const double sx0 = -zq / (h * zs), sy0 = -zr / (h * zs), sz0 = 1 / zs;
   const double hi = 1 / h;
   const double* M = src.m;
   for (int k = kc - 2; k <= kc + 3; k++)
      for (int j = jc - 2; j <= jc + 3; j++)
         for (int i = ic - 2; i <= ic + 3; i++) {
            const int a = i - ic + 2, b = j - jc + 2, c = k - kc + 2;
            const double wFx = hi * dwi[a] * wj[b] * wk[c] + wi[a] * wj[b] * sx0 * dwk[c];
            const double wFy = wi[a] * hi * dwj[b] * wk[c] + wi[a] * wj[b] * sy0 * dwk[c];
            const double wFz = wi[a] * wj[b] * sz0 * dwk[c];
            // J = h^2 z_r (curvilinear-c.C:159)
            const double jaci = 1.0 / (h * h * G.zr(i, j, k));
            GridPointSource p;
            p.i = i;
            p.j = j;
            p.k = k;
            p.f[0] = -(M[0] * wFx + M[1] * wFy + M[2] * wFz) * jaci;
            p.f[1] = -(M[1] * wFx + M[3] * wFy + M[4] * wFz) * jaci;
            p.f[2] = -(M[2] * wFx + M[4] * wFy + M[5] * wFz) * jaci;
            p.freq = src.freq;
            p.t0 = src.t0;
            p.tfkind = src.tfkind;
            if (k >= 1 && k <= Nz) out.push_back(p);
         }
   return 0;
}

// EW::sort_grid_point_sources (EW.C:6370-6433): order by grid index and mark
// runs that share one point.  A stable sort keeps the input order within a
// run, so the summation order at a shared point is deterministic.
void sort_grid_point_sources(const GridSetup& g, std::vector<GridPointSource>& src,
                             std::vector<int>& first)
{
   const long ni = g.nx + 4, nij = ni * (g.ny + 4);
   auto key = [&](const GridPointSource& p) { return (p.i + 1) + ni * (p.j + 1) + nij * (p.k + 1); };
   std::stable_sort(src.begin(), src.end(),
                    [&](const GridPointSource& a, const GridPointSource& b) { return key(a) < key(b); });
   first.clear();
   for (size_t s = 0; s < src.size(); s++)
      if (s == 0 || key(src[s]) != key(src[s - 1])) first.push_back((int)s);
   first.push_back((int)src.size());
}

}  // namespace sw4amd
