// topo.cpp -- set-up of the curvilinear grid under a Gaussian-hill
// topography stacked on the Cartesian grid (see topo.h for the reference
// routines each function follows).
#include "topo.h"

#include <cmath>

#include "sw4hip_internal.h"

namespace sw4amd {

namespace {
constexpr int kGhost = 2;     // m_ghost_points
constexpr int kExtGhost = 2;  // m_ext_ghost_points: the sampled hill reaches 2 points past the ghosts
const double kNormWgh[4] = {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0};

// Interpolation weights of the 8-point stencil i-3..i+4 at offset a in
// [0, 1) and their derivative in a (Source::getmetwgh, Source.C:682-720;
// the values are EW::gettopowgh, EW.C:5667-5691).  The polynomials fix
// where the metric at a source is evaluated, so they are kept term for term.
void metric_weights(double a, double w[8], double dw[8])
{
   double pol = a * a * a * a * a * a * a * (-251 + 135 * a + 25 * a * a - 33 * a * a * a + 6 * a * a * a * a) / 720;
   w[0] = -1.0 / 60 * a + 1.0 / 180 * a * a + 1.0 / 48 * a * a * a + 23.0 / 144 * a * a * a * a -
          (17.0 * a + 223.0) * a * a * a * a * a / 720 - pol;
   w[1] = 3.0 / 20 * a - 3.0 / 40 * a * a - 1.0 / 6 * a * a * a - 13.0 / 12 * a * a * a * a +
          97.0 / 45 * a * a * a * a * a + 1.0 / 6 * a * a * a * a * a * a + 7 * pol;
   w[2] = -0.75 * a + 0.75 * a * a + (13.0 + 155 * a) * a * a * a / 48 - 103.0 / 16 * a * a * a * a * a -
          121.0 / 240 * a * a * a * a * a * a - 21 * pol;
   w[3] = 1 - 49.0 / 36 * a * a - 49.0 / 9 * a * a * a * a + 385.0 / 36 * a * a * a * a * a +
          61.0 / 72 * a * a * a * a * a * a + 35 * pol;
   w[4] = 0.75 * a + 0.75 * a * a - 13.0 / 48 * a * a * a + 89.0 / 16 * a * a * a * a -
          1537.0 / 144 * a * a * a * a * a - 41.0 / 48 * a * a * a * a * a * a - 35 * pol;
   w[5] = -3.0 / 20 * a - 3.0 / 40 * a * a + 1.0 / 6 * a * a * a - 41.0 / 12 * a * a * a * a +
          6.4 * a * a * a * a * a + 31.0 / 60 * a * a * a * a * a * a + 21 * pol;
   w[6] = 1.0 / 60 * a + 1.0 / 180 * a * a - 1.0 / 48 * a * a * a + 167.0 / 144 * a * a * a * a -
          1537.0 / 720 * a * a * a * a * a - 25.0 / 144 * a * a * a * a * a * a - 7 * pol;
   w[7] = -1.0 / 6 * a * a * a * a + 11.0 / 36 * a * a * a * a * a + 1.0 / 40 * a * a * a * a * a * a + pol;

   pol = a * a * a * a * a * a * (-1757.0 / 720 + 1.5 * a + 0.31250 * a * a - (1.375 * a * a * a - 0.275 * a * a * a * a) / 3);
   dw[0] = -1.0 / 60 + 1.0 / 90 * a + a * a / 16 + 23.0 / 36 * a * a * a - 223.0 / 144 * a * a * a * a -
           17.0 / 120 * a * a * a * a * a - pol;
   dw[1] = 3.0 / 20 - 3.0 / 20 * a - 0.5 * a * a - 13.0 / 3 * a * a * a + 97.0 / 9 * a * a * a * a +
           a * a * a * a * a + 7 * pol;
   dw[2] = -0.75 + 1.5 * a + 13.0 / 16 * a * a + 155.0 * a * a * a / 12 - 103.0 * 5.0 / 16 * a * a * a * a -
           121.0 / 40 * a * a * a * a * a - 21 * pol;
   dw[3] = -49.0 / 18 * a - 4 * 49.0 / 9.0 * a * a * a + 385.0 * 5.0 / 36 * a * a * a * a +
           61.0 / 12 * a * a * a * a * a + 35 * pol;
   dw[4] = 0.75 + 1.5 * a - 13.0 / 16 * a * a + 89.0 / 4 * a * a * a - 1537.0 * 5 / 144.0 * a * a * a * a -
           41.0 / 8 * a * a * a * a * a - 35 * pol;
   dw[5] = -3.0 / 20 - 3.0 / 20 * a + 0.5 * a * a - 41.0 / 3 * a * a * a + 32 * a * a * a * a +
           3.1 * a * a * a * a * a + 21 * pol;
   dw[6] = 1.0 / 60 + 1.0 / 90 * a - 1.0 / 16 * a * a + 167.0 / 36 * a * a * a - 1537.0 / 144 * a * a * a * a -
           25.0 / 24 * a * a * a * a * a - 7 * pol;
   dw[7] = -2.0 / 3 * a * a * a + 55.0 / 36 * a * a * a * a + 3.0 / 20 * a * a * a * a * a + pol;
}

// (1 - a)^n as the reference forms it, one factor at a time
double pow_by_factors(double a, int n)
{
   double p = 1 - a;
   for (int l = 2; l <= n; l++) p *= (1 - a);
   return p;
}

bool in_topo_domain(const StackedGrids& s, double q, double r)
{
   return q >= 1 - kGhost && q <= s.nx + kGhost && r >= 1 - kGhost && r <= s.ny + kGhost;
}

// z of the curvilinear grid at index coordinates (EW::grid_mapping, EW.C:5694-5719)
double grid_mapping_z(const Topography& t, const StackedGrids& s, double q, double r, double sk)
{
   if (!in_topo_domain(s, q, r)) return -1e38;
   const double ztopo = -t.elevation(q, r, s.h);
   const int nz = s.nz_curv;
   const double h = s.h;
   const double izb = 1.0 / (t.zeta_break * (nz - 1));
   const double sa = (sk - 1) * izb;
   const double omsm = pow_by_factors(sa, t.order);
   if (sa >= 1) return t.zmax - (nz - sk) * h;
   return t.zmax - (nz - sk) * h - omsm * (t.zmax - (nz - 1) * h - ztopo);
}

// z derivatives of the curvilinear mapping at a source, in index units
// (Source::compute_metric_at_source with the analytic k-derivative,
// Source.C:2191-2330): zq, zr, zs
void metric_at_source(const Topography& t, const StackedGrids& s, double q, double r, double sk, int ic, int jc,
                      int kc, double& zq, double& zr, double& zs)
{
   const int Nz = s.nz_curv;
   const double h = s.h;
   zq = zr = 0;
   zs = h;
   if (kc > Nz - 3) return;  // the lowest grid lines are treated as Cartesian
   const double zpar = (sk - 1) / (t.zeta_break * (Nz - 1));
   const double kbreak = 1 + t.zeta_break * (Nz - 1);
   if (zpar >= 1) return;
   double a6i[8], d6i[8], a6j[8], d6j[8];
   metric_weights(q - ic, a6i, d6i);
   metric_weights(r - jc, a6j, d6j);
   const double pp = std::pow(1 - zpar, t.order - 1);
   const double powo = (1 - zpar) * pp;
   const double dpowo = -t.order * pp / t.zeta_break;
   double tauavg = 0, tauq = 0, taur = 0;
   for (int j = jc - 3; j <= jc + 4; j++)
      for (int i = ic - 3; i <= ic + 4; i++) {
         const double tau = t.sample(i, j, h);
         tauavg += a6i[i - (ic - 3)] * a6j[j - (jc - 3)] * tau;
         tauq += d6i[i - (ic - 3)] * a6j[j - (jc - 3)] * tau;
         taur += a6i[i - (ic - 3)] * d6j[j - (jc - 3)] * tau;
      }
   zq = (-tauq) * powo;
   zr = (-taur) * powo;
   const double zmaxc = s.zmin_cart - (Nz - kbreak) * h;
   const double c1 = zmaxc + tauavg - h * (kbreak - 1);
   zs = h + c1 * (-dpowo) / (Nz - 1);
}

// EW::computeNearestGridPoint (EW.C:5149-5226) for the stacked grids:
// grid 0 (Cartesian) below zmin_cart, else grid 1 with k undefined
void nearest_point(const StackedGrids& s, double x, double y, double z, int& grid, int& i, int& j, int& k)
{
   const double h = s.h;
   i = (int)std::floor(x / h) + 1;
   if (x - ((i - 0.5) * h) > 0.) i++;
   j = (int)std::floor(y / h) + 1;
   if (y - ((j - 0.5) * h) > 0.) j++;
   if (z > s.zmin_cart) {
      grid = 0;
      k = (int)std::floor((z - s.zmin_cart) / h) + 1;
      if (z - (s.zmin_cart + (k - 0.5) * h) > 0.) k++;
   } else {
      grid = 1;
      k = 1;
   }
   if (z >= s.zmax) {
      k = s.nz_cart + kGhost;
      grid = 0;
   }
}
}  // namespace

double Topography::elevation(double q, double r, double h) const
{
   const double X0 = (q - 1.0) * h, Y0 = (r - 1.0) * h;
   const double igx2 = 1.0 / (lx * lx), igy2 = 1.0 / (ly * ly);
   return amp * std::exp(-(X0 - xc) * (X0 - xc) * igx2 - (Y0 - yc) * (Y0 - yc) * igy2);
}

double Topography::sample(int i, int j, double h) const
{
   const double x = (i - 1) * h, y = (j - 1) * h;
   const double a = (x - xc) / lx, b = (y - yc) / ly;
   return amp * std::exp(-(a) * (a) - (b) * (b));
}

int stacked_grid_sizes(const Topography& t, int nx, int ny, double h, double zdomain, StackedGrids& s)
{
   if (!(h > 0) || nx < 12 || ny < 12) return set_error(E_INVAL, "topography: grid too small");
   if (t.order < 2 || t.order > 7) return set_error(E_INVAL, "topography: order must be 2..7");
   if (!(t.zeta_break > 0 && t.zeta_break <= 1)) return set_error(E_INVAL, "topography: zetabreak must be in (0, 1]");
   if (!(t.lx > 0 && t.ly > 0)) return set_error(E_INVAL, "topography: gaussianLx and gaussianLy must be positive");
   // EW::compute_minmax_topography (EW.C:6127-6160): z = -elevation over the
   // sampled hill, which reaches kExtGhost points past the ghost points
   double zlo = 0, zhi = 0;
   bool first = true;
   for (int i = 1 - kGhost - kExtGhost; i <= nx + kGhost + kExtGhost; i++)
      for (int j = 1 - kGhost - kExtGhost; j <= ny + kGhost + kExtGhost; j++) {
         const double z = -t.sample(i, j, h);
         if (first || z > zhi) zhi = z;
         if (first || z < zlo) zlo = z;
         first = false;
      }
   const double topo_avg = 0.5 * (zlo + zhi);
   s.nx = nx;
   s.ny = ny;
   s.h = h;
   const int nz = static_cast<int>(1 + std::round((zdomain - t.zmax) / h));
   s.nz_cart = nz;
   s.zmax = t.zmax + (nz - 1) * h;
   s.zmin_cart = t.zmax;
   s.nz_curv = static_cast<int>(1 + std::round((t.zmax - topo_avg) / h));
   if (s.nz_cart < 12) return set_error(E_INVAL, "topography: the Cartesian grid below zmax needs at least 12 planes");
   if (s.nz_curv < 12) return set_error(E_INVAL, "topography: the curvilinear grid needs at least 12 planes");
   return OK;
}

void curvilinear_grid_z(const Topography& t, const StackedGrids& s, std::vector<double>& z)
{
   const long ni = s.nx + 4, nj = s.ny + 4, nk = s.nz_curv + 4;
   z.resize(ni * nj * nk);
   for (int k = -1; k <= s.nz_curv + 2; k++)
      for (int j = -1; j <= s.ny + 2; j++)
         for (int i = -1; i <= s.nx + 2; i++)
            z[(i + 1) + ni * (j + 1) + ni * nj * (k + 1)] = grid_mapping_z(t, s, i, j, k);
}

bool topo_surface_z(const Topography& t, const StackedGrids& s, double x, double y, double& z)
{
   const double q = x / s.h + 1.0, r = y / s.h + 1.0;
   if (!in_topo_domain(s, q, r)) {
      z = 0;
      return false;
   }
   z = -t.elevation(q, r, s.h);
   return true;
}

bool curvilinear_invert(const Topography& t, const StackedGrids& s, double x, double y, double z, double& q,
                        double& r, double& sk)
{
   const int maxit = 10;
   const double tol = 1e-9;
   const double h = s.h;
   const int nz = s.nz_curv;
   q = x / h + 1;
   r = y / h + 1;
   if (!in_topo_domain(s, q, r)) {
      sk = -1e38;
      return false;
   }
   const double ztopo = -t.elevation(q, r, h);
   const double izb = 1.0 / t.zeta_break;
   // normalised s, 0 at the surface and 1 at the common plane
   double sun = 1 - (t.zmax - z) / ((nz - 1) * h);
   if (sun >= t.zeta_break) {
      sk = (nz - 1) * sun + 1;
      return true;
   }
   int it = 0;
   double err = tol + 1;
   while (err > tol && it < maxit) {
      double omsm = 1 - izb * sun;
      for (int l = 2; l <= t.order - 1; l++) omsm *= (1 - izb * sun);
      const double fp = h * (nz - 1) + izb * t.order * omsm * (t.zmax - (nz - 1) * h - ztopo);
      omsm *= (1 - izb * sun);
      const double f = t.zmax - (nz - 1) * h * (1 - sun) - omsm * (t.zmax - (nz - 1) * h - ztopo) - z;
      const double ds = f / fp;
      err = std::fabs(ds);
      sun = sun - ds;
      it++;
   }
   sk = (nz - 1) * sun + 1;
   if (err >= tol) {
      sk = -1e38;
      return false;
   }
   return true;
}

int stacked_receiver_point(const Topography& t, const StackedGrids& s, double x, double y, double z, bool topodepth,
                           int& grid, int& i, int& j, int& k)
{
   nearest_point(s, x, y, z, grid, i, j, k);
   if (grid == 1 || topodepth) {
      double ztopo = 0;
      if (!topo_surface_z(t, s, x, y, ztopo)) return set_error(E_INVAL, "receiver off the topography grid");
      if (topodepth) z += ztopo;
      if (z < ztopo - 1.0e-9) return 1;  // above the surface: the reference ignores it
      // TimeSeries::computeNearestGridPointZ (TimeSeries.C:307-330)
      int i2, j2;
      nearest_point(s, x, y, z, grid, i2, j2, k);
      if (grid == 1) {
         double q, r, sk;
         if (curvilinear_invert(t, s, x, y, z, q, r, sk)) {
            k = static_cast<int>(std::round(sk));
            if (k < 1) k = 1;
            if (k > s.nz_curv) k = s.nz_curv;
         } else {
            k = 1;  // placed on the surface (TimeSeries.C:161-168)
         }
      }
   }
   const int nz = grid == 0 ? s.nz_cart : s.nz_curv;
   if (i < 1 || i > s.nx || j < 1 || j > s.ny || k < 1 || k > nz)
      return set_error(E_INVAL, "receiver outside the grid");
   return OK;
}

int stacked_moment_source(const Topography& t, const StackedGrids& s, const double* jac, const MomentSource& src,
                          std::vector<GridPointSource>& cart, std::vector<GridPointSource>& curv)
{
   const double h = s.h;
   double ztopo = 0;
   if (!topo_surface_z(t, s, src.x, src.y, ztopo)) return set_error(E_INVAL, "source off the topography grid");
   if (src.z < ztopo - 1e-9) return 1;  // above the surface: ignored (Source::adjust_zcoord)
   // grid and mapped coordinates (Source::compute_mapped_coordinates, Source.C:533-553)
   int grid, i0, j0, k0;
   nearest_point(s, src.x, src.y, src.z, grid, i0, j0, k0);
   double q0, r0, s0;
   if (grid == 0) {
      q0 = src.x / h + 1;
      r0 = src.y / h + 1;
      s0 = (src.z - s.zmin_cart) / h + 1;
   } else if (!curvilinear_invert(t, s, src.x, src.y, src.z, q0, r0, s0)) {
      return set_error(E_INVAL, "source: cannot invert the curvilinear grid mapping at the source");
   }
   const bool curvilinear = grid == 1;
   const int Ni = s.nx, Nj = s.ny, Nz = curvilinear ? s.nz_curv : s.nz_cart;
   // stencil position (Source.C:1155-1200)
   int ic = static_cast<int>(std::floor(q0)), jc = static_cast<int>(std::floor(r0)),
       kc = static_cast<int>(std::floor(s0));
   if (ic <= 2) ic = 3;
   if (ic >= Ni - 2) ic = Ni - 3;
   if (jc <= 2) jc = 3;
   if (jc >= Nj - 2) jc = Nj - 3;
   if (kc >= Nz) kc = Nz - 1;
   if (kc < 1) kc = 1;
   const bool upper = kc == 1 || kc == 2 || kc == 3;
   const bool lower = kc == Nz - 1 || kc == Nz - 2 || kc == Nz - 3;
   // the stencil reaches the plane shared by the two grids
   const bool ccbndry = (upper && !curvilinear) || (lower && curvilinear);
   if (!ccbndry) {
      if (kc <= 2) kc = 3;
      if (kc >= Nz - 2) kc = Nz - 3;
   }
   const double ai = q0 - ic, bi = r0 - jc, ci = s0 - kc;
   double wi[6], wj[6], wk[6], dwi[6], dwj[6], dwk[6];
   source_delta_weights(ai, wi);
   source_delta_weights(bi, wj);
   source_delta_weights(ci, wk);
   source_ddelta_weights(ai, dwi);
   source_ddelta_weights(bi, dwj);
   source_ddelta_weights(ci, dwk);
   if (!ccbndry && kc == 3 && ci <= 0) {
      source_delta_weights_low(ci, wk);
      source_ddelta_weights_low(ci, dwk);
   }
   // SBP norm at the one-sided (free-surface) top of the curvilinear grid
   if (curvilinear)
      for (int m = 0; m <= 5; m++)
         if (1 <= m + kc - 2 && m + kc - 2 <= 4) {
            wk[m] /= kNormWgh[m + kc - 3];
            dwk[m] /= kNormWgh[m + kc - 3];
         }
   // gradient of the index coordinates with respect to (x, y, z)
   double qx = 1 / h, ry = 1 / h, sx = 0, sy = 0, sz = 1 / h;
   if (curvilinear) {
      double zq, zr, zs;
      metric_at_source(t, s, q0, r0, s0, ic, jc, kc, zq, zr, zs);
      sx = -zq / (h * zs);
      sy = -zr / (h * zs);
      sz = 1 / zs;
   }
   const long ni = s.nx + 4, nij = ni * (s.ny + 4);
   const double* M = src.m;  // xx, xy, xz, yy, yz, zz
   for (int k = kc - 2; k <= kc + 3; k++)
      for (int j = jc - 2; j <= jc + 3; j++)
         for (int i = ic - 2; i <= ic + 3; i++) {
            const int a = i - ic + 2, b = j - jc + 2, c = k - kc + 2;
            double wFx = qx * dwi[a] * wj[b] * wk[c];
            double wFy = wi[a] * ry * dwj[b] * wk[c];
            wFx += wi[a] * wj[b] * sx * dwk[c];
            wFy += wi[a] * wj[b] * sy * dwk[c];
            const double wFz = wi[a] * wj[b] * sz * dwk[c];
            const double jaci = curvilinear ? 1 / jac[(i + 1) + ni * (j + 1) + nij * (k + 1)] : 1.0 / (h * h * h);
            GridPointSource p;
            p.i = i;
            p.j = j;
            p.f[0] = -(M[0] * wFx + M[1] * wFy + M[2] * wFz) * jaci;
            p.f[1] = -(M[1] * wFx + M[3] * wFy + M[4] * wFz) * jaci;
            p.f[2] = -(M[2] * wFx + M[4] * wFy + M[5] * wFz) * jaci;
            p.freq = src.freq;
            p.t0 = src.t0;
            p.tfkind = src.tfkind;
            std::vector<GridPointSource>& own = curvilinear ? curv : cart;
            if (1 <= k && k <= Nz) {
               p.k = k;
               own.push_back(p);
            }
            // the part of the stencil on the other side of the common plane
            // (k = 1 of the Cartesian grid is k = nz_curv of the curvilinear one)
            if (!curvilinear && ccbndry && k <= 1) {
               p.k = s.nz_curv - 1 + k;
               curv.push_back(p);
            }
            if (curvilinear && ccbndry && k >= Nz) {
               p.k = k - Nz + 1;
               cart.push_back(p);
            }
         }
   return OK;
}

}  // namespace sw4amd
