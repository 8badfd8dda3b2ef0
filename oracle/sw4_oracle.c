/*
 * sw4_oracle.c -- CPU restatement (C99 + OpenMP, FP64) of the sw4lite
 * elastic-wave time-step kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library in sw4lite_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Every function restates the algorithm of a reference routine (cited
 * file:line, paths relative to the sw4lite source tree) in code of its own.
 * Array conventions are the reference's:
 *   - index boxes [ifirst,ilast]x[jfirst,jlast]x[kfirst,klast] include the
 *     two ghost points on every side;
 *   - 3-component arrays are either component-interleaved (corder=0,
 *     u[c + 3*(i + ni*j + nij*k)], Sarray.h:84-135) or component-blocked
 *     (corder=1, u[i + ni*j + nij*k + npts*c], Sarray.C:753-778);
 *   - 1-D supergrid arrays start at ifirst/jfirst/kfirst;
 *   - SBP tables acof[6][8][8] (row fastest), bope[8][6], ghcof[6], sbop[5]
 *     as produced by GetStencilCoefficients (ew-cfromfort.C:1487-1944).
 *
 * Pinning (see tests/): the kernels are compared with the reference sources
 * compiled by oracle/Makefile (rhs4sg.C, rhs4sg_rev.C, tests/testil/
 * addsgd4fort*.C) and with committed golden vectors made from them; the
 * pointwise updates, the ghost-point closure, the forcing scatter and the
 * time functions are pinned by closed forms and invariants in
 * tests/test_oracle_*.py; the whole chain is pinned by the reference's
 * point-source error fixture (PointSourceErr.txt).
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* layout helpers                                                      */
/* ------------------------------------------------------------------ */
typedef struct {
   int ib, ie, jb, je, kb, ke;
   ptrdiff_t ni, nij, npts;
   int corder;
} box_t;

static box_t mkbox(int ib, int ie, int jb, int je, int kb, int ke, int corder)
{
   box_t b;
   b.ib = ib; b.ie = ie; b.jb = jb; b.je = je; b.kb = kb; b.ke = ke;
   b.ni = ie - ib + 1;
   b.nij = b.ni * (je - jb + 1);
   b.npts = b.nij * (ke - kb + 1);
   b.corder = corder;
   return b;
}

/* scalar field offset */
static inline ptrdiff_t P(const box_t *b, int i, int j, int k)
{
   return (i - b->ib) + b->ni * (j - b->jb) + b->nij * (k - b->kb);
}
/* vector field offset, c = 0,1,2 */
static inline ptrdiff_t V(const box_t *b, int c, int i, int j, int k)
{
   ptrdiff_t p = P(b, i, j, k);
   return b->corder ? p + b->npts * c : 3 * p + c;
}

/* 4th-order centred first difference, times 12h:
 * f(-2) - f(+2) + 8 (f(+1) - f(-1)).  Used for every mixed derivative
 * (rhs4sg_rev.C:226-336). */
static inline double d4(double fm2, double fm1, double fp1, double fp2)
{
   return fm2 - fp2 + 8 * (fp1 - fm1);
}

/* Coefficients of the 4th-order variable-coefficient second difference
 * (a u_x)_x for a 5-point window a[0..4] = a(i-2..i+2)  (rhs4sg_rev.C:120-145). */
static inline void varcoef4(const double a[5], double m[4])
{
   m[0] = a[1] - 0.75 * (a[2] + a[0]);
   m[1] = a[0] + a[3] + 3 * (a[2] + a[1]);
   m[2] = a[1] + a[4] + 3 * (a[3] + a[2]);
   m[3] = a[3] - 0.75 * (a[2] + a[4]);
}

/* sum_q m[q] (f(q) - f(centre)) over the window f[0..4] */
static inline double varsum4(const double m[4], const double f[5])
{
   return m[0] * (f[0] - f[2]) + m[1] * (f[1] - f[2]) + m[2] * (f[3] - f[2]) +
          m[3] * (f[4] - f[2]);
}

#define ACOF(a, k, q, m) (a)[((k)-1) + 6 * ((q)-1) + 48 * ((m)-1)]
#define BOPE(b, k, q) (b)[((k)-1) + 6 * ((q)-1)]

/* ------------------------------------------------------------------ */
/* Cartesian divergence of stress, rhs4sg.C:38-849 / rhs4sg_rev.C:43-864 */
/* ------------------------------------------------------------------ */

/* In-plane (x and y) second-derivative parts of r1,r2,r3, already with the
 * 1/6 factor, plus the z part when withz != 0.  rhs4sg_rev.C:120-221. */
static void second_derivs(const box_t *b, int i, int j, int k, const double *u,
                          const double *mu, const double *la, const double *strx,
                          const double *stry, const double *strz, int withz, double r[3])
{
   double am[5], al[5], mx[4], lx[4], my[4], ly[4], mz[4], lz[4], cx[4], cy[4], cz[4];
   double f[5];
   int s, q;
   const double sx = strx[i - b->ib], sy = stry[j - b->jb];
   for (s = 0; s < 5; s++) {
      am[s] = mu[P(b, i + s - 2, j, k)] * strx[i + s - 2 - b->ib];
      al[s] = la[P(b, i + s - 2, j, k)] * strx[i + s - 2 - b->ib];
   }
   varcoef4(am, mx);
   varcoef4(al, lx);
   for (s = 0; s < 5; s++) {
      am[s] = mu[P(b, i, j + s - 2, k)] * stry[j + s - 2 - b->jb];
      al[s] = la[P(b, i, j + s - 2, k)] * stry[j + s - 2 - b->jb];
   }
   varcoef4(am, my);
   varcoef4(al, ly);
   if (withz) {
      for (s = 0; s < 5; s++) {
         am[s] = mu[P(b, i, j, k + s - 2)] * strz[k + s - 2 - b->kb];
         al[s] = la[P(b, i, j, k + s - 2)] * strz[k + s - 2 - b->kb];
      }
      varcoef4(am, mz);
      varcoef4(al, lz);
   }
   for (q = 0; q < 4; q++) {
      cx[q] = 2 * mx[q] + lx[q];
      cy[q] = 2 * my[q] + ly[q];
      if (withz) cz[q] = 2 * mz[q] + lz[q];
   }
   for (int c = 0; c < 3; c++) {
      double acc;
      for (s = 0; s < 5; s++) f[s] = u[V(b, c, i + s - 2, j, k)];
      acc = sx * varsum4(c == 0 ? cx : mx, f);
      for (s = 0; s < 5; s++) f[s] = u[V(b, c, i, j + s - 2, k)];
      acc += sy * varsum4(c == 1 ? cy : my, f);
      if (withz) {
         for (s = 0; s < 5; s++) f[s] = u[V(b, c, i, j, k + s - 2)];
         acc += strz[k - b->kb] * varsum4(c == 2 ? cz : mz, f);
      }
      r[c] = acc / 6;
   }
}

/* (coef * D_b u_c)_a with centred 4th-order differences in both directions,
 * times 144 h^2.  dir a/b: 0=i,1=j,2=k; coef is la or mu sampled along a. */
static double mixed(const box_t *b, int i, int j, int k, const double *coef, const double *u,
                    int c, int da, int db)
{
   double g[4];
   static const int off[4] = {-2, -1, 1, 2};
   for (int s = 0; s < 4; s++) {
      int ia = i + (da == 0) * off[s], ja = j + (da == 1) * off[s], ka = k + (da == 2) * off[s];
      double fb[4];
      for (int t = 0; t < 4; t++) {
         int ib = ia + (db == 0) * off[t], jb = ja + (db == 1) * off[t], kb = ka + (db == 2) * off[t];
         fb[t] = u[V(b, c, ib, jb, kb)];
      }
      g[s] = coef[P(b, ia, ja, ka)] * d4(fb[0], fb[1], fb[2], fb[3]);
   }
   return d4(g[0], g[1], g[2], g[3]);
}

static void rhs_interior_point(const box_t *b, int i, int j, int k, const double *u,
                               const double *mu, const double *la, const double *strx,
                               const double *stry, const double *strz, double r[3])
{
   const double i144 = 1.0 / 144;
   const double sx = strx[i - b->ib], sy = stry[j - b->jb], sz = strz[k - b->kb];
   second_derivs(b, i, j, k, u, mu, la, strx, stry, strz, 1, r);
   /* rhs4sg_rev.C:226-336 */
   r[0] += sx * sy * i144 * mixed(b, i, j, k, la, u, 1, 0, 1)   /* (la v_y)_x */
         + sx * sz * i144 * mixed(b, i, j, k, la, u, 2, 0, 2)   /* (la w_z)_x */
         + sx * sy * i144 * mixed(b, i, j, k, mu, u, 1, 1, 0)   /* (mu v_x)_y */
         + sx * sz * i144 * mixed(b, i, j, k, mu, u, 2, 2, 0);  /* (mu w_x)_z */
   r[1] += sx * sy * i144 * mixed(b, i, j, k, mu, u, 0, 0, 1)   /* (mu u_y)_x */
         + sx * sy * i144 * mixed(b, i, j, k, la, u, 0, 1, 0)   /* (la u_x)_y */
         + sy * sz * i144 * mixed(b, i, j, k, la, u, 2, 1, 2)   /* (la w_z)_y */
         + sy * sz * i144 * mixed(b, i, j, k, mu, u, 2, 2, 1);  /* (mu w_y)_z */
   r[2] += sx * sz * i144 * mixed(b, i, j, k, mu, u, 0, 0, 2)   /* (mu u_z)_x */
         + sy * sz * i144 * mixed(b, i, j, k, mu, u, 1, 1, 2)   /* (mu v_z)_y */
         + sx * sz * i144 * mixed(b, i, j, k, la, u, 0, 2, 0)   /* (la u_x)_z */
         + sy * sz * i144 * mixed(b, i, j, k, la, u, 1, 2, 1);  /* (la v_y)_z */
}

/* SBP boundary row.  kr = row index 1..6 counted from the boundary,
 * kk(m) maps a boundary-relative index m=0..8 to the grid k, sgn = +1 at
 * the top (k = 1..6, rhs4sg_rev.C:349-601), -1 at the bottom
 * (k = nk-5..nk, rhs4sg_rev.C:602-863). */
static void rhs_sbp_point(const box_t *b, int i, int j, int k, int kr, int kbase, int sgn,
                          const double *acof, const double *bope, const double *ghcof,
                          const double *u, const double *mu, const double *la,
                          const double *strx, const double *stry, double r[3])
{
#define KK(m) (kbase + sgn * ((m)-1))
   const double i12 = 1.0 / 12, i144 = 1.0 / 144;
   const double sx = strx[i - b->ib], sy = stry[j - b->jb];
   static const int off[4] = {-2, -1, 1, 2};
   double zz[3] = {0, 0, 0};
   second_derivs(b, i, j, k, u, mu, la, strx, stry, NULL, 0, r);

   /* (mu u_z)_z, (mu v_z)_z, ((2mu+la) w_z)_z through the 8x8 acof blocks */
   for (int q = 1; q <= 8; q++) {
      double cmu = 0, c2ml = 0;
      for (int m = 1; m <= 8; m++) {
         double a = ACOF(acof, kr, q, m);
         ptrdiff_t pm = P(b, i, j, KK(m));
         cmu += a * mu[pm];
         c2ml += a * (2 * mu[pm] + la[pm]);
      }
      zz[0] += cmu * u[V(b, 0, i, j, KK(q))];
      zz[1] += cmu * u[V(b, 1, i, j, KK(q))];
      zz[2] += c2ml * u[V(b, 2, i, j, KK(q))];
   }
   {
      ptrdiff_t p1 = P(b, i, j, KK(1));
      double gh = ghcof[kr - 1];
      r[0] += zz[0] + gh * mu[p1] * u[V(b, 0, i, j, KK(0))];
      r[1] += zz[1] + gh * mu[p1] * u[V(b, 1, i, j, KK(0))];
      r[2] += zz[2] + gh * (2 * mu[p1] + la[p1]) * u[V(b, 2, i, j, KK(0))];
   }
   /* centred in-plane cross terms */
   r[0] += sx * sy * i144 * (mixed(b, i, j, k, la, u, 1, 0, 1) + mixed(b, i, j, k, mu, u, 1, 1, 0));
   r[1] += sx * sy * i144 * (mixed(b, i, j, k, mu, u, 0, 0, 1) + mixed(b, i, j, k, la, u, 0, 1, 0));

   /* one-sided z derivative via bope (negated at the bottom) */
   {
      double w3x[4], w3y[4], w1x[4], w2y[4];
      double t3x = 0, t3y = 0, t1x = 0, t2y = 0;
      for (int s = 0; s < 4; s++) {
         double a3x = 0, a3y = 0, a1x = 0, a2y = 0;
         for (int q = 1; q <= 8; q++) {
            double bq = sgn * BOPE(bope, kr, q);
            a3x += bq * u[V(b, 2, i + off[s], j, KK(q))];
            a3y += bq * u[V(b, 2, i, j + off[s], KK(q))];
            a1x += bq * u[V(b, 0, i + off[s], j, KK(q))];
            a2y += bq * u[V(b, 1, i, j + off[s], KK(q))];
         }
         w3x[s] = la[P(b, i + off[s], j, k)] * a3x;   /* la w_z along x */
         w3y[s] = la[P(b, i, j + off[s], k)] * a3y;   /* la w_z along y */
         w1x[s] = mu[P(b, i + off[s], j, k)] * a1x;   /* mu u_z along x */
         w2y[s] = mu[P(b, i, j + off[s], k)] * a2y;   /* mu v_z along y */
      }
      for (int q = 1; q <= 8; q++) {
         double bq = sgn * BOPE(bope, kr, q);
         int kq = KK(q);
         ptrdiff_t pq = P(b, i, j, kq);
         double dx3 = d4(u[V(b, 2, i - 2, j, kq)], u[V(b, 2, i - 1, j, kq)], u[V(b, 2, i + 1, j, kq)],
                         u[V(b, 2, i + 2, j, kq)]);
         double dy3 = d4(u[V(b, 2, i, j - 2, kq)], u[V(b, 2, i, j - 1, kq)], u[V(b, 2, i, j + 1, kq)],
                         u[V(b, 2, i, j + 2, kq)]);
         double dx1 = d4(u[V(b, 0, i - 2, j, kq)], u[V(b, 0, i - 1, j, kq)], u[V(b, 0, i + 1, j, kq)],
                         u[V(b, 0, i + 2, j, kq)]);
         double dy2 = d4(u[V(b, 1, i, j - 2, kq)], u[V(b, 1, i, j - 1, kq)], u[V(b, 1, i, j + 1, kq)],
                         u[V(b, 1, i, j + 2, kq)]);
         t3x += bq * (mu[pq] * i12 * dx3);   /* (mu w_x)_z */
         t3y += bq * (mu[pq] * i12 * dy3);   /* (mu w_y)_z */
         t1x += bq * (la[pq] * i12 * dx1);   /* (la u_x)_z */
         t2y += bq * (la[pq] * i12 * dy2);   /* (la v_y)_z */
      }
      r[0] += sx * (i12 * d4(w3x[0], w3x[1], w3x[2], w3x[3]));
      r[0] += sx * t3x;
      r[1] += sy * (i12 * d4(w3y[0], w3y[1], w3y[2], w3y[3]));
      r[1] += sy * t3y;
      r[2] += sx * (i12 * d4(w1x[0], w1x[1], w1x[2], w1x[3]));
      r[2] += sy * (i12 * d4(w2y[0], w2y[1], w2y[2], w2y[3]));
      r[2] += sx * t1x;
      r[2] += sy * t2y;
   }
#undef KK
}

/* lu = L(u): rhs4sg.C:38-849 (corder=0) and rhs4sg_rev.C:43-864 (corder=1).
 * Writes lu only at i,j in [first+2,last-2] and k in [1 or kfirst+2, ...]. */
void orc_rhs4sg(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                const int *onesided, const double *acof, const double *bope, const double *ghcof,
                double *lu, const double *u, const double *mu, const double *la, double h,
                const double *strx, const double *stry, const double *strz, int corder)
{
   box_t b = mkbox(ifirst, ilast, jfirst, jlast, kfirst, klast, corder);
   const double cof = 1.0 / (h * h);
   int k1 = onesided[4] == 1 ? 7 : kfirst + 2;
   int k2 = onesided[5] == 1 ? nk - 6 : klast - 2;
#pragma omp parallel for schedule(static)
   for (int k = kfirst; k <= klast; k++) {
      int top = onesided[4] == 1 && k >= 1 && k <= 6;
      int bot = onesided[5] == 1 && k >= nk - 5 && k <= nk;
      if (!(top || bot || (k >= k1 && k <= k2))) continue;
      for (int j = jfirst + 2; j <= jlast - 2; j++)
         for (int i = ifirst + 2; i <= ilast - 2; i++) {
            double r[3];
            if (top)
               rhs_sbp_point(&b, i, j, k, k, 1, +1, acof, bope, ghcof, u, mu, la, strx, stry, r);
            else if (bot)
               rhs_sbp_point(&b, i, j, k, nk - k + 1, nk, -1, acof, bope, ghcof, u, mu, la, strx,
                             stry, r);
            else
               rhs_interior_point(&b, i, j, k, u, mu, la, strx, stry, strz, r);
            for (int c = 0; c < 3; c++) lu[V(&b, c, i, j, k)] = cof * r[c];
         }
   }
}

/* ------------------------------------------------------------------ */
/* Pointwise time update, ew-cfromfort.C:40-141                        */
/* ------------------------------------------------------------------ */
/* up = 2u - um + dt2/rho (lu + f) over the whole box (ew-cfromfort.C:70-113) */
void orc_predfort(int ib, int ie, int jb, int je, int kb, int ke, double *up, const double *u,
                  const double *um, const double *lu, const double *fo, const double *rho,
                  double dt2, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
#pragma omp parallel for schedule(static)
   for (ptrdiff_t p = 0; p < b.npts; p++) {
      double s = dt2 / rho[p];
      for (int c = 0; c < 3; c++) {
         ptrdiff_t q = corder ? p + b.npts * c : 3 * p + c;
         up[q] = 2 * u[q] - um[q] + s * (lu[q] + fo[q]);
      }
   }
}

/* up += dt4/(12 rho) (lu + f)  (ew-cfromfort.C:40-67) */
void orc_corrfort(int ib, int ie, int jb, int je, int kb, int ke, double *up, const double *lu,
                  const double *fo, const double *rho, double dt4, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
   const double c12 = dt4 / 12;
#pragma omp parallel for schedule(static)
   for (ptrdiff_t p = 0; p < b.npts; p++) {
      double s = c12 / rho[p];
      for (int c = 0; c < 3; c++) {
         ptrdiff_t q = corder ? p + b.npts * c : 3 * p + c;
         up[q] += s * (lu[q] + fo[q]);
      }
   }
}

/* uacc = (up - 2u + um)/dt^2, layout independent (ew-cfromfort.C:116-141) */
void orc_dpdmtfort(int ib, int ie, int jb, int je, int kb, int ke, const double *up,
                   const double *u, const double *um, double *uacc, double dt2i)
{
   ptrdiff_t n = 3 * (ptrdiff_t)(ie - ib + 1) * (je - jb + 1) * (ke - kb + 1);
#pragma omp parallel for schedule(static)
   for (ptrdiff_t q = 0; q < n; q++) uacc[q] = dt2i * (up[q] - 2 * u[q] + um[q]);
}

/* ------------------------------------------------------------------ */
/* 4th-order supergrid damping, ew-cfromfort.C:748-849 / :952-1056     */
/* ------------------------------------------------------------------ */
void orc_addsgd4(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, double *up,
                 const double *u, const double *um, const double *rho, const double *dcx,
                 const double *dcy, const double *dcz, const double *strx, const double *stry,
                 const double *strz, const double *cox, const double *coy, const double *coz,
                 double beta, int corder)
{
   if (beta == 0) return;
   box_t b = mkbox(ifirst, ilast, jfirst, jlast, kfirst, klast, corder);
#pragma omp parallel for schedule(static)
   for (int k = kfirst + 2; k <= klast - 2; k++)
      for (int j = jfirst + 2; j <= jlast - 2; j++)
         for (int i = ifirst + 2; i <= ilast - 2; i++) {
            const int ii = i - ifirst, jj = j - jfirst, kk = k - kfirst;
            const double birho = beta / rho[P(&b, i, j, k)];
            for (int c = 0; c < 3; c++) {
               /* 1-D operator sum_s w_s rho dc (D+D- (u-um)) at s-1,s,s+1 with
                * weights (1,-2,1); d(t) = u - um at offset t along the axis */
               double acc[3];
               for (int dir = 0; dir < 3; dir++) {
                  const int di = dir == 0, dj = dir == 1, dk = dir == 2;
                  const double *dc = dir == 0 ? dcx : dir == 1 ? dcy : dcz;
                  const int base = dir == 0 ? ii : dir == 1 ? jj : kk;
                  double d[5];
                  for (int t = 0; t < 5; t++) {
                     ptrdiff_t q = V(&b, c, i + di * (t - 2), j + dj * (t - 2), k + dk * (t - 2));
                     d[t] = u[q] - um[q];
                  }
                  double rp = rho[P(&b, i + di, j + dj, k + dk)] * dc[base + 1];
                  double r0 = rho[P(&b, i, j, k)] * dc[base];
                  double rm = rho[P(&b, i - di, j - dj, k - dk)] * dc[base - 1];
                  acc[dir] = rp * (d[4] - 2 * d[3] + d[2]) - 2 * r0 * (d[3] - 2 * d[2] + d[1]) +
                             rm * (d[2] - 2 * d[1] + d[0]);
               }
               up[V(&b, c, i, j, k)] -=
                   birho * (strx[ii] * coy[jj] * coz[kk] * acc[0] + stry[jj] * cox[ii] * coz[kk] * acc[1] +
                            strz[kk] * cox[ii] * coy[jj] * acc[2]);
            }
         }
}

/* ------------------------------------------------------------------ */
/* Ghost-point boundary conditions, ew-cfromfort.C:205-747             */
/* ------------------------------------------------------------------ */
/* bccnd[s]: 0 free surface (sides 4,5 only), 1 Dirichlet, 2 supergrid,
 * 3 periodic; others untouched.  wind[6*s..6*s+5] = ghost window of side s
 * (EW.C:3347-3443).  bforce[s] holds 3 values per window point; for the
 * free surface it holds the traction on the boundary plane, indexed by
 * (i - ib) + ni (j - jb). */
int orc_bcfortsg(int ib, int ie, int jb, int je, int kb, int ke, const int *wind, int nx, int ny,
                 int nz, double *u, double h, const int *bccnd, const double *sbop,
                 const double *mu, const double *la, const double *const *bforce,
                 const double *strx, const double *stry, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
   const double d4a = 2.0 / 3.0, d4b = -1.0 / 12.0;
   for (int s = 0; s < 6; s++) {
      const int *w = wind + 6 * s;
      if (bccnd[s] == 1 || bccnd[s] == 2) {
         const double *bf = bforce[s];
         const ptrdiff_t wi = w[1] - w[0] + 1, wj = w[3] - w[2] + 1;
#pragma omp parallel for schedule(static)
         for (int k = w[4]; k <= w[5]; k++)
            for (int j = w[2]; j <= w[3]; j++)
               for (int i = w[0]; i <= w[1]; i++) {
                  ptrdiff_t q = (i - w[0]) + wi * (j - w[2]) + wi * wj * (k - w[4]);
                  for (int c = 0; c < 3; c++) u[V(&b, c, i, j, k)] = bf[3 * q + c];
               }
      } else if (bccnd[s] == 3) {
         /* periodic: copy from the point one period inside */
         const int pi = s == 0 ? nx : s == 1 ? -nx : 0;
         const int pj = s == 2 ? ny : s == 3 ? -ny : 0;
         const int pk = s == 4 ? nz : s == 5 ? -nz : 0;
#pragma omp parallel for schedule(static)
         for (int k = w[4]; k <= w[5]; k++)
            for (int j = w[2]; j <= w[3]; j++)
               for (int i = w[0]; i <= w[1]; i++)
                  for (int c = 0; c < 3; c++)
                     u[V(&b, c, i, j, k)] = u[V(&b, c, i + pi, j + pj, k + pk)];
      } else if (bccnd[s] == 0) {
         if (s != 4 && s != 5) return -1;
         const double *bf = bforce[s];
         const int k = s == 4 ? 1 : nz, kl = s == 4 ? 1 : -1;
#pragma omp parallel for schedule(static)
         for (int j = jb + 2; j <= je - 2; j++)
            for (int i = ib + 2; i <= ie - 2; i++) {
               const ptrdiff_t q = (i - ib) + b.ni * (j - jb);
               const ptrdiff_t p = P(&b, i, j, k);
               const double sx = strx[i - ib], sy = stry[j - jb];
#define UU(c, di, dj, dk) u[V(&b, c, i + (di), j + (dj), k + (dk))]
               double wx = sx * (d4a * (UU(2, 1, 0, 0) - UU(2, -1, 0, 0)) + d4b * (UU(2, 2, 0, 0) - UU(2, -2, 0, 0)));
               double ux = sx * (d4a * (UU(0, 1, 0, 0) - UU(0, -1, 0, 0)) + d4b * (UU(0, 2, 0, 0) - UU(0, -2, 0, 0)));
               double wy = sy * (d4a * (UU(2, 0, 1, 0) - UU(2, 0, -1, 0)) + d4b * (UU(2, 0, 2, 0) - UU(2, 0, -2, 0)));
               double vy = sy * (d4a * (UU(1, 0, 1, 0) - UU(1, 0, -1, 0)) + d4b * (UU(1, 0, 2, 0) - UU(1, 0, -2, 0)));
               double dz[3] = {0, 0, 0};
               for (int m = 1; m <= 4; m++)
                  for (int c = 0; c < 3; c++) dz[c] += sbop[m] * UU(c, 0, 0, kl * (m - 1));
               /* one-sided normal derivative sbop[0] u(ghost) + dz = -(tangential) + h*traction */
               UU(0, 0, 0, -kl) = (-dz[0] - kl * wx + kl * h * bf[3 * q] / mu[p]) / sbop[0];
               UU(1, 0, 0, -kl) = (-dz[1] - kl * wy + kl * h * bf[3 * q + 1] / mu[p]) / sbop[0];
               UU(2, 0, 0, -kl) =
                   (-dz[2] + (-kl * la[p] * (ux + vy) + kl * h * bf[3 * q + 2]) / (2 * mu[p] + la[p])) /
                   sbop[0];
#undef UU
            }
      }
   }
   return 0;
}

/* ------------------------------------------------------------------ */
/* Source time functions, time_functions.C                             */
/* ------------------------------------------------------------------ */
/* tfkind: 0 Ricker (time_functions.C:175-182, _tt :202-209),
 *         1 Gaussian (:284-290, _tt :311-317),
 *         2 C6SmoothBump (:1637-1648, _tt :1676-1687).
 * par0 is the exponent cut-off (-700, Source.C:2183-2187). */
double orc_timefunc(int tfkind, int deriv2, double freq, double t, double par0)
{
   if (tfkind == 0) {
      double x = M_PI * freq * t, f = x * x;
      if (!(-f > par0)) return 0;
      if (!deriv2) return (2 * f - 1) * exp(-f);
      return M_PI * M_PI * freq * freq * (6 - 24 * f + 8 * f * f) * exp(-f);
   } else if (tfkind == 1) {
      double tf = t * freq, f = tf * tf / 2;
      if (!(-f > par0)) return 0;
      if (!deriv2) return freq / sqrt(2 * M_PI) * exp(-f);
      return freq / sqrt(2 * M_PI) * freq * freq * (2 * f - 1) * exp(-f);
   } else if (tfkind == 2) {
      double s = t * freq;
      if (s < 0 || s > 1) return 0;
      double p = s * (1 - s);
      if (!deriv2) return 51480 * pow(p, 7);
      double q = 1 - 2 * s;
      return 51480 * freq * freq * 7 * (6 * q * q * pow(p, 5) - 2 * pow(p, 6));
   }
   return 0;
}

/* ------------------------------------------------------------------ */
/* Point-source forcing, EW.C:3085-3123 + GridPointSource.C:330-586   */
/* ------------------------------------------------------------------ */
/* F := 0 over the whole box, then for each grid-point source s (sorted by
 * grid index), F(:, i_s, j_s, k_s) += force_s * g(t - t0_s). */
void orc_force(int ib, int ie, int jb, int je, int kb, int ke, double *F, int nsrc,
               const int *ijk, const double *force, const double *freq, const double *t0,
               const int *tfkind, double t, int deriv2, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
   memset(F, 0, sizeof(double) * 3 * b.npts);
   for (int s = 0; s < nsrc; s++) {
      double g = orc_timefunc(tfkind[s], deriv2, freq[s], t - t0[s], -700.0);
      for (int c = 0; c < 3; c++)
         F[V(&b, c, ijk[3 * s], ijk[3 * s + 1], ijk[3 * s + 2])] += force[3 * s + c] * g;
   }
}

/* ------------------------------------------------------------------ */
/* Error norms, ew-cfromfort.C:144-202                                 */
/* ------------------------------------------------------------------ */
/* Linf and h^3-weighted L2 of u - uex over [imin,imax]x... outside a ball of
 * radius `radius` around (x0,y0,z0); xli = max of uex. */
void orc_solerr3(int ib, int ie, int jb, int je, int kb, int ke, double h, const double *uex,
                 const double *u, double *li, double *l2, double *xli, double zmin, double x0,
                 double y0, double z0, double radius, int imin, int imax, int jmin, int jmax,
                 int kmin, int kmax, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
   double r2 = radius * radius;
   if (radius < 0) r2 = -r2;
   const double h3 = h * h * h;
   double vli = 0, vl2 = 0, vxli = 0;
   for (int c = 0; c < 3; c++)
      for (int k = kmin; k <= kmax; k++)
         for (int j = jmin; j <= jmax; j++)
            for (int i = imin; i <= imax; i++) {
               double dx = (i - 1) * h - x0, dy = (j - 1) * h - y0, dz = (k - 1) * h + zmin - z0;
               if (dx * dx + dy * dy + dz * dz <= r2) continue;
               ptrdiff_t q = V(&b, c, i, j, k);
               double e = u[q] - uex[q];
               if (fabs(e) > vli) vli = fabs(e);
               if (uex[q] > vxli) vxli = uex[q];
               vl2 += h3 * e * e;
            }
   *li = vli;
   *l2 = vl2;
   *xli = vxli;
}

/* ------------------------------------------------------------------ */
/* Analytic whole-space solution of a moment-tensor point source,      */
/* EW::get_exact_point_source (EW.C:3899-4523).                        */
/* ------------------------------------------------------------------ */
/* Time functions in the scaled time of EW.C:3929-3934:
 *   kind 2 (C6SmoothBump): g(s) = 51480 (s(1-s))^7 on (0,1);
 *   kind 1 (Gaussian):     g(s) = exp(-s^2/(2 f^2)) / (f sqrt(2 pi)).
 * The near-field integral int_{R/alpha}^{R/beta} tau g(time - tau) dtau is
 * evaluated here with 8-point Gauss-Legendre quadrature on the part of the
 * interval inside the support (exact for the degree-15 polynomial integrand
 * of the bump; for the Gaussian the interval is split into 64 panels).  This
 * is a different method from the closed forms of the HIP library. */
static double ex_g(int kind, double s, double f)
{
   if (kind == 2) {
      if (!(s > 0 && s < 1)) return 0;
      return 51480.0 * pow(s * (1 - s), 7);
   }
   return exp(-s * s / (2 * f * f)) / (f * sqrt(2 * M_PI));
}
static double ex_gp(int kind, double s, double f)
{
   if (kind == 2) {
      if (!(s > 0 && s < 1)) return 0;
      return 51480.0 * 7 * (1 - 2 * s) * pow(s * (1 - s), 6);
   }
   return -s / (f * f) * ex_g(kind, s, f);
}
static double ex_tau_integral(int kind, double time, double R, double alpha, double beta, double f)
{
   static const double xg[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                                0.7966664774136267,  0.9602898564975363};
   static const double wg[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                                0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                                0.2223810344533745, 0.1012285362903763};
   double lo = R / alpha, hi = R / beta;
   int npanel = 64;
   if (kind == 2) {
      if (time - 1 > lo) lo = time - 1;
      if (time < hi) hi = time;
      if (!(lo < hi)) return 0;
      npanel = 1;
   }
   double sum = 0, w = (hi - lo) / npanel;
   for (int p = 0; p < npanel; p++) {
      double a = lo + p * w, c = 0.5 * w, m = a + c;
      for (int q = 0; q < 8; q++) {
         double tau = m + c * xg[q];
         sum += c * wg[q] * tau * ex_g(kind, time - tau, f);
      }
   }
   return sum;
}

/* u_i = -1/(4 pi rho) sum_kj M_kj d/dx_j G_ik(x - x0) with the Stokes
 * solution G (Aki & Richards eq. 4.23).  Written over the whole box. */
void orc_exact_moment(int ib, int ie, int jb, int je, int kb, int ke, double h, double zmin,
                      double x0, double y0, double z0, double rho, double alpha, double beta,
                      double time, double f, const double *m6, int kind, double *out, int corder)
{
   box_t b = mkbox(ib, ie, jb, je, kb, ke, corder);
   const double Mt[3][3] = {{m6[0], m6[1], m6[2]}, {m6[1], m6[3], m6[4]}, {m6[2], m6[4], m6[5]}};
   const double eps = 1e-3 * h;
#pragma omp parallel for schedule(dynamic, 1)
   for (int k = kb; k <= ke; k++)
      for (int j = jb; j <= je; j++)
         for (int i = ib; i <= ie; i++) {
            double X[3] = {(i - 1) * h - x0, (j - 1) * h - y0, (k - 1) * h + zmin - z0};
            double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
            double res[3] = {0, 0, 0};
            if (R >= eps) {
               double a2 = alpha * alpha, b2 = beta * beta;
               double Aa = ex_g(kind, time - R / alpha, f) / a2;
               double Bb = ex_g(kind, time - R / beta, f) / b2;
               double Cc = ex_tau_integral(kind, time, R, alpha, beta, f);
               double Dd = ex_gp(kind, time - R / alpha, f) / (a2 * alpha) / R;
               double Ee = ex_gp(kind, time - R / beta, f) / (b2 * beta) / R;
               /* radial derivatives of the three parts of G:
                *   dC/dx_j = X_j (Bb - Aa), d(Aa)/dx_j = -X_j Dd, d(Bb)/dx_j = -X_j Ee */
               double R2 = R * R, R3 = R2 * R, R5 = R3 * R2, R7 = R5 * R2;
               for (int ii = 0; ii < 3; ii++) {
                  double acc = 0;
                  for (int kk = 0; kk < 3; kk++)
                     for (int jj = 0; jj < 3; jj++) {
                        double dij = ii == jj, dkj = kk == jj, dik = ii == kk;
                        double xx = X[ii] * X[kk], dxx = dij * X[kk] + dkj * X[ii], xj = X[jj];
                        /* G1 = (3 xx/R^5 - dik/R^3) C */
                        double dG1 = (3 * dxx / R5 - 15 * xx * xj / R7 + 3 * dik * xj / R5) * Cc +
                                     (3 * xx / R5 - dik / R3) * xj * (Bb - Aa);
                        /* G2 = xx/R^3 Aa */
                        double dG2 = (dxx / R3 - 3 * xx * xj / R5) * Aa - xx / R3 * xj * Dd;
                        /* G3 = -(xx/R^3 - dik/R) Bb */
                        double dG3 = -(dxx / R3 - 3 * xx * xj / R5 + dik * xj / R3) * Bb +
                                     (xx / R3 - dik / R) * xj * Ee;
                        acc += Mt[kk][jj] * (dG1 + dG2 + dG3);
                     }
                  res[ii] = -acc / (4 * M_PI * rho);
               }
            }
            for (int c = 0; c < 3; c++) out[V(&b, c, i, j, k)] = res[c];
         }
}

/* Parallel first-touch fill, as EW::timesteploop does with set_value(0.0)
 * on every array before stepping (EW.C:2390-2398). */
void orc_fill(ptrdiff_t n, double *a, double v)
{
#pragma omp parallel for schedule(static)
   for (ptrdiff_t q = 0; q < n; q++) a[q] = v;
}
