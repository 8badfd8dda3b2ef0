/*
 * sw4_oracle_curv.c -- CPU restatement (C99 + OpenMP, FP64) of the sw4lite
 * curvilinear-grid kernels: the metric of a topography-following grid, the
 * curvilinear divergence of stress, its supergrid damping and the traction-
 * free ghost-point closure on the curved top surface.
 *
 * TEST INFRASTRUCTURE ONLY.  This file belongs to the parity checker for the
 * HIP library in sw4lite_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Conventions are those of sw4_oracle.c (index boxes with two ghost points,
 * corder 0 = component-interleaved, 1 = component-blocked).  The metric is a
 * 4-component array met(1..4) = (sqrt(z_r), -z_p/sqrt(z_r), -z_q/sqrt(z_r),
 * h/sqrt(z_r)) laid out like a 4-component Sarray (curvilinear-c.C:53-172),
 * here indexed with c = 0..3; jac = h^2 z_r.
 *
 * Pinning (tests/test_oracle_curv.py): orc_metric and orc_rhs4sgcurv are
 * compared with the reference's own metric (tests/testil/metricC.C) and
 * rhs4sgcurv/rhs4sgcurv_rev sources compiled by oracle/Makefile; on a flat
 * grid the curvilinear operator, damping and closure reduce to the Cartesian
 * ones (pinned against rhs4sg.C / addsgd4fort.C); the closure is pinned by
 * the traction it leaves (zero up to the given forcing) on curved grids.
 */
#include <math.h>
#include <stddef.h>

typedef struct {
   int ib, ie, jb, je, kb, ke;
   ptrdiff_t ni, nij, npts;
   int corder;
} cbox_t;

static cbox_t cmkbox(int ib, int ie, int jb, int je, int kb, int ke, int corder)
{
   cbox_t b;
   b.ib = ib; b.ie = ie; b.jb = jb; b.je = je; b.kb = kb; b.ke = ke;
   b.ni = ie - ib + 1;
   b.nij = b.ni * (je - jb + 1);
   b.npts = b.nij * (ke - kb + 1);
   b.corder = corder;
   return b;
}

static inline ptrdiff_t SP(const cbox_t *b, int i, int j, int k)
{
   return (i - b->ib) + b->ni * (j - b->jb) + b->nij * (k - b->kb);
}
/* component c of an nc-component array (nc = 3 displacement, 4 metric) */
static inline ptrdiff_t CP(const cbox_t *b, int nc, int c, int i, int j, int k)
{
   const ptrdiff_t p = SP(b, i, j, k);
   return b->corder ? p + b->npts * c : nc * p + c;
}

#define ACOF(a, k, q, m) (a)[((k)-1) + 6 * ((q)-1) + 48 * ((m)-1)]
#define BOPE(b, k, q) (b)[((k)-1) + 6 * ((q)-1)]

/* ------------------------------------------------------------------ */
/* Metric, curvilinear-c.C:53-172 (metric) / :175-295 (metric_rev)     */
/* ------------------------------------------------------------------ */

/* d/ds of f along one grid line at position s of n points (0-based), f(t)
 * returned by get(t): centred 4th order inside, the 6-point one-sided
 * closures on the first and last two points (curvilinear-c.C:78-146). */
static double line_derivative(const double *f, ptrdiff_t stride, int s, int n)
{
#define F(t) f[(ptrdiff_t)(t) * stride]
   const double c1 = 2.0 / 3, c2 = -1.0 / 12;
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
#undef F
}

/* met, jac from the grid depth z over the whole box; h is the horizontal
 * spacing.  Returns -1 (and stops) where z_r <= 0, else 0. */
int orc_metric(int ib, int ie, int jb, int je, int kb, int ke, double h, const double *z, double *met,
               double *jac, int corder)
{
   cbox_t b = cmkbox(ib, ie, jb, je, kb, ke, corder);
   const int ni = ie - ib + 1, nj = je - jb + 1, nk = ke - kb + 1;
   int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
   for (int k = kb; k <= ke; k++)
      for (int j = jb; j <= je; j++)
         for (int i = ib; i <= ie; i++) {
            const ptrdiff_t p = SP(&b, i, j, k);
            const double zr = line_derivative(z + (p - b.nij * (k - kb)), b.nij, k - kb, nk);
            const double zq = line_derivative(z + (p - b.ni * (j - jb)), b.ni, j - jb, nj);
            const double zp = line_derivative(z + (p - (i - ib)), 1, i - ib, ni);
            if (!(zr > 0)) {
               bad = 1;
               continue;
            }
            const double sq = sqrt(zr);
            jac[p] = h * h * zr;
            met[CP(&b, 4, 0, i, j, k)] = sq;
            met[CP(&b, 4, 1, i, j, k)] = -zp / sq;
            met[CP(&b, 4, 2, i, j, k)] = -zq / sq;
            met[CP(&b, 4, 3, i, j, k)] = h / sq;
         }
   return bad ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Curvilinear divergence of stress, rhs4sgcurv.C:34-1406 and          */
/* rhs4sgcurv_rev.C (same arithmetic, component-blocked arrays)        */
/* ------------------------------------------------------------------ */
/*
 * With M1..M4 the metric, K = 2mu+lambda, sx = strx(i), sy = stry(j) and
 * D_p, D_q the centred 4th-order first differences along i and j,
 * L(u)_a = sx sy / jac * r_a where r_a is the sum of
 *   pp:  (1/6) (A u_a,p)_p / sy,  A = (a==0 ? K : mu) M1^2 strx
 *   qq:  (1/6) (B u_a,q)_q / sx,  B = (a==1 ? K : mu) M1^2 stry
 *   rr:  sum_b S_ab (R_ab u_b,r)_r with the 3x3 coupling R of mu, lambda,
 *        M2 sx, M3 sy, M4 and the scalings S (rhs4sgcurv.C:226-283);
 *        4th-order variable-coefficient second difference inside, the SBP
 *        acof/ghcof closure on the top six rows;
 *   pq, qp: centred products in the (i,j) plane (rhs4sgcurv.C:285-365);
 *   pr, qr: k-derivative (bope rows or centred) of T_a, T built from D_p u
 *           and D_q u on each plane (rhs4sgcurv.C:540-578);
 *   rp, rq: D_p resp. D_q of P_a resp. Q_a, built from the k-derivatives of
 *           u on the neighbouring columns (rhs4sgcurv.C:367-538).
 */
typedef struct {
   cbox_t b;
   const double *u, *mu, *la, *met, *strx, *stry;
} curv_t;

static inline double cu(const curv_t *s, int c, int i, int j, int k) { return s->u[CP(&s->b, 3, c, i, j, k)]; }
static inline double cm(const curv_t *s, int c, int i, int j, int k) { return s->met[CP(&s->b, 4, c, i, j, k)]; }
static inline double cmu(const curv_t *s, int i, int j, int k) { return s->mu[SP(&s->b, i, j, k)]; }
static inline double cla(const curv_t *s, int i, int j, int k) { return s->la[SP(&s->b, i, j, k)]; }

static const double C1 = 2.0 / 3, C2 = -1.0 / 12;
static const int OFF[4] = {-2, -1, 1, 2};
static const double WC[4] = {-(-1.0 / 12), -(2.0 / 3), 2.0 / 3, -1.0 / 12}; /* weights of OFF */

/* centred first differences of u_c along i and j at (i,j,k) */
static double dp(const curv_t *s, int c, int i, int j, int k)
{
   return C2 * (cu(s, c, i + 2, j, k) - cu(s, c, i - 2, j, k)) + C1 * (cu(s, c, i + 1, j, k) - cu(s, c, i - 1, j, k));
}
static double dq(const curv_t *s, int c, int i, int j, int k)
{
   return C2 * (cu(s, c, i, j + 2, k) - cu(s, c, i, j - 2, k)) + C1 * (cu(s, c, i, j + 1, k) - cu(s, c, i, j - 1, k));
}

static void vc4(const double a[5], double m[4])
{
   m[0] = a[1] - 0.75 * (a[2] + a[0]);
   m[1] = a[0] + a[3] + 3 * (a[2] + a[1]);
   m[2] = a[1] + a[4] + 3 * (a[3] + a[2]);
   m[3] = a[3] - 0.75 * (a[2] + a[4]);
}
static double vs4(const double m[4], const double f[5])
{
   return m[0] * (f[0] - f[2]) + m[1] * (f[1] - f[2]) + m[2] * (f[3] - f[2]) + m[3] * (f[4] - f[2]);
}

/* rr coupling R_ab at (i,j,m) (symmetric; index ab = 00,11,22,01,02,12) */
static void rr_coef(const curv_t *s, int i, int j, int m, double sx, double sy, double R[6])
{
   const double mu = cmu(s, i, j, m), la = cla(s, i, j, m);
   const double m2 = cm(s, 1, i, j, m), m3 = cm(s, 2, i, j, m), m4 = cm(s, 3, i, j, m);
   const double x2 = m2 * sx * m2 * sx, y2 = m3 * sy * m3 * sy, z2 = m4 * m4;
   R[0] = (2 * mu + la) * x2 + mu * (y2 + z2);
   R[1] = (2 * mu + la) * y2 + mu * (x2 + z2);
   R[2] = (2 * mu + la) * z2 + mu * (x2 + y2);
   R[3] = (mu + la) * m2 * m3;
   R[4] = (mu + la) * m2 * m4;
   R[5] = (mu + la) * m3 * m4;
}

/* T_a of the pr/qr terms on plane m at column (i,j) */
static void t_terms(const curv_t *s, int i, int j, int m, double sx, double sy, double isx, double isy, double T[3])
{
   const double mu = cmu(s, i, j, m), la = cla(s, i, j, m), K = 2 * mu + la;
   const double m1 = cm(s, 0, i, j, m), m2 = cm(s, 1, i, j, m), m3 = cm(s, 2, i, j, m), m4 = cm(s, 3, i, j, m);
   const double p0 = dp(s, 0, i, j, m), p1 = dp(s, 1, i, j, m), p2 = dp(s, 2, i, j, m);
   const double q0 = dq(s, 0, i, j, m), q1 = dq(s, 1, i, j, m), q2 = dq(s, 2, i, j, m);
   T[0] = K * m2 * m1 * p0 * sx * isy + mu * m3 * m1 * p1 + mu * m4 * m1 * p2 * isy + mu * m3 * m1 * q0 * sy * isx +
          la * m2 * m1 * q1;
   T[1] = la * m3 * m1 * p0 + mu * m2 * m1 * p1 * sx * isy + mu * m2 * m1 * q0 + K * m3 * m1 * q1 * sy * isx +
          mu * m4 * m1 * q2 * isx;
   T[2] = la * m4 * m1 * p0 * isy + mu * m2 * m1 * p2 * sx * isy + mu * m3 * m1 * q2 * sy * isx +
          la * m4 * m1 * q1 * isx;
}

/* k-derivative weights of row k: bope (top six rows) or centred */
static int kweights(int k, int bnd, const double *bope, int kq[8], double wq[8])
{
   if (bnd) {
      for (int q = 1; q <= 8; q++) {
         kq[q - 1] = q;
         wq[q - 1] = BOPE(bope, k, q);
      }
      return 8;
   }
   for (int t = 0; t < 4; t++) {
      kq[t] = k + OFF[t];
      wq[t] = WC[t];
   }
   return 4;
}

static void curv_point(const curv_t *s, int i, int j, int k, int bnd, const double *acof, const double *bope,
                       const double *ghcof, double r[3])
{
   const cbox_t *b = &s->b;
   const double i6 = 1.0 / 6;
   const double sx = s->strx[i - b->ib], sy = s->stry[j - b->jb];
   const double isx = 1 / sx, isy = 1 / sy, isxy = isy * isx;
   double a[5], m[4], f[5];
   r[0] = r[1] = r[2] = 0;

   /* pp and qq */
   for (int c = 0; c < 3; c++) {
      for (int t = 0; t < 5; t++) {
         const int ii = i + t - 2;
         const double K = c == 0 ? 2 * cmu(s, ii, j, k) + cla(s, ii, j, k) : cmu(s, ii, j, k);
         a[t] = K * cm(s, 0, ii, j, k) * cm(s, 0, ii, j, k) * s->strx[ii - b->ib];
         f[t] = cu(s, c, ii, j, k);
      }
      vc4(a, m);
      r[c] += i6 * vs4(m, f) * isy;
      for (int t = 0; t < 5; t++) {
         const int jj = j + t - 2;
         const double K = c == 1 ? 2 * cmu(s, i, jj, k) + cla(s, i, jj, k) : cmu(s, i, jj, k);
         a[t] = K * cm(s, 0, i, jj, k) * cm(s, 0, i, jj, k) * s->stry[jj - b->jb];
         f[t] = cu(s, c, i, jj, k);
      }
      vc4(a, m);
      r[c] += i6 * vs4(m, f) * isx;
   }

   /* rr: pair index of (a,b) into R[] and the scaling of each pair */
   static const int PAIR[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
   const double S[3][3] = {{isxy, 1, isy}, {1, isxy, isx}, {isy, isx, isxy}};
   if (!bnd) {
      double R[5][6];
      for (int t = 0; t < 5; t++) rr_coef(s, i, j, k + t - 2, sx, sy, R[t]);
      for (int c = 0; c < 3; c++)
         for (int d = 0; d < 3; d++) {
            for (int t = 0; t < 5; t++) {
               a[t] = R[t][PAIR[c][d]];
               f[t] = cu(s, d, i, j, k + t - 2);
            }
            vc4(a, m);
            r[c] += i6 * vs4(m, f) * S[c][d];
         }
   } else {
      double R[9][6]; /* planes 1..8 */
      for (int q = 1; q <= 8; q++) rr_coef(s, i, j, q, sx, sy, R[q]);
      for (int q = 1; q <= 8; q++) {
         double Cq[6] = {0, 0, 0, 0, 0, 0};
         for (int mm = 1; mm <= 8; mm++)
            for (int e = 0; e < 6; e++) Cq[e] += ACOF(acof, k, q, mm) * R[mm][e];
         for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) r[c] += S[c][d] * Cq[PAIR[c][d]] * cu(s, d, i, j, q);
      }
      const double g = ghcof[k - 1];
      for (int c = 0; c < 3; c++)
         for (int d = 0; d < 3; d++) r[c] += S[c][d] * (g * R[1][PAIR[c][d]]) * cu(s, d, i, j, 0);
   }

   /* pq and qp (u and v equations) */
   for (int t = 0; t < 4; t++) {
      const int d = OFF[t];
      const double mq = cm(s, 0, i, j + d, k), mp = cm(s, 0, i + d, j, k);
      r[0] += WC[t] * (cmu(s, i, j + d, k) * mq * mq * dp(s, 1, i, j + d, k) +
                       cla(s, i + d, j, k) * mp * mp * dq(s, 1, i + d, j, k));
      r[1] += WC[t] * (cla(s, i, j + d, k) * mq * mq * dp(s, 0, i, j + d, k) +
                       cmu(s, i + d, j, k) * mp * mp * dq(s, 0, i + d, j, k));
   }

   /* pr and qr */
   int kq[8];
   double wq[8];
   const int nq = kweights(k, bnd, bope, kq, wq);
   for (int t = 0; t < nq; t++) {
      double T[3];
      t_terms(s, i, j, kq[t], sx, sy, isx, isy, T);
      for (int c = 0; c < 3; c++) r[c] += wq[t] * T[c];
   }

   /* rp and rq */
   for (int t = 0; t < 4; t++) {
      const int d = OFF[t];
      double dr[2][3] = {{0, 0, 0}, {0, 0, 0}}; /* column (i+d,j), column (i,j+d) */
      for (int e = 0; e < nq; e++)
         for (int c = 0; c < 3; c++) {
            dr[0][c] += wq[e] * cu(s, c, i + d, j, kq[e]);
            dr[1][c] += wq[e] * cu(s, c, i, j + d, kq[e]);
         }
      {
         const int ii = i + d;
         const double mu = cmu(s, ii, j, k), la = cla(s, ii, j, k), sxi = s->strx[ii - b->ib];
         const double m1 = cm(s, 0, ii, j, k), m2 = cm(s, 1, ii, j, k), m3 = cm(s, 2, ii, j, k),
                      m4 = cm(s, 3, ii, j, k);
         r[0] += WC[t] * ((2 * mu + la) * m2 * m1 * sxi * dr[0][0] + la * m3 * m1 * dr[0][1] * sy +
                          la * m4 * m1 * dr[0][2]) * isy;
         r[1] += WC[t] * (mu * m3 * m1 * dr[0][0] + mu * m2 * m1 * dr[0][1] * sxi * isy);
         r[2] += WC[t] * (mu * m4 * m1 * dr[0][0] + mu * m2 * m1 * dr[0][2] * sxi) * isy;
      }
      {
         const int jj = j + d;
         const double mu = cmu(s, i, jj, k), la = cla(s, i, jj, k), syj = s->stry[jj - b->jb];
         const double m1 = cm(s, 0, i, jj, k), m2 = cm(s, 1, i, jj, k), m3 = cm(s, 2, i, jj, k),
                      m4 = cm(s, 3, i, jj, k);
         r[0] += WC[t] * (mu * m3 * m1 * dr[1][0] * syj * isx + mu * m2 * m1 * dr[1][1]);
         r[1] += WC[t] * (la * m2 * m1 * dr[1][0] + (2 * mu + la) * m3 * m1 * dr[1][1] * syj * isx +
                          la * m4 * m1 * dr[1][2] * isx);
         r[2] += WC[t] * (mu * m3 * m1 * dr[1][2] * syj + mu * m4 * m1 * dr[1][1]) * isx;
      }
   }
}

/* lu = L(u) at i,j in [first+2,last-2], k in [1,6] (onesided[4] == 1, SBP
 * rows) and [kstart, klast-2], kstart = 7 or kfirst+2 (rhs4sgcurv.C:86-94,
 * 587-590).  No bottom closure and no z stretching: the curvilinear grid is
 * the top grid. */
void orc_rhs4sgcurv(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, const double *u,
                    const double *mu, const double *la, const double *met, const double *jac, double *lu,
                    const int *onesided, const double *acof, const double *bope, const double *ghcof,
                    const double *strx, const double *stry, int corder)
{
   curv_t s;
   s.b = cmkbox(ifirst, ilast, jfirst, jlast, kfirst, klast, corder);
   s.u = u; s.mu = mu; s.la = la; s.met = met; s.strx = strx; s.stry = stry;
   const int top = onesided[4] == 1;
   const int kstart = top ? 7 : kfirst + 2;
#pragma omp parallel for schedule(static)
   for (int k = kfirst; k <= klast - 2; k++) {
      const int bnd = top && k >= 1 && k <= 6;
      if (!bnd && k < kstart) continue;
      for (int j = jfirst + 2; j <= jlast - 2; j++)
         for (int i = ifirst + 2; i <= ilast - 2; i++) {
            double r[3];
            curv_point(&s, i, j, k, bnd, acof, bope, ghcof, r);
            const double ij = strx[i - ifirst] * stry[j - jfirst] / jac[SP(&s.b, i, j, k)];
            for (int c = 0; c < 3; c++) lu[CP(&s.b, 3, c, i, j, k)] = r[c] * ij;
         }
   }
}

/* ------------------------------------------------------------------ */
/* Curvilinear supergrid damping, ew-cfromfort.C:1160-1240 / :1323-    */
/* ------------------------------------------------------------------ */
/* up -= beta/(rho jac) [strx coy Dx(rho dc jac Dx^2 (u-um))
 *                       + stry cox Dy(rho dc jac Dy^2 (u-um))]: the x and y
 * parts of the Cartesian damping with rho -> rho jac; no z part. */
void orc_addsgd4c(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, double *up,
                  const double *u, const double *um, const double *rho, const double *dcx, const double *dcy,
                  const double *strx, const double *stry, const double *jac, const double *cox,
                  const double *coy, double beta, int corder)
{
   if (beta == 0) return;
   cbox_t b = cmkbox(ifirst, ilast, jfirst, jlast, kfirst, klast, corder);
#pragma omp parallel for schedule(static)
   for (int k = kfirst + 2; k <= klast - 2; k++)
      for (int j = jfirst + 2; j <= jlast - 2; j++)
         for (int i = ifirst + 2; i <= ilast - 2; i++) {
            const int ii = i - ifirst, jj = j - jfirst;
            const double w = beta / (rho[SP(&b, i, j, k)] * jac[SP(&b, i, j, k)]);
            for (int c = 0; c < 3; c++) {
               double acc[2];
               for (int dir = 0; dir < 2; dir++) {
                  const int di = dir == 0, dj = dir == 1;
                  const double *dc = dir == 0 ? dcx : dcy;
                  const int s0 = dir == 0 ? ii : jj;
                  double d[5], rj[3];
                  for (int t = 0; t < 5; t++) {
                     const ptrdiff_t q = CP(&b, 3, c, i + di * (t - 2), j + dj * (t - 2), k);
                     d[t] = u[q] - um[q];
                  }
                  for (int t = 0; t < 3; t++) {
                     const ptrdiff_t p = SP(&b, i + di * (t - 1), j + dj * (t - 1), k);
                     rj[t] = rho[p] * dc[s0 + t - 1] * jac[p];
                  }
                  acc[dir] = rj[2] * (d[4] - 2 * d[3] + d[2]) - 2 * rj[1] * (d[3] - 2 * d[2] + d[1]) +
                             rj[0] * (d[2] - 2 * d[1] + d[0]);
               }
               up[CP(&b, 3, c, i, j, k)] -= w * (strx[ii] * coy[jj] * acc[0] + stry[jj] * cox[ii] * acc[1]);
            }
         }
}

/* ------------------------------------------------------------------ */
/* Traction-free ghost points on the curved top (or bottom) plane,     */
/* curvilinear-c.C:465-618 (freesurfcurvisg) / :621-764 (_rev)         */
/* ------------------------------------------------------------------ */
/* side 5 (k = 1, ghost k = 0) or 6 (k = nz, ghost nz+1), numbered from 1
 * as in the reference; forcing holds the
 * boundary traction, 3 interleaved values per (i,j) of the plane.  The
 * tangential part g of the traction comes from centred differences in the
 * plane; the normal part is N (sbop . u along k) with
 * N = mu ac I + (mu+lambda) v v^T, v = (sqrt(sx/sy) M2, sqrt(sy/sx) M3,
 * M4/sqrt(sx sy)), ac = |v|^2, so the ghost value solves
 * N (s0 u_ghost + sum_w s_w u_w) = -(g - forcing) with
 * N^-1 = I/(mu ac) - (mu+lambda)/((2mu+lambda) mu ac^2) v v^T. */
void orc_freesurfcurvisg(int ib, int ie, int jb, int je, int kb, int ke, int nz, int side, double *u,
                         const double *mu, const double *la, const double *met, const double *sbop,
                         const double *forcing, const double *strx, const double *stry, int corder)
{
   curv_t s;
   s.b = cmkbox(ib, ie, jb, je, kb, ke, corder);
   s.u = u; s.mu = mu; s.la = la; s.met = met; s.strx = strx; s.stry = stry;
   const int k = side == 5 ? 1 : nz, kl = side == 5 ? 1 : -1;
   const ptrdiff_t ni = ie - ib + 1;
#pragma omp parallel for schedule(static)
   for (int j = jb + 2; j <= je - 2; j++)
      for (int i = ib + 2; i <= ie - 2; i++) {
         const double sx = strx[i - ib], sy = stry[j - jb], isx = 1 / sx, isy = 1 / sy;
         double T[3];
         t_terms(&s, i, j, k, sx, sy, isx, isy, T);
         const ptrdiff_t f = 3 * ((i - ib) + ni * (j - jb));
         double g[3];
         for (int c = 0; c < 3; c++) g[c] = T[c] - forcing[f + c];
         const double m2 = cm(&s, 1, i, j, k), m3 = cm(&s, 2, i, j, k), m4 = cm(&s, 3, i, j, k);
         const double mu0 = cmu(&s, i, j, k), la0 = cla(&s, i, j, k);
         const double ac = sx * isy * m2 * m2 + sy * isx * m3 * m3 + m4 * m4 * isy * isx;
         const double bc = 1 / (mu0 * ac);
         const double cc = (mu0 + la0) / (2 * mu0 + la0) * bc / ac;
         const double xy = sqrt(sx * isy);
         const double v[3] = {xy * m2, m3 / xy, isx * xy * m4};
         const double dc = cc * (v[0] * g[0] + v[1] * g[1] + v[2] * g[2]);
         for (int c = 0; c < 3; c++) {
            double acc = 0;
            for (int w = 1; w <= 4; w++) acc += sbop[w] * cu(&s, c, i, j, k + kl * (w - 1));
            u[CP(&s.b, 3, c, i, j, k - kl)] = -(acc + bc * g[c] - dc * v[c]) / sbop[0];
         }
      }
}
