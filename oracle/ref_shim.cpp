// ref_shim.cpp -- C-linkage forwarders to the reference kernels compiled from
// their own sources by oracle/Makefile into oracle/_ref/libsw4ref.so.
// TEST INFRASTRUCTURE ONLY (used by tests/ and tests/golden/make_golden.py).
// The prototypes restate the reference signatures:
//   rhs4sg           src/rhs4sg.C:38-41
//   rhs4sg_rev       src/rhs4sg_rev.C:43-48
//   addsgd4fort      tests/testil/addsgd4fort.C:4-11
//   addsgd4fort_rev  tests/testil/addsgd4fort_rev.C:4-11
//   rhs4sgcurv       src/rhs4sgcurv.C:34-39
//   rhs4sgcurv_rev   src/rhs4sgcurv_rev.C:34-40
//   metric           tests/testil/metricC.C:5-6
//   metric_rev       tests/testil/metricC.C:127-128
typedef double R;

void rhs4sg(int, int, int, int, int, int, int, int*, R*, R*, R*, R*, R*, R*, R*, R, R*, R*, R*);
void rhs4sg_rev(int, int, int, int, int, int, int, int*, R*, R*, R*, R*, R*, R*, R*, R, R*, R*, R*);
void addsgd4fort(int, int, int, int, int, int, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R);
void addsgd4fort_rev(int, int, int, int, int, int, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R*, R);
void rhs4sgcurv(int, int, int, int, int, int, R*, R*, R*, R*, R*, R*, int*, R*, R*, R*, R*, R*);
void rhs4sgcurv_rev(int, int, int, int, int, int, R*, R*, R*, R*, R*, R*, int*, R*, R*, R*, R*, R*);
int metric(int, int, int, int, int, int, R*, R*, R*, R*, R*);
int metric_rev(int, int, int, int, int, int, R*, R*, R*, R*, R*);

extern "C" {
void ref_rhs4sg(int corder, int ib, int ie, int jb, int je, int kb, int ke, int nk, int* onesided,
                R* acof, R* bope, R* ghcof, R* lu, R* u, R* mu, R* la, R h, R* strx, R* stry, R* strz)
{
   if (corder)
      rhs4sg_rev(ib, ie, jb, je, kb, ke, nk, onesided, acof, bope, ghcof, lu, u, mu, la, h, strx, stry, strz);
   else
      rhs4sg(ib, ie, jb, je, kb, ke, nk, onesided, acof, bope, ghcof, lu, u, mu, la, h, strx, stry, strz);
}
void ref_addsgd4(int corder, int ib, int ie, int jb, int je, int kb, int ke, R* up, R* u, R* um, R* rho,
                 R* dcx, R* dcy, R* dcz, R* strx, R* stry, R* strz, R* cox, R* coy, R* coz, R beta)
{
   if (corder)
      addsgd4fort_rev(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta);
   else
      addsgd4fort(ib, ie, jb, je, kb, ke, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta);
}
void ref_rhs4sgcurv(int corder, int ib, int ie, int jb, int je, int kb, int ke, R* u, R* mu, R* la, R* met,
                    R* jac, R* lu, int* onesided, R* acof, R* bope, R* ghcof, R* strx, R* stry)
{
   if (corder)
      rhs4sgcurv_rev(ib, ie, jb, je, kb, ke, u, mu, la, met, jac, lu, onesided, acof, bope, ghcof, strx, stry);
   else
      rhs4sgcurv(ib, ie, jb, je, kb, ke, u, mu, la, met, jac, lu, onesided, acof, bope, ghcof, strx, stry);
}
int ref_metric(int corder, int ib, int ie, int jb, int je, int kb, int ke, R* x, R* y, R* z, R* met, R* jac)
{
   return corder ? metric_rev(ib, ie, jb, je, kb, ke, x, y, z, met, jac)
                 : metric(ib, ie, jb, je, kb, ke, x, y, z, met, jac);
}
}
