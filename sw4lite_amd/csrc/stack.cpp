// stack.cpp -- two stacked grids stepped in the reference stage order.
#include "stack.h"

#include <algorithm>

#include "sw4hip_internal.h"

namespace sw4amd {

int GridStack::setup(double tmax, int nsteps)
{
   const GridSetup& a = m_cart.m_opt.grid;
   const GridSetup& b = m_curv.m_opt.grid;
   if (!m_curv.curvilinear() || m_cart.curvilinear())
      return set_error(E_INVAL, "grid stack: the upper grid must be curvilinear and the lower one Cartesian");
   if (a.nx != b.nx || a.ny != b.ny || a.h != b.h || m_cart.m_opt.corder != m_curv.m_opt.corder)
      return set_error(E_INVAL, "grid stack: the two grids must share nx, ny, h and the array layout");
   if (a.bc[4] != BC_CCINTERFACE || b.bc[5] != BC_CCINTERFACE)
      return set_error(E_INVAL, "grid stack: the common plane must be side 4 of the Cartesian grid and side 5 of the "
                                "curvilinear one");
   if (m_cart.m_stream != m_curv.m_stream) return set_error(E_INVAL, "grid stack: both grids must use one stream");
   double dta = 0, dtb = 0;
   int rc;
   if ((rc = m_cart.stable_dt(&dta)) || (rc = m_curv.stable_dt(&dtb))) return rc;
   const double dt = std::min(dta, dtb);
   if ((rc = m_cart.setup(tmax, nsteps, dt)) || (rc = m_curv.setup(tmax, nsteps, dt))) return rc;
   // the initial data of the receivers were recorded by each setup; both
   // grids start at zero, so no interface pass is needed before the first step
   return OK;
}

int GridStack::enforce_bc(double t)
{
   int rc;
   if ((rc = m_cart.enforceBC(m_cart.mUp, m_cart.mMu, m_cart.mLambda, t)) ||
       (rc = m_curv.enforceBC(m_curv.mUp, m_curv.mMu, m_curv.mLambda, t)))
      return rc;
   return launch_enforce_cart_topo(m_cart.m_opt.corder, m_cart.m_box, m_cart.mUp.dev, m_curv.m_box, m_curv.mUp.dev, 2,
                                   m_cart.m_stream);
}

int GridStack::step_reference(int n)
{
   EW* g[2] = {&m_cart, &m_curv};
   int rc;
   for (EW* e : g) {
      if ((rc = e->comp_end())) return rc;
      if (!e->m_setup_done) return set_error(E_STATE, "grid stack: step before setup");
      if ((rc = e->ensure_reference_arrays())) return rc;
   }
   for (int s = 0; s < n; s++) {
      const double t = m_cart.mT, dt = m_cart.mDt;
      for (EW* e : g)
         if ((rc = e->Force(t, e->mF, false)) || (rc = e->evalRHS(e->mU, e->mMu, e->mLambda, e->mLu)) ||
             (rc = e->evalPredictor(e->mUp, e->mU, e->mUm, e->mRho, e->mLu, e->mF)))
            return rc;
      if ((rc = enforce_bc(t + dt))) return rc;
      for (EW* e : g)
         if ((rc = e->Force(t, e->mF, true)) || (rc = e->evalDpDmInTime(e->mUp, e->mU, e->mUm, e->mUacc)) ||
             (rc = e->evalRHS(e->mUacc, e->mMu, e->mLambda, e->mLu)) ||
             (rc = e->evalCorrector(e->mUp, e->mRho, e->mLu, e->mF)) ||
             (rc = e->addSuperGridDamping(e->mUp, e->mU, e->mUm, e->mRho)))
            return rc;
      if ((rc = enforce_bc(t + dt))) return rc;
      for (EW* e : g) {
         if ((rc = e->record(e->mUp, e->mUm))) return rc;
         e->mT += e->mDt;
         e->mCycle++;
         e->cycleSolutionArrays();
      }
   }
   return OK;
}

}  // namespace sw4amd
