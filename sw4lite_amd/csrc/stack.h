// stack.h -- a curvilinear grid under the surface stacked on a Cartesian
// grid, stepped together as the reference steps its grids when topography
// exists (EW::timesteploop, EW.C:2527-2842): every stage runs on both grids,
// and each boundary pass (EW::enforceBC, EW.C:3446-3501) ends with the
// interface injection EW::enforceCartTopo (EW.C:3504-3531), which gives the
// Cartesian grid its ghost planes above the common plane from the
// curvilinear grid and the curvilinear grid its common plane and the ghost
// planes below it from the Cartesian grid.
#pragma once
#include "ew.h"

namespace sw4amd {

class GridStack {
public:
   // cart: the Cartesian grid (side 4 BC_CCINTERFACE); curv: the curvilinear
   // grid (side 5 BC_CCINTERFACE, topography set), same nx, ny, h, layout and
   // stream.  Both set up by setup() below.
   GridStack(EW& cart, EW& curv) : m_cart(cart), m_curv(curv) {}

   // dt: the smaller stable step of the two grids (EW::computeDT,
   // EW.C:5041-5146), then each grid's setup with it
   int setup(double tmax, int nsteps);
   // n steps in the reference stage order on both grids
   int step_reference(int n);
   EW& cart() { return m_cart; }
   EW& curv() { return m_curv; }

private:
   int enforce_bc(double t);  // both grids' boundary passes, then the interface
   EW& m_cart;
   EW& m_curv;
};

}  // namespace sw4amd
