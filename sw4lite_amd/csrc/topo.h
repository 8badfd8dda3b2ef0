// topo.h -- the curvilinear grid under a topography surface, stacked on the
// Cartesian grid as the reference builds it for `topography input=gaussian`
// (EW::processTopography, EW.C:535-725): the analytic Gaussian hill, the grid
// sizes (EW::defineDimensionsZ, EW.C:2011-2060), the grid mapping and its
// inverse (EW::grid_mapping / invert_grid_mapping, EW.C:5694-5790), the
// nearest grid point of a location (EW::computeNearestGridPoint,
// EW.C:5149-5226; TimeSeries.C:112-183, 307-330) and the moment-source
// discretisation on the curvilinear grid (Source::set_grid_point_sources4
// with Source::compute_metric_at_source, Source.C:1074-1960, 2191-2330).
// Host code, run once at set-up.
#pragma once
#include <vector>

#include "setup.h"

namespace sw4amd {

// Gaussian hill: elevation amp exp(-((x-xc)/lx)^2 - ((y-yc)/ly)^2), z = -elevation
struct Topography {
   double zmax = 0;            // zmax=: z of the plane shared with the Cartesian grid
   int order = 4;              // order=: m_grid_interpolation_order (2..7)
   double zeta_break = 0.95;   // zetabreak=
   double amp = 0, xc = 0, yc = 0, lx = 1, ly = 1;
   // Elevation at index coordinates (q, r) as EW::interpolate_topography
   // evaluates the analytic hill (EW.C:5617-5625).
   double elevation(double q, double r, double h) const;
   // The sampled hill at grid point (i, j) (mTopoGridExt of
   // EW::buildGaussianHillTopography, EW.C:6114-6122).
   double sample(int i, int j, double h) const;
};

struct StackedGrids {
   int nx = 0, ny = 0;
   double h = 0;
   int nz_cart = 0;    // Cartesian grid: k = 1 at z = zmin_cart
   int nz_curv = 0;    // curvilinear grid: k = 1 on the surface, k = nz_curv at z = zmin_cart
   double zmin_cart = 0;
   double zmax = 0;    // bottom of the Cartesian grid (m_global_zmax)
};

// Sizes of both grids for a domain of depth zdomain (grid z=) and nx x ny
// points of spacing h.  Returns 0, or E_INVAL with the message set.
int stacked_grid_sizes(const Topography& t, int nx, int ny, double h, double zdomain, StackedGrids& s);

// z of the curvilinear grid over its box [-1, nx+2] x [-1, ny+2] x
// [-1, nz_curv+2], i fastest (EW::generate_grid, EW.C:6164-6205).
void curvilinear_grid_z(const Topography& t, const StackedGrids& s, std::vector<double>& z);

// z of the surface at (x, y) (EW::find_topo_zcoord_all, EW.C:6336-6358);
// false outside the grid and its ghost points.
bool topo_surface_z(const Topography& t, const StackedGrids& s, double x, double y, double& z);

// Index coordinates (q, r, sk) of (x, y, z) in the curvilinear grid
// (EW::invert_grid_mapping, EW.C:5722-5790); false where the Newton iteration
// does not converge or (x, y) is off the grid.
bool curvilinear_invert(const Topography& t, const StackedGrids& s, double x, double y, double z, double& q,
                        double& r, double& sk);

// The grid (0 Cartesian, 1 curvilinear) and nearest grid point of a
// receiver at (x, y, z), z below the surface when topodepth (as
// TimeSeries::TimeSeries does, TimeSeries.C:112-170).  Returns 0; 1 when
// the reference ignores the station (above the surface); E_INVAL off the
// grid.
int stacked_receiver_point(const Topography& t, const StackedGrids& s, double x, double y, double z, bool topodepth,
                           int& grid, int& i, int& j, int& k);

// Grid-point sources of a moment tensor at (x, y, z) in the stacked grids
// (z already below the surface, Source::adjust_zcoord).  cart, curv receive
// the grid-point sources of each grid (a source whose stencil reaches the
// common plane contributes to both, Source.C:1927-1948).  jac: the
// curvilinear grid's Jacobian over its box (metric, EW::metric_rev).
// Returns 0, 1 when the reference ignores the source (above the surface),
// or E_INVAL with the message set.
int stacked_moment_source(const Topography& t, const StackedGrids& s, const double* jac, const MomentSource& src,
                          std::vector<GridPointSource>& cart, std::vector<GridPointSource>& curv);

}  // namespace sw4amd
