// setup.h -- host-side set-up of one Cartesian grid (see setup.cpp).
#pragma once
#include <vector>

namespace sw4amd {

// boundaryConditionType values of the reference (src/double/sw4.h:37-38);
// they are part of the bcfortsg ABI.
// BC_CCINTERFACE: the side of a grid that a curvilinear grid shares with
// the Cartesian grid below it, set by EW::enforceCartTopo, not by bcfortsg.
enum BcType { BC_FREE = 0, BC_DIRICHLET = 1, BC_SUPERGRID = 2, BC_PERIODIC = 3, BC_CCINTERFACE = 4, BC_NONE = 8 };

// Time functions supported on the hot path (timeDep subset).
enum TimeFn { TF_RICKER = 0, TF_GAUSSIAN = 1, TF_C6SMOOTHBUMP = 2 };

struct GridSetup {
   int nx, ny, nz;     // interior points; arrays span [-1, n+2] per axis
   double h;
   double zmin;        // z of the plane k=1
   int bc[6];          // BcType per side (x-,x+,y-,y+,z-,z+)
   int sg_gp_thickness;
};

struct Taper {
   bool left = false, right = false;
   double x0 = 0, x1 = 1, width = 0.1, trans = 0.05, eps = 1e-4;
   void define(bool l, double start, bool r, double end, double w);
   static double psi0(double xi);
   double ramp(double x, double w) const;
   double lin(double x) const;
   double stretching(double x) const;
   double damping(double x) const;
   double corner(double x) const;
};

struct SupergridArrays {
   std::vector<double> dcx, dcy, dcz, strx, stry, strz, cox, coy, coz;
};

struct MomentSource {
   double x, y, z;
   double m[6];  // xx, xy, xz, yy, yz, zz (already scaled by m0)
   double t0, freq;
   int tfkind;
};

struct GridPointSource {
   int i, j, k;
   double f[3];
   double freq, t0;
   int tfkind;
};

void make_supergrid_arrays(const GridSetup& g, SupergridArrays& a);
// Weights of the regularised delta and its derivative on the 6-point source
// stencil at offset a, and their one-sided forms next to a free surface
// (Source::getsourcewgh / getsourcedwgh / ...low, Source.C:553-680).
void source_delta_weights(double a, double w[6]);
void source_ddelta_weights(double a, double w[6]);
void source_delta_weights_low(double a, double w[6]);
void source_ddelta_weights_low(double a, double w[6]);
void make_boundary_windows(const GridSetup& g, int wind[36], int npts[6]);
int moment_source_points(const GridSetup& g, const MomentSource& src, std::vector<GridPointSource>& out);
// Curvilinear grid of depth zgrid (over the box, i fastest): 0, or -1 grid
// too small, -2 source outside the grid, -3 folded grid at the source.
int moment_source_points_curv(const GridSetup& g, const double* zgrid, const MomentSource& src,
                              std::vector<GridPointSource>& out);
void sort_grid_point_sources(const GridSetup& g, std::vector<GridPointSource>& src, std::vector<int>& first);

}  // namespace sw4amd
