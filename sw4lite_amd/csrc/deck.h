// deck.h -- input-deck front end: the reference's deck commands
// (EW::parseInputFile, EW.C:2171-2285) for one Cartesian grid, or with a
// Gaussian-hill topography for a curvilinear grid stacked on the Cartesian
// one, run through EW objects, with receiver output (TimeSeries) and
// checkpoint/restart files (CheckPoint.C) in the reference's formats.
#pragma once
#include <string>
#include <vector>

#include "timeseries.h"
#include "topo.h"

namespace sw4amd {

struct DeckBlock {  // block (EW::processMaterialBlock, EW.C:1385-1600)
   double vp = -1, vs = -1, rho = -1, vpgrad = 0, vsgrad = 0, rhograd = 0;
   double x1 = -1e38, x2 = 1e38, y1 = -1e38, y2 = 1e38, z1 = -1e38, z2 = 1e38;
   bool absdepth = false;  // absdepth=1: z bounds are z, not depth below the surface
};

struct DeckSource {  // source (EW::processSource, EW.C:951-1232), moment tensors
   double x = 0, y = 0, z = 0, m0 = 1, t0 = 0, freq = 1;
   double m[6] = {0, 0, 0, 0, 0, 0};  // xx, xy, xz, yy, yz, zz
   int tfkind = -1;                   // 0 Ricker, 1 Gaussian, 2 C6SmoothBump
   std::string type = "RickerInt";
   bool topodepth = false;            // depth= given: z below the surface
};

struct DeckCheckPoint {  // checkpoint (EW::processCheckPoint, EW.C:783-848)
   double time = 0, time_interval = 0;
   int cycle = -1, cycle_interval = 0;
   std::string file = "restart";
};

struct Deck {
   // grid
   double h = 0, x = 0, y = 0, z = 0;
   int nx = 0, ny = 0, nz = 0;
   // time: t (tmax > 0) or steps
   double tmax = 0;
   int steps = -1;
   // fileio
   std::string path = ".";
   // supergrid
   int sg_gp = 30;
   double sg_dc = 0.02;
   // developer
   double cfl = 1.3;
   int corder = 0;
   // testpointsource
   bool point_source_test = false, halfspace = false;
   double tp_rho = 1.0, tp_cs = 1.0, tp_cp = 1.7320508075688772;
   std::vector<DeckBlock> blocks;
   std::vector<DeckSource> sources;
   std::vector<Station> stations;
   std::vector<char> station_topodepth;  // per station: depth= given
   // topography input=gaussian (EW::processTopography, EW.C:535-725)
   bool topography = false;
   Topography topo;
   std::vector<DeckCheckPoint> checkpoints;
   std::string restart;  // restart file=
};

// Parse deck text; commands outside this front end (topography, dgalerkin,
// options the EW object does not model) are refused with a message naming
// the line.  Returns 0 or a status code.
int parse_deck(const std::string& text, Deck& d);

struct DeckResult {
   int nx = 0, ny = 0, nz = 0;  // nz: the Cartesian grid's
   double h = 0, dt = 0, t = 0;
   int nsteps = 0, cycle = 0, nstations = 0, ncheckpoints = 0;
   int ngrids = 1, nz_curv = 0;  // with topography: 2 and the curvilinear grid's nz
   bool point_source_test = false;
   double linf = 0, l2 = 0, solinf = 0;
};

// Run a parsed deck on the current GPU: grid, material, sources,
// receivers, dt, optional restart, the time loop (mode 0 fused sweeps, 1 the
// reference stage order, 2 FP32 with the compensated update), checkpoints,
// then the receiver files and, for a point-source test, PointSourceErr.txt.
// path_override, when not empty, replaces the deck's fileio path.
int run_deck(const Deck& d, const std::string& path_override, int mode, DeckResult& out);

// CheckPoint::write_checkpoint / read_checkpoint (CheckPoint.C:251-480): one
// grid, header [precision 8, grids 1, time, cycle, 1 nx 1 ny 1 nz], then Um
// and U over the interior points, components of a point together, i
// fastest, then j, then k.  um, u: host arrays over the box [-1, n+2] in the
// component-blocked layout (component c at c * npts).
int write_checkpoint(const std::string& file, double time, int cycle, int nx, int ny, int nz, const double* um,
                     const double* u);
int read_checkpoint(const std::string& file, int nx, int ny, int nz, double* time, int* cycle, double* um,
                    double* u);
// file name of a checkpoint: prefix.cycle=<cycle, zero padded to the digits
// of nsteps>.sw4checkpoint (CheckPoint::compute_file_suffix)
std::string checkpoint_name(const std::string& prefix, int cycle, int nsteps);

}  // namespace sw4amd
