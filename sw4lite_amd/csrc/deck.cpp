// deck.cpp -- the reference's input-deck commands (EW::parseInputFile,
// EW.C:2171-2285) for one Cartesian grid, or for a curvilinear grid under a
// Gaussian-hill topography stacked on the Cartesian one, run through EW
// objects.
#include "deck.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "ew.h"
#include "stack.h"

namespace sw4amd {

namespace {

bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

// split a line into whitespace-separated tokens, dropping a trailing comment
std::vector<std::string> tokens(const std::string& line)
{
   std::vector<std::string> out;
   std::istringstream is(line);
   std::string t;
   while (is >> t) {
      if (t[0] == '#') break;
      out.push_back(t);
   }
   return out;
}

// value of "key=value" as a double / int / string (atof / atoi as the
// reference reads them)
double fval(const std::string& t, size_t n) { return std::atof(t.c_str() + n); }
int ival(const std::string& t, size_t n) { return std::atoi(t.c_str() + n); }
std::string sval(const std::string& t, size_t n) { return t.substr(n); }

int bad(int lineno, const std::string& what)
{
   return set_error(E_INVAL, "deck line " + std::to_string(lineno) + ": " + what);
}

// EW::computeEndGridPoint (EW.C:227-238)
int end_grid_point(double maxval, double h)
{
   const double reltol = 1e-5, abstol = 1e-12;
   const double fnpts = std::round(maxval / h + 1);
   if (std::fabs((fnpts - 1) * h - maxval) < reltol * std::fabs(maxval) + abstol) return (int)fnpts;
   return (int)fnpts + 1;
}

int parse_grid(const std::vector<std::string>& tk, int ln, Deck& d)
{
   double h = 0, x = 0, y = 0, z = 0;
   int nx = 0, ny = 0, nz = 0;
   for (size_t q = 1; q < tk.size(); q++) {
      const std::string& t = tk[q];
      if (starts(t, "ny=")) ny = ival(t, 3);
      else if (starts(t, "nx=")) nx = ival(t, 3);
      else if (starts(t, "nz=")) nz = ival(t, 3);
      else if (starts(t, "x=")) x = fval(t, 2);
      else if (starts(t, "y=")) y = fval(t, 2);
      else if (starts(t, "z=")) z = fval(t, 2);
      else if (starts(t, "h=")) h = fval(t, 2);
      else return bad(ln, "grid: option not supported: " + t);
   }
   // EW::processGrid (EW.C:343-474)
   const char* err = "grid: give h with x, y, z, or h with nx, ny, nz, or x, y, z with one of nx, ny, nz";
   if (h != 0.0) {
      if (nx > 0 || nz > 0 || ny > 0) {
         if (!(nx && nz) || !(x == 0.0 && y == 0.0 && z == 0.0)) return bad(ln, err);
      } else if (!(x > 0.0 && z > 0.0)) {
         return bad(ln, err);
      }
   } else if (!(x > 0.0 && z > 0.0) || (nx > 0) + (ny > 0) + (nz > 0) != 1) {
      return bad(ln, err);
   }
   int nxp, nyp, nzp;
   if (nx > 0 && h == 0.0) {
      h = x / (nx - 1);
      nxp = nx;
      nzp = end_grid_point(z, h);
      nyp = end_grid_point(y, h);
   } else if (ny > 0 && h == 0.0) {
      h = y / (ny - 1);
      nyp = ny;
      nxp = end_grid_point(x, h);
      nzp = end_grid_point(z, h);
   } else if (nz > 0 && h == 0.0) {
      h = z / (nz - 1);
      nzp = nz;
      nxp = end_grid_point(x, h);
      nyp = end_grid_point(y, h);
   } else {
      if (nx == 0 && x != 0.0) nxp = end_grid_point(x, h);
      else if (nx != 0) nxp = nx;
      else return bad(ln, err);
      if (nz == 0 && z != 0.0) nzp = end_grid_point(z, h);
      else if (nz != 0) nzp = nz;
      else return bad(ln, err);
      if (ny == 0 && y != 0.0) nyp = end_grid_point(y, h);
      else if (ny != 0) nyp = ny;
      else return bad(ln, err);
   }
   d.h = h;
   d.nx = nxp;
   d.ny = nyp;
   d.nz = nzp;
   d.x = (nxp - 1) * h;
   d.y = (nyp - 1) * h;
   d.z = (nzp - 1) * h;
   return OK;
}

int parse_source(const std::vector<std::string>& tk, int ln, Deck& d)
{
   DeckSource s;
   bool zset = false, depthset = false;
   double depth = 0;
   for (size_t q = 1; q < tk.size(); q++) {
      const std::string& t = tk[q];
      if (starts(t, "m0=")) s.m0 = fval(t, 3);
      else if (starts(t, "x=")) s.x = fval(t, 2);
      else if (starts(t, "y=")) s.y = fval(t, 2);
      else if (starts(t, "z=")) { s.z = fval(t, 2); zset = true; }
      else if (starts(t, "depth=")) { depth = fval(t, 6); depthset = true; }
      else if (starts(t, "Mxx=") || starts(t, "mxx=")) s.m[0] = fval(t, 4);
      else if (starts(t, "Mxy=") || starts(t, "mxy=")) s.m[1] = fval(t, 4);
      else if (starts(t, "Mxz=") || starts(t, "mxz=")) s.m[2] = fval(t, 4);
      else if (starts(t, "Myy=") || starts(t, "myy=")) s.m[3] = fval(t, 4);
      else if (starts(t, "Myz=") || starts(t, "myz=")) s.m[4] = fval(t, 4);
      else if (starts(t, "Mzz=") || starts(t, "mzz=")) s.m[5] = fval(t, 4);
      else if (starts(t, "t0=")) s.t0 = fval(t, 3);
      else if (starts(t, "freq=")) {
         s.freq = fval(t, 5);
         if (!(s.freq > 0)) return bad(ln, "source: freq must be positive");
      } else if (starts(t, "type=")) s.type = sval(t, 5);
      else if (starts(t, "Fx=") || starts(t, "fx=") || starts(t, "Fy=") || starts(t, "fy=") || starts(t, "Fz=") ||
               starts(t, "fz=") || starts(t, "f0="))
         return bad(ln, "source: point forces are not supported (moment tensors only)");
      else return bad(ln, "source: option not supported: " + t);
   }
   if (!zset && !depthset) return bad(ln, "source: depth or z must be given");
   if (depthset) {
      // depth below the surface (EW.C:1026-1034); without topography it is z
      if (depth < 0) return bad(ln, "source: depth below the surface must be >= 0");
      s.z = depth;
      s.topodepth = true;
   }
   if (s.type == "Ricker") s.tfkind = 0;
   else if (s.type == "Gaussian") s.tfkind = 1;
   else if (s.type == "C6SmoothBump") s.tfkind = 2;
   else return bad(ln, "source: time function " + s.type + " is not supported (Ricker, Gaussian, C6SmoothBump)");
   for (double& v : s.m) v *= s.m0;  // EW.C:1189-1195
   d.sources.push_back(s);
   return OK;
}

int parse_receiver(const std::vector<std::string>& tk, int ln, Deck& d)
{
   Station s;
   s.file = "station";
   bool named = false, zset = false, topodepth = false;
   for (size_t q = 1; q < tk.size(); q++) {
      const std::string& t = tk[q];
      if (starts(t, "x=")) s.x = fval(t, 2);
      else if (starts(t, "y=")) s.y = fval(t, 2);
      else if (starts(t, "z=")) { s.z = fval(t, 2); zset = true; topodepth = false; }
      else if (starts(t, "depth=") || starts(t, "topodepth=")) {
         s.z = fval(t, starts(t, "depth=") ? 6 : 10);
         if (s.z < 0) return bad(ln, "rec: depth must be >= 0");
         zset = true;
         topodepth = true;
      }
      else if (starts(t, "file=")) s.file = sval(t, 5);
      else if (starts(t, "sta=")) { s.sta = sval(t, 4); named = true; }
      else if (starts(t, "nsew=")) s.xyz = ival(t, 5) != 1;
      else if (starts(t, "writeEvery=")) continue;  // the files are written at the end of the run
      else if (starts(t, "usgsformat=")) s.usgs = ival(t, 11) != 0;
      else if (starts(t, "sacformat=")) s.sac = ival(t, 10) != 0;
      else if (starts(t, "variables=")) {
         const std::string v = sval(t, 10);
         if (v == "displacement") s.mode = 0;
         else if (v == "velocity") s.mode = 1;
         else return bad(ln, "rec: variables=" + v + " is not supported (displacement, velocity)");
      } else return bad(ln, "rec: option not supported: " + t);
   }
   (void)zset;
   if (!named) s.sta = s.file;
   d.stations.push_back(s);
   d.station_topodepth.push_back(topodepth);
   return OK;
}

int parse_block(const std::vector<std::string>& tk, int ln, Deck& d)
{
   DeckBlock b;
   for (size_t q = 1; q < tk.size(); q++) {
      const std::string& t = tk[q];
      if (starts(t, "rhograd=")) b.rhograd = fval(t, 8);
      else if (starts(t, "vpgrad=")) b.vpgrad = fval(t, 7);
      else if (starts(t, "vsgrad=")) b.vsgrad = fval(t, 7);
      else if (starts(t, "vp=")) b.vp = fval(t, 3);
      else if (starts(t, "vs=")) b.vs = fval(t, 3);
      else if (starts(t, "rho=")) b.rho = fval(t, 4);
      else if (starts(t, "r=")) b.rho = fval(t, 2);
      else if (starts(t, "x1=")) b.x1 = fval(t, 3);
      else if (starts(t, "x2=")) b.x2 = fval(t, 3);
      else if (starts(t, "y1=")) b.y1 = fval(t, 3);
      else if (starts(t, "y2=")) b.y2 = fval(t, 3);
      else if (starts(t, "z1=")) b.z1 = fval(t, 3);
      else if (starts(t, "z2=")) b.z2 = fval(t, 3);
      else if (starts(t, "absdepth=")) b.absdepth = ival(t, 9) != 0;
      else if (starts(t, "Qs=") || starts(t, "qs=") || starts(t, "Qp=") || starts(t, "qp=")) continue;  // no attenuation
      else return bad(ln, "block: option not supported: " + t);
   }
   if (b.x2 < b.x1 || b.y2 < b.y1 || b.z2 < b.z1) return bad(ln, "block: upper bound below lower bound");
   d.blocks.push_back(b);
   return OK;
}

// topography (EW::processTopography, EW.C:535-725): the Gaussian hill
int parse_topography(const std::vector<std::string>& tk, int ln, Deck& d)
{
   Topography t;
   bool gaussian = false, zmax = false;
   for (size_t q = 1; q < tk.size(); q++) {
      const std::string& v = tk[q];
      if (starts(v, "zmax=")) { t.zmax = fval(v, 5); zmax = true; }
      else if (starts(v, "order=")) {
         t.order = ival(v, 6);
         if (t.order < 2 || t.order > 7) return bad(ln, "topography: order must be 2, 3, 4, 5, 6 or 7");
      } else if (starts(v, "zetabreak=")) {
         t.zeta_break = fval(v, 10);
         if (!(t.zeta_break > 0 && t.zeta_break <= 1)) return bad(ln, "topography: zetabreak must be in (0, 1]");
      } else if (starts(v, "input=")) {
         if (sval(v, 6) != "gaussian") return bad(ln, "topography: input=" + sval(v, 6) + " is not supported (gaussian)");
         gaussian = true;
      } else if (starts(v, "file=")) continue;  // not read for the analytic hill
      else if (starts(v, "gaussianAmp=")) t.amp = fval(v, 12);
      else if (starts(v, "gaussianXc=")) t.xc = fval(v, 11);
      else if (starts(v, "gaussianYc=")) t.yc = fval(v, 11);
      else if (starts(v, "gaussianLx=")) t.lx = fval(v, 11);
      else if (starts(v, "gaussianLy=")) t.ly = fval(v, 11);
      else if (starts(v, "analyticalMetric=")) {
         const std::string a = sval(v, 17);
         if (a == "1" || a == "true" || a == "yes")
            return bad(ln, "topography: analyticalMetric (metricexgh) is not supported; the metric is computed from "
                           "the grid");
      } else return bad(ln, "topography: option not supported: " + v);
   }
   if (!gaussian) return bad(ln, "topography: input=gaussian must be given");
   if (!zmax) return bad(ln, "topography: zmax must be given");
   if (!(t.lx > 0 && t.ly > 0)) return bad(ln, "topography: gaussianLx and gaussianLy must be positive");
   d.topography = true;
   d.topo = t;
   return OK;
}

}  // namespace

int parse_deck(const std::string& text, Deck& d)
{
   std::istringstream in(text);
   std::string line;
   int ln = 0, rc;
   bool have_grid = false, have_time = false;
   while (std::getline(in, line)) {
      ln++;
      const std::vector<std::string> tk = tokens(line);
      if (tk.empty()) continue;
      const std::string& c = tk[0];
      if (c == "grid") {
         if ((rc = parse_grid(tk, ln, d))) return rc;
         have_grid = true;
      } else if (c == "time") {
         double t = 0;
         int steps = -1;
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "t=")) t = fval(tk[q], 2);
            else if (starts(tk[q], "steps=")) steps = ival(tk[q], 6);
            else return bad(ln, "time: option not supported: " + tk[q]);
         }
         if (t > 0.0 && steps >= 0) return bad(ln, "time: cannot set both t and steps");
         if (t < 0 || (t == 0 && steps < 0)) return bad(ln, "time: give t > 0 or steps >= 0");
         d.tmax = t > 0 ? t : 0;
         d.steps = t > 0 ? -1 : steps;
         have_time = true;
      } else if (c == "fileio") {
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "path=")) d.path = sval(tk[q], 5) + "/";
            else if (starts(tk[q], "verbose=") || starts(tk[q], "printcycle=") || starts(tk[q], "pfs=") ||
                     starts(tk[q], "nwriters="))
               continue;
            else return bad(ln, "fileio: option not supported: " + tk[q]);
         }
      } else if (c == "supergrid") {
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "gp=")) {
               d.sg_gp = ival(tk[q], 3);
               if (d.sg_gp <= 0) return bad(ln, "supergrid: gp must be positive");
            } else if (starts(tk[q], "dc=")) {
               d.sg_dc = fval(tk[q], 3);
               if (d.sg_dc < 0) return bad(ln, "supergrid: dc must be non-negative");
            } else return bad(ln, "supergrid: option not supported: " + tk[q]);
         }
      } else if (c == "developer") {
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "cfl=")) d.cfl = fval(tk[q], 4);
            else if (starts(tk[q], "corder=")) {
               const std::string v = sval(tk[q], 7);  // EW.C:1364-1369
               d.corder = v == "yes" || v == "1" || v == "on";
            }
            else if (starts(tk[q], "checkfornan=") || starts(tk[q], "reporttiming=") || starts(tk[q], "trace=") ||
                     starts(tk[q], "thblocki=") || starts(tk[q], "thblockj=") || starts(tk[q], "thblockk="))
               continue;  // diagnostics and CPU thread blocking of the reference
            else return bad(ln, "developer: option not supported: " + tk[q]);
         }
      } else if (c == "testpointsource") {
         d.point_source_test = true;
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "cp=")) d.tp_cp = fval(tk[q], 3);
            else if (starts(tk[q], "cs=")) d.tp_cs = fval(tk[q], 3);
            else if (starts(tk[q], "rho=")) d.tp_rho = fval(tk[q], 4);
            else if (starts(tk[q], "halfspace=")) d.halfspace = sval(tk[q], 10) == "1" || sval(tk[q], 10) == "true";
            else if (starts(tk[q], "diractest=")) return bad(ln, "testpointsource: diractest is not supported");
            else return bad(ln, "testpointsource: option not supported: " + tk[q]);
         }
      } else if (c == "source") {
         if ((rc = parse_source(tk, ln, d))) return rc;
      } else if (c == "rec") {
         if ((rc = parse_receiver(tk, ln, d))) return rc;
      } else if (c == "block") {
         if ((rc = parse_block(tk, ln, d))) return rc;
      } else if (c == "checkpoint") {
         DeckCheckPoint p;
         bool timing = false;
         for (size_t q = 1; q < tk.size(); q++) {
            const std::string& t = tk[q];
            if (starts(t, "time=")) { p.time = fval(t, 5); timing = true; }
            else if (starts(t, "timeInterval=")) { p.time_interval = fval(t, 13); timing = true; }
            else if (starts(t, "cycle=")) { p.cycle = ival(t, 6); timing = true; }
            else if (starts(t, "cycleInterval=")) { p.cycle_interval = ival(t, 14); timing = true; }
            else if (starts(t, "file=")) p.file = sval(t, 5);
            else if (starts(t, "bufsize=")) continue;
            else return bad(ln, "checkpoint: option not supported: " + t);
         }
         if (!timing) return bad(ln, "checkpoint: give time, timeInterval, cycle or cycleInterval");
         d.checkpoints.push_back(p);
      } else if (c == "restart") {
         for (size_t q = 1; q < tk.size(); q++) {
            if (starts(tk[q], "file=")) d.restart = sval(tk[q], 5);
            else if (starts(tk[q], "bufsize=")) continue;
            else return bad(ln, "restart: option not supported: " + tk[q]);
         }
         if (d.restart.empty()) return bad(ln, "restart: file must be given");
      } else if (c == "topography") {
         if ((rc = parse_topography(tk, ln, d))) return rc;
      } else {
         return bad(ln, "command not supported: " + c);
      }
   }
   if (!have_grid) return set_error(E_INVAL, "deck: no grid command");
   if (!have_time) return set_error(E_INVAL, "deck: no time command");
   return OK;
}

// ---------------------------------------------------------------------------
// checkpoint files
// ---------------------------------------------------------------------------
std::string checkpoint_name(const std::string& prefix, int cycle, int nsteps)
{
   char buf[50];
   const int zeros = std::snprintf(buf, sizeof(buf), "%d", nsteps);
   std::ostringstream s;
   s << prefix << ".cycle=";
   int temp = (int)std::pow(10.0, zeros - 1);
   int test = cycle == 0 ? 1 : cycle;
   while (test < temp) {
      s << "0";
      temp /= 10;
   }
   s << cycle << ".sw4checkpoint";
   return s.str();
}

namespace {
// interior points of a component-blocked box array, components of a point together
void pack_interior(int nx, int ny, int nz, const double* a, std::vector<double>& out)
{
   const long ni = nx + 4, nij = ni * (ny + 4), npts = nij * (nz + 4);
   out.resize((size_t)3 * nx * ny * nz);
   size_t s = 0;
   for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
         for (int i = 1; i <= nx; i++) {
            const long p = (i + 1) + ni * (j + 1) + nij * (k + 1);
            for (int c = 0; c < 3; c++) out[s++] = a[p + npts * c];
         }
}
void unpack_interior(int nx, int ny, int nz, const std::vector<double>& in, double* a)
{
   const long ni = nx + 4, nij = ni * (ny + 4), npts = nij * (nz + 4);
   size_t s = 0;
   for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
         for (int i = 1; i <= nx; i++) {
            const long p = (i + 1) + ni * (j + 1) + nij * (k + 1);
            for (int c = 0; c < 3; c++) a[p + npts * c] = in[s++];
         }
}
}  // namespace

int write_checkpoint(const std::string& file, double time, int cycle, int nx, int ny, int nz, const double* um,
                     const double* u)
{
   FILE* fp = std::fopen(file.c_str(), "wb");
   if (!fp) return set_error(E_INVAL, "cannot open checkpoint file " + file + " for writing");
   const int prec = 8, ng = 1;
   const int sizes[6] = {1, nx, 1, ny, 1, nz};
   bool ok = std::fwrite(&prec, sizeof(int), 1, fp) == 1 && std::fwrite(&ng, sizeof(int), 1, fp) == 1 &&
             std::fwrite(&time, sizeof(double), 1, fp) == 1 && std::fwrite(&cycle, sizeof(int), 1, fp) == 1 &&
             std::fwrite(sizes, sizeof(int), 6, fp) == 6;
   std::vector<double> buf;
   for (const double* a : {um, u}) {
      if (!ok) break;
      pack_interior(nx, ny, nz, a, buf);
      ok = std::fwrite(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
   }
   if (std::fclose(fp) != 0 || !ok) return set_error(E_INVAL, "error writing checkpoint file " + file);
   return OK;
}

int read_checkpoint(const std::string& file, int nx, int ny, int nz, double* time, int* cycle, double* um, double* u)
{
   FILE* fp = std::fopen(file.c_str(), "rb");
   if (!fp) return set_error(E_INVAL, "cannot open checkpoint file " + file);
   int prec = 0, ng = 0, sizes[6];
   double t = 0;
   int cyc = 0;
   bool ok = std::fread(&prec, sizeof(int), 1, fp) == 1 && std::fread(&ng, sizeof(int), 1, fp) == 1 &&
             std::fread(&t, sizeof(double), 1, fp) == 1 && std::fread(&cyc, sizeof(int), 1, fp) == 1 &&
             std::fread(sizes, sizeof(int), 6, fp) == 6;
   if (!ok) {
      std::fclose(fp);
      return set_error(E_INVAL, "checkpoint file " + file + ": short header");
   }
   if (prec != 8 || ng != 1 || sizes[0] != 1 || sizes[1] != nx || sizes[2] != 1 || sizes[3] != ny ||
       sizes[4] != 1 || sizes[5] != nz) {
      std::fclose(fp);
      return set_error(E_INVAL, "checkpoint file " + file + ": precision, grid count or sizes do not match the grid");
   }
   std::vector<double> buf((size_t)3 * nx * ny * nz);
   for (double* a : {um, u}) {
      if (std::fread(buf.data(), sizeof(double), buf.size(), fp) != buf.size()) {
         std::fclose(fp);
         return set_error(E_INVAL, "checkpoint file " + file + ": short data");
      }
      if (a) unpack_interior(nx, ny, nz, buf, a);
   }
   std::fclose(fp);
   *time = t;
   *cycle = cyc;
   return OK;
}

// ---------------------------------------------------------------------------
// running a deck
// ---------------------------------------------------------------------------
namespace {

int make_dirs(const std::string& path)
{
   if (path.empty() || path == "." || path == "./") return OK;
   std::string cur;
   for (size_t q = 0; q < path.size(); q++) {
      cur.push_back(path[q]);
      if (path[q] == '/' || q + 1 == path.size()) {
         if (cur == "/" || cur == "./") continue;
         if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            return set_error(E_INVAL, "cannot create output directory " + cur);
      }
   }
   return OK;
}

// One grid for the material blocks: where its points are and which
// extrapolations the reference applies to it (EW::setup_materials,
// EW.C:5443-5466).
struct MatGrid {
   int nx = 0, ny = 0, nz = 0;
   double h = 0, zmin = 0;
   const double* zgrid = nullptr;     // curvilinear grid: z over the box
   const Topography* topo = nullptr;  // surface of a stacked Cartesian grid
   const StackedGrids* stack = nullptr;
   bool ext_low = true, ext_high = true;
};

// material of the blocks (MaterialBlock::set_material_properties,
// MaterialBlock.C:115-262): a Cartesian grid's interior columns at every k
// (depth below the surface, or z with absdepth=1), every point of a
// curvilinear grid (depth below the grid's top plane); -1 where no block
// reaches; ghosts extrapolated in z (the ends the grid has), then in x and y
// (EW::extrapolateInZ / extrapolateInXY, EW.C:5493-5550); then
// mu = rho cs^2, lambda = rho cp^2 - 2 mu
int block_material(const Deck& d, const MatGrid& g, std::vector<double>& rho, std::vector<double>& mu,
                   std::vector<double>& la)
{
   const int nx = g.nx, ny = g.ny, nz = g.nz;
   const long ni = nx + 4, nj = ny + 4, nk = nz + 4, nij = ni * nj, n = nij * nk;
   std::vector<double> cs(n, -1.0), cp(n, -1.0);
   rho.assign(n, -1.0);
   auto P = [&](int i, int j, int k) { return (i + 1) + ni * (j + 1) + nij * (k + 1); };
   const double tol = 1e-5;
   const int ij0 = g.zgrid ? -1 : 1, i1 = g.zgrid ? nx + 2 : nx, j1 = g.zgrid ? ny + 2 : ny;
   // surface z of each interior column of a stacked Cartesian grid
   std::vector<double> ztop;
   if (!g.zgrid && g.topo) {
      ztop.assign(ni * nj, 0.0);
      for (int j = 1; j <= ny; j++)
         for (int i = 1; i <= nx; i++) {
            double zt = 0;
            if (!topo_surface_z(*g.topo, *g.stack, (i - 1) * g.h, (j - 1) * g.h, zt))
               return set_error(E_INVAL, "deck: surface undefined over the grid");
            ztop[(i + 1) + ni * (j + 1)] = zt;
         }
   }
   for (const DeckBlock& b : d.blocks)
      for (int k = -1; k <= nz + 2; k++)
         for (int j = ij0; j <= j1; j++)
            for (int i = ij0; i <= i1; i++) {
               const double x = (i - 1) * g.h, y = (j - 1) * g.h;
               double depth;
               if (g.zgrid) {
                  const double z = g.zgrid[P(i, j, k)];
                  depth = b.absdepth ? z : z - g.zgrid[P(i, j, 1)];
               } else {
                  const double z = g.zmin + (k - 1) * g.h;
                  depth = b.absdepth || !g.topo ? z : z - ztop[(i + 1) + ni * (j + 1)];
               }
               if (b.x1 - tol <= x && x <= b.x2 + tol && b.y1 - tol <= y && y <= b.y2 + tol &&
                   b.z1 - tol <= depth && depth <= b.z2 + tol) {
                  const long p = P(i, j, k);
                  if (b.rho != -1) rho[p] = b.rho + b.rhograd * depth;
                  if (b.vs != -1) cs[p] = b.vs + b.vsgrad * depth;
                  if (b.vp != -1) cp[p] = b.vp + b.vpgrad * depth;
               }
            }
   for (std::vector<double>* f : {&rho, &cp, &cs}) {
      std::vector<double>& a = *f;
      for (int k = -1; k <= nz + 2; k++) {
         if (k >= 1 && k <= nz) continue;
         if ((k < 1 && !g.ext_low) || (k > nz && !g.ext_high)) continue;
         const int kk = k < 1 ? 1 : nz;
         for (int j = -1; j <= ny + 2; j++)
            for (int i = -1; i <= nx + 2; i++)
               if (a[P(i, j, k)] == -1) a[P(i, j, k)] = a[P(i, j, kk)];
      }
      for (int k = -1; k <= nz + 2; k++)
         for (int j = -1; j <= ny + 2; j++) {
            for (int i = -1; i < 1; i++)
               if (a[P(i, j, k)] == -1) a[P(i, j, k)] = a[P(1, j, k)];
            for (int i = nx + 1; i <= nx + 2; i++)
               if (a[P(i, j, k)] == -1) a[P(i, j, k)] = a[P(nx, j, k)];
         }
      for (int k = -1; k <= nz + 2; k++) {
         for (int j = -1; j < 1; j++)
            for (int i = -1; i <= nx + 2; i++)
               if (a[P(i, j, k)] == -1) a[P(i, j, k)] = a[P(i, 1, k)];
         for (int j = ny + 1; j <= ny + 2; j++)
            for (int i = -1; i <= nx + 2; i++)
               if (a[P(i, j, k)] == -1) a[P(i, j, k)] = a[P(i, ny, k)];
      }
   }
   for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
         for (int i = 1; i <= nx; i++) {
            const long p = P(i, j, k);
            if (rho[p] == -1 || cs[p] == -1 || cp[p] == -1)
               return set_error(E_INVAL, "deck: the blocks leave the material undefined at grid point (" +
                                             std::to_string(i) + "," + std::to_string(j) + "," + std::to_string(k) +
                                             ")" + (g.zgrid ? " of the curvilinear grid" : ""));
         }
   mu.resize(n);
   la.resize(n);
   for (long p = 0; p < n; p++) {
      mu[p] = rho[p] * cs[p] * cs[p];
      la[p] = rho[p] * cp[p] * cp[p] - 2 * mu[p];
   }
   return OK;
}

// The receiver files of the stations recorded by one EW object: station r
// of st is receiver idx[r] of e (idx[r] < 0: not recorded).  Sample i is
// labelled t = i dt from the start time 0, also on a restarted run
// (TimeSeries::allocateRecordingArrays, m_shift = mTstart - m_t0 = 0).
// Velocity stations record nothing for the initial data (EW.C:2488-2490):
// their first sample is the centred difference of the first step, so the
// object's initial sample is dropped.
int write_recorded(const std::string& path, EW& e, const std::vector<Station>& st, const std::vector<int>& idx)
{
   const int nrec = (int)e.m_rec.size() / 4;
   int ns = 0, rc;
   if ((rc = e.get_receivers(&ns, nullptr))) return rc;
   std::vector<double> rec((size_t)3 * nrec * ns);
   if (ns > 0 && nrec > 0 && (rc = e.get_receivers(&ns, rec.data()))) return rc;
   GeoFrame geo;
   for (size_t r = 0; r < st.size(); r++) {
      if (idx[r] < 0) continue;
      const int skip = st[r].mode == 1 ? 1 : 0;
      const int n = std::max(ns - skip, 0);
      std::vector<double> s((size_t)3 * n);
      for (int q = 0; q < n; q++)
         for (int c = 0; c < 3; c++) s[3 * q + c] = rec[((size_t)(q + skip) * nrec + idx[r]) * 3 + c];
      if ((rc = write_station(path == "./" ? std::string(".") : path, st[r], geo, 0.0, e.mDt, n, s.data())))
         return rc;
   }
   return OK;
}

// CheckPoint::timeToWrite (CheckPoint.C:211-232)
struct CkptState {
   DeckCheckPoint p;
   bool time_done = false;
   double next_time = 0.0;
   bool due(double time, int cycle, double dt)
   {
      bool doit = cycle == p.cycle;
      if (p.cycle_interval != 0 && cycle % p.cycle_interval == 0 && time >= 0.0) doit = true;
      if (p.time > 0.0 && p.time <= time + dt * 0.5 && !time_done) {
         time_done = true;
         doit = true;
      }
      if (p.time_interval != 0.0 && next_time <= time + dt * 0.5) {
         next_time += p.time_interval;
         if (time >= 0.0) doit = true;
      }
      return doit;
   }
};

}  // namespace

namespace {

// A deck with topography: the curvilinear grid under the Gaussian hill
// stacked on the Cartesian grid below topography zmax (EW::parseInputFile
// and EW::setupRun with two grids, EW.C:2171-2336), stepped in the
// reference stage order.
int run_deck_topo(const Deck& d, const std::string& path, int mode, DeckResult& out)
{
   if (mode == 2) return set_error(E_INVAL, "deck: the FP32 form steps one Cartesian grid; topography needs mode 0 or 1");
   if (d.point_source_test) return set_error(E_INVAL, "deck: testpointsource with topography is not supported");
   if (!d.checkpoints.empty() || !d.restart.empty())
      return set_error(E_INVAL, "deck: checkpoint and restart files of two grids are not supported");
   StackedGrids sg;
   int rc;
   if ((rc = stacked_grid_sizes(d.topo, d.nx, d.ny, d.h, d.z, sg))) return rc;
   EWOptions oc, ov;
   oc.grid.nx = ov.grid.nx = d.nx;
   oc.grid.ny = ov.grid.ny = d.ny;
   oc.grid.h = ov.grid.h = d.h;
   oc.grid.nz = sg.nz_cart;
   oc.grid.zmin = sg.zmin_cart;
   ov.grid.nz = sg.nz_curv;
   ov.grid.zmin = 0.0;
   // sides (EW::default_bcs, EW.C:6077-6082): supergrid around, the free
   // surface on top of the curvilinear grid, the common plane between them
   for (int q = 0; q < 6; q++) oc.grid.bc[q] = ov.grid.bc[q] = BC_SUPERGRID;
   oc.grid.bc[4] = BC_CCINTERFACE;
   ov.grid.bc[4] = BC_FREE;
   ov.grid.bc[5] = BC_CCINTERFACE;
   oc.grid.sg_gp_thickness = ov.grid.sg_gp_thickness = d.sg_gp;
   oc.corder = ov.corder = mode == 1 ? d.corder : 1;
   oc.beta = ov.beta = d.sg_dc;
   oc.cfl = ov.cfl = d.cfl;
   EW cart(oc), curv(ov);
   if ((rc = cart.init()) || (rc = curv.init())) return rc;
   std::vector<double> zc;
   curvilinear_grid_z(d.topo, sg, zc);
   if ((rc = curv.set_topography(zc.data(), false))) return rc;
   {
      std::vector<double> rho, mu, la;
      MatGrid mc;
      mc.nx = d.nx;
      mc.ny = d.ny;
      mc.nz = sg.nz_cart;
      mc.h = d.h;
      mc.zmin = sg.zmin_cart;
      mc.topo = &d.topo;
      mc.stack = &sg;
      mc.ext_low = false;
      if ((rc = block_material(d, mc, rho, mu, la)) ||
          (rc = cart.set_material(rho.data(), mu.data(), la.data(), false)))
         return rc;
      MatGrid mv = mc;
      mv.nz = sg.nz_curv;
      mv.zmin = 0;
      mv.zgrid = zc.data();
      mv.ext_low = true;
      mv.ext_high = false;
      if ((rc = block_material(d, mv, rho, mu, la)) ||
          (rc = curv.set_material(rho.data(), mu.data(), la.data(), false)))
         return rc;
   }
   // sources: the curvilinear grid's Jacobian weights them (Source.C:1798-1801)
   if (!d.sources.empty()) {
      std::vector<double> jac(curv.mJ.npts());
      if ((rc = check_hip(hipMemcpy(jac.data(), curv.mJ.dev, sizeof(double) * jac.size(), hipMemcpyDeviceToHost),
                          "Jacobian readback")))
         return rc;
      for (const DeckSource& s : d.sources) {
         if (s.x < 0 || s.x > d.x || s.y < 0 || s.y > d.y)
            return set_error(E_INVAL, "deck: source positioned outside the grid");
         MomentSource m;
         m.x = s.x;
         m.y = s.y;
         m.z = s.z;
         if (s.topodepth) {  // Source::adjust_zcoord (Source.C:517-531)
            double zt = 0;
            if (!topo_surface_z(d.topo, sg, s.x, s.y, zt)) return set_error(E_INVAL, "deck: source off the surface");
            m.z += zt;
         }
         if (m.z > sg.zmax) return set_error(E_INVAL, "deck: source below the grid");
         for (int q = 0; q < 6; q++) m.m[q] = s.m[q];
         m.t0 = s.t0;
         m.freq = s.freq;
         m.tfkind = s.tfkind;
         std::vector<GridPointSource> pc, pv;
         rc = stacked_moment_source(d.topo, sg, jac.data(), m, pc, pv);
         if (rc == 1) continue;  // above the surface: ignored as the reference does
         if (rc) return rc;
         for (const GridPointSource& p : pc)
            if ((rc = cart.add_grid_point_source(p))) return rc;
         for (const GridPointSource& p : pv)
            if ((rc = curv.add_grid_point_source(p))) return rc;
      }
   }
   std::vector<Station> st = d.stations;
   std::vector<int> in_cart(st.size(), -1), in_curv(st.size(), -1);
   int utc[7];
   utc_now(utc);
   for (size_t r = 0; r < st.size(); r++) {
      Station& s = st[r];
      if (s.x < 0 || s.x > d.x || s.y < 0 || s.y > d.y)
         return set_error(E_INVAL, "deck: receiver " + s.file + " outside the grid");
      int grid = 0, i = 0, j = 0, k = 0;
      rc = stacked_receiver_point(d.topo, sg, s.x, s.y, s.z, d.station_topodepth[r] != 0, grid, i, j, k);
      if (rc == 1) continue;  // above the surface: the reference writes no file
      if (rc) return set_error(E_INVAL, "deck: receiver " + s.file + " outside the grid");
      EW& e = grid == 0 ? cart : curv;
      int index = 0;
      if ((rc = e.add_receiver_ijk(i, j, k, s.mode, &index))) return rc;
      (grid == 0 ? in_cart : in_curv)[r] = index;
      // the grid point recorded (TimeSeries.C:172-187)
      s.gx = (i - 1) * d.h;
      s.gy = (j - 1) * d.h;
      s.gz = grid == 0 ? sg.zmin_cart + (k - 1) * d.h
                       : zc[(i + 1) + (long)(d.nx + 4) * (j + 1) + (long)(d.nx + 4) * (d.ny + 4) * (k + 1)];
      std::memcpy(s.utc, utc, sizeof(utc));
   }
   GridStack stack(cart, curv);
   if ((rc = stack.setup(d.tmax, d.steps >= 0 ? d.steps : 0))) return rc;
   if ((rc = make_dirs(path))) return rc;
   const int total = cart.mNumberOfTimeSteps;
   if ((rc = stack.step_reference(total))) return rc;
   out = DeckResult{};
   out.nx = d.nx;
   out.ny = d.ny;
   out.nz = sg.nz_cart;
   out.ngrids = 2;
   out.nz_curv = sg.nz_curv;
   out.h = d.h;
   out.dt = cart.mDt;
   out.t = cart.mT;
   out.nsteps = total;
   out.cycle = cart.mCycle;
   out.nstations = (int)st.size();
   if ((rc = write_recorded(path, cart, st, in_cart)) || (rc = write_recorded(path, curv, st, in_curv))) return rc;
   return OK;
}

}  // namespace

int run_deck(const Deck& d, const std::string& path_override, int mode, DeckResult& out)
{
   if (mode < 0 || mode > 2) return set_error(E_INVAL, "deck: mode must be 0, 1 or 2");
   const std::string path = path_override.empty() ? d.path
                                                  : (path_override.back() == '/' ? path_override : path_override + "/");
   if (d.topography) return run_deck_topo(d, path, mode, out);
   EWOptions o;
   o.grid.nx = d.nx;
   o.grid.ny = d.ny;
   o.grid.nz = d.nz;
   o.grid.h = d.h;
   o.grid.zmin = 0.0;
   // sides (EW::default_bcs, EW.C:6077-6082; processTestPointSource, EW.C:944-948)
   for (int s = 0; s < 6; s++) o.grid.bc[s] = 2;
   o.grid.bc[4] = d.point_source_test && !d.halfspace ? 2 : 0;
   o.grid.sg_gp_thickness = d.sg_gp;
   // the fused sweeps need the component-blocked layout; the arithmetic and
   // the results do not depend on it
   o.corder = mode == 1 ? d.corder : 1;
   o.beta = d.sg_dc;
   o.cfl = d.cfl;
   if (d.nx < 12 || d.ny < 12 || d.nz < 12) return set_error(E_INVAL, "deck: grid needs at least 12 points per axis");
   EW e(o);
   int rc;
   if ((rc = e.init())) return rc;
   if (d.point_source_test) {
      if ((rc = e.set_material_const(d.tp_rho, d.tp_cs, d.tp_cp))) return rc;
   } else {
      std::vector<double> rho, mu, la;
      MatGrid mg;
      mg.nx = d.nx;
      mg.ny = d.ny;
      mg.nz = d.nz;
      mg.h = d.h;
      if ((rc = block_material(d, mg, rho, mu, la)) || (rc = e.set_material(rho.data(), mu.data(), la.data(), false)))
         return rc;
   }
   for (const DeckSource& s : d.sources) {
      if (s.x < 0 || s.x > d.x || s.y < 0 || s.y > d.y || s.z < 0 || s.z > d.z)
         return set_error(E_INVAL, "deck: source positioned outside the grid");
      MomentSource m;
      m.x = s.x;
      m.y = s.y;
      m.z = s.z;
      for (int q = 0; q < 6; q++) m.m[q] = s.m[q];
      m.t0 = s.t0;
      m.freq = s.freq;
      m.tfkind = s.tfkind;
      if ((rc = e.add_moment_source(m))) return rc;
   }
   std::vector<Station> st = d.stations;
   int utc[7];
   utc_now(utc);
   for (Station& s : st) {
      if (s.x < 0 || s.x > d.x || s.y < 0 || s.y > d.y || s.z < 0)
         return set_error(E_INVAL, "deck: receiver " + s.file + " outside the grid");
      if ((rc = e.add_receiver(s.x, s.y, s.z, s.mode, nullptr))) return rc;
      const int* r = e.m_rec.data() + e.m_rec.size() - 4;
      s.gx = (r[0] - 1) * d.h;
      s.gy = (r[1] - 1) * d.h;
      s.gz = (r[2] - 1) * d.h;
      std::memcpy(s.utc, utc, sizeof(utc));
   }
   if ((rc = e.setup(d.tmax, d.steps >= 0 ? d.steps : 0))) return rc;
   if ((rc = make_dirs(path))) return rc;
   const long n3 = 3 * e.mU.npts();
   std::vector<double> hu, hum;
   if (!d.restart.empty()) {
      // CheckPoint::read_checkpoint, then the boundary conditions of both
      // time levels (EW.C:2403-2414)
      hu.assign(n3, 0.0);
      hum.assign(n3, 0.0);
      double t = 0;
      int cycle = 0;
      if ((rc = read_checkpoint(d.restart, d.nx, d.ny, d.nz, &t, &cycle, hum.data(), hu.data()))) return rc;
      if ((rc = e.restart_from(hu.data(), hum.data(), t, cycle))) return rc;
   }
   std::vector<CkptState> cks;
   for (const DeckCheckPoint& p : d.checkpoints) cks.push_back(CkptState{p});
   int nck = 0;
   const int total = e.mNumberOfTimeSteps;
   while (e.mCycle < total) {
      // step to the next cycle at which a checkpoint is due, or to the end
      int next = total;
      if (!cks.empty()) next = e.mCycle + 1;
      const int n = next - e.mCycle;
      if (mode == 0) rc = e.step_fused(n);
      else if (mode == 1) rc = e.step_reference(n);
      else rc = e.step_comp(n);
      if (rc) return rc;
      for (CkptState& c : cks)
         if (c.due(e.mT, e.mCycle, e.mDt)) {
            hu.resize(n3);
            hum.resize(n3);
            if ((rc = e.host_solution(hu.data(), hum.data()))) return rc;
            const std::string prefix = (path != "." && path != "./" ? path : std::string()) + c.p.file;
            if ((rc = write_checkpoint(checkpoint_name(prefix, e.mCycle, total), e.mT, e.mCycle, d.nx, d.ny, d.nz,
                                       hum.data(), hu.data())))
               return rc;
            nck++;
         }
   }
   if ((rc = e.comp_end())) return rc;
   out = DeckResult{};
   out.nx = d.nx;
   out.ny = d.ny;
   out.nz = d.nz;
   out.h = d.h;
   out.dt = e.mDt;
   out.t = e.mT;
   out.nsteps = total;
   out.cycle = e.mCycle;
   out.nstations = (int)st.size();
   out.ncheckpoints = nck;
   out.point_source_test = d.point_source_test;
   if (d.point_source_test) {
      if ((rc = e.point_source_error(out.linf, out.l2, out.solinf))) return rc;
      // EW::timesteploop end (EW.C:2884-2903)
      std::ofstream es((path + "PointSourceErr.txt").c_str());
      es.precision(12);
      es << out.t << " " << out.linf << " " << out.l2 << " " << out.solinf << std::endl;
      if (!es) return set_error(E_INVAL, "cannot write " + path + "PointSourceErr.txt");
   }
   if (!st.empty()) {
      std::vector<int> idx(st.size());
      for (size_t r = 0; r < st.size(); r++) idx[r] = (int)r;
      if ((rc = write_recorded(path, e, st, idx))) return rc;
   }
   return OK;
}

}  // namespace sw4amd
