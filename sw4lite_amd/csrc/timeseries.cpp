// timeseries.cpp -- USGS text and binary SAC output of a receiver station
// (TimeSeries::writeFile, TimeSeries.C:416-1060; SAC header of sacsubc.C).
#include "timeseries.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>

#include "sw4hip_internal.h"

namespace sw4amd {

void GeoFrame::geographic(double x, double y, double& lon, double& lat) const
{
   const double deg2rad = M_PI / 180.0, phi = az * deg2rad;
   lat = lat0 + (x * cos(phi) - y * sin(phi)) / mpd;
   if (const_mpl)
      lon = lon0 + (x * sin(phi) + y * cos(phi)) / mpl;
   else
      lon = lon0 + (x * sin(phi) + y * cos(phi)) / (mpd * cos(lat * deg2rad));
}

void utc_now(int utc[7])
{
   const time_t now = time(nullptr);
   struct tm t;
   gmtime_r(&now, &t);
   utc[0] = t.tm_year + 1900;
   utc[1] = t.tm_mon + 1;
   utc[2] = t.tm_mday;
   utc[3] = t.tm_hour;
   utc[4] = t.tm_min;
   utc[5] = t.tm_sec;
   utc[6] = 0;
}

namespace {

// Derived quantities of the TimeSeries constructor (TimeSeries.C:224-246):
// geographic positions and the rotation to east-west / north-south.
struct StationFrame {
   double lon, lat, gp_lon, gp_lat;
   double calpha, salpha, thxnrm, thynrm;
   double x_azimuth;
};

StationFrame frame_of(const Station& st, const GeoFrame& geo)
{
   StationFrame f;
   f.x_azimuth = geo.az;
   geo.geographic(st.x, st.y, f.lon, f.lat);
   geo.geographic(st.gx, st.gy, f.gp_lon, f.gp_lat);
   f.calpha = cos(M_PI * f.x_azimuth / 180.0);
   f.salpha = sin(M_PI * f.x_azimuth / 180.0);
   const double cphi = cos(M_PI * f.lat / 180.0), sphi = sin(M_PI * f.lat / 180.0);
   const double mpd = geo.mpd;
   f.thxnrm = f.salpha + (st.x * f.salpha + st.y * f.calpha) / cphi / mpd * (M_PI / 180.0) * sphi * f.calpha;
   f.thynrm = f.calpha - (st.x * f.salpha + st.y * f.calpha) / cphi / mpd * (M_PI / 180.0) * sphi * f.salpha;
   const double nrm = sqrt(f.thxnrm * f.thxnrm + f.thynrm * f.thynrm);
   f.thxnrm /= nrm;
   f.thynrm /= nrm;
   return f;
}

// ---------------------------------------------------------------------------
// SAC binary file: header of 70 floats, 40 ints and 24 eight-character
// fields, then the samples as floats (the layout sacsubc.C writes).
// ---------------------------------------------------------------------------
struct SacHeader {
   float r[70];
   int i[40];
   char c[24][8];
};
static_assert(sizeof(SacHeader) == 632, "SAC header is 632 bytes");

// indices of the header words this writer sets
enum {
   R_DELTA = 0, R_DEPMIN = 1, R_DEPMAX = 2, R_B = 5, R_E = 6, R_O = 7, R_STLA = 31, R_STLO = 32, R_EVLA = 35,
   R_EVLO = 36, R_EVDP = 38, R_DEPMEN = 56, R_CMPAZ = 57, R_CMPINC = 58
};
enum {
   I_NZYEAR = 0, I_NZJDAY = 1, I_NZHOUR = 2, I_NZMIN = 3, I_NZSEC = 4, I_NZMSEC = 5, I_NVHDR = 6, I_NINF = 7,
   I_NHST = 8, I_NPTS = 9, I_IFTYPE = 15, I_LEVEN = 35, I_LOVROK = 37, I_LCALDA = 38
};
enum { C_KSTNM = 0, C_KCMPNM = 20 };

// a fresh header of an evenly sampled time series (inihdr + newhdr,
// sacsubc.C:671-696): undefined values -12345, version 6
void sac_new_header(SacHeader& h)
{
   for (float& v : h.r) v = -12345.0f;
   for (int& v : h.i) v = -12345;
   h.i[I_NVHDR] = 6;
   h.i[I_NINF] = 0;
   h.i[I_NHST] = 0;
   for (int q = 35; q < 40; q++) h.i[q] = 0;
   for (auto& s : h.c) std::memcpy(s, "-12345  ", 8);
   h.i[I_IFTYPE] = 1;  // ITIME
   h.i[I_LEVEN] = 1;
   h.i[I_LOVROK] = 1;
   h.i[I_LCALDA] = 1;
}

void sac_text(char dst[8], const std::string& s)
{
   for (int j = 0; j < 8; j++) dst[j] = j < (int)s.size() ? s[j] : ' ';
}

int days_in_month(int year, int month)
{
   if (month == 2) return 28 + ((year % 400 == 0) || ((year % 4 == 0) && year % 100 != 0));
   if (month == 4 || month == 6 || month == 9 || month == 11) return 30;
   return 31;
}

// TimeSeries::write_sac_format (TimeSeries.C:845-921)
int write_sac(const std::string& name, const Station& st, const StationFrame& f, const std::vector<float>& y,
              float btime, float dt, const char* var, float cmpinc, float cmpaz)
{
   SacHeader h;
   sac_new_header(h);
   const int npts = (int)y.size();
   float depmax = -1.0e+38f, depmin = 1.0e+38f, depmen = -12345.0f;
   double sum = 0.0;
   for (float v : y) {
      if (v > depmax) depmax = v;
      if (v < depmin) depmin = v;
      sum += v;
   }
   if (npts > 0) depmen = (float)(sum / npts);
   h.r[R_DEPMAX] = depmax;
   h.r[R_DEPMIN] = depmin;
   h.r[R_DEPMEN] = depmen;
   h.i[I_NPTS] = npts;
   h.r[R_DELTA] = dt;
   h.r[R_B] = btime;
   h.r[R_E] = btime + (npts - 1) * dt;
   h.i[I_LEVEN] = 1;
   h.i[I_LOVROK] = 1;
   h.i[I_LCALDA] = 1;
   int days = 0;
   for (int m = 1; m < st.utc[1]; m++) days += days_in_month(st.utc[0], m);
   days += st.utc[2];
   h.i[I_NZYEAR] = st.utc[0];
   h.i[I_NZJDAY] = days;
   h.i[I_NZHOUR] = st.utc[3];
   h.i[I_NZMIN] = st.utc[4];
   h.i[I_NZSEC] = st.utc[5];
   h.i[I_NZMSEC] = st.utc[6];
   sac_text(h.c[C_KCMPNM], var);
   h.r[R_STLA] = (float)f.lat;
   h.r[R_STLO] = (float)f.lon;
   // epicentre and its time offset: not known to the writer (TimeSeries.C:437)
   h.r[R_EVLA] = 0.0f;
   h.r[R_EVLO] = 0.0f;
   h.r[R_EVDP] = 0.0f;
   h.r[R_O] = 0.0f;
   h.r[R_CMPINC] = cmpinc;
   h.r[R_CMPAZ] = cmpaz;
   sac_text(h.c[C_KSTNM], st.sta);
   FILE* fp = fopen(name.c_str(), "wb");
   if (!fp) return set_error(E_INVAL, "cannot open SAC file " + name + " for writing");
   const bool ok = fwrite(&h, sizeof(h), 1, fp) == 1 &&
                   (npts == 0 || fwrite(y.data(), sizeof(float), (size_t)npts, fp) == (size_t)npts);
   if (fclose(fp) != 0 || !ok) return set_error(E_INVAL, "error writing SAC file " + name);
   return OK;
}

// TimeSeries::write_usgs_format (TimeSeries.C:924-1060), displacement and
// velocity
int write_usgs(const std::string& name, const Station& st, const StationFrame& f, double t0, double dt,
               int n, const double* u)
{
   FILE* fd = fopen(name.c_str(), "w");
   if (!fd) return set_error(E_INVAL, "cannot open USGS file " + name + " for writing");
   fprintf(fd, "# Author: SW4\n");
   fprintf(fd, "# Scenario: %s\n", "test");
   fprintf(fd, "# Date: UTC  %02i/%02i/%i:%i:%i:%02i.%.3i\n", st.utc[1], st.utc[2], st.utc[0], st.utc[3], st.utc[4],
           st.utc[5], st.utc[6]);
   fprintf(fd, "# Bandwith (Hz): %e\n", 1.234);
   fprintf(fd, "# Station: %s\n", st.file.c_str());
   fprintf(fd, "# Target location (WGS84 longitude, latitude) (deg): %e %e\n", f.lon, f.lat);
   fprintf(fd, "# Actual location (WGS84 longitude, latitude) (deg): %e %e\n", f.gp_lon, f.gp_lat);
   fprintf(fd, "# Distance from target to actual location (m): %e\n",
           sqrt((st.x - st.gx) * (st.x - st.gx) + (st.y - st.gy) * (st.y - st.gy)));
   fprintf(fd, "# nColumns: %i\n", 4);
   fprintf(fd, "# Column 1: Time (s)\n");
   if (st.mode == 0 && st.xyz) {
      fprintf(fd, "# Column 2: X displacement (m)\n");
      fprintf(fd, "# Column 3: Y displacement (m)\n");
      fprintf(fd, "# Column 4: Z displacement (m)\n");
   } else if (st.mode == 0) {
      fprintf(fd, "# Column 2: East-west displacement (m)\n");
      fprintf(fd, "# Column 3: North-sourth displacement (m)\n");
      fprintf(fd, "# Column 4: Up-down displacement (m)\n");
   } else if (st.xyz) {
      fprintf(fd, "# Column 2: X velocity (m/s)\n");
      fprintf(fd, "# Column 3: Y velocity (m/s)\n");
      fprintf(fd, "# Column 4: Z velocity (m/s)\n");
   } else {
      fprintf(fd, "# Column 2: East-west velocity (m/s)\n");
      fprintf(fd, "# Column 3: Nort-south velocity (m/s)\n");
      fprintf(fd, "# Column 4: Up-down velocity (m/s)\n");
   }
   for (int i = 0; i < n; i++) {
      const double* s = u + 3 * (long)i;
      fprintf(fd, "%e", t0 + i * dt);
      if (st.xyz) {
         for (int q = 0; q < 3; q++) fprintf(fd, " %24.17e", s[q]);
      } else {
         const double uns = f.thynrm * s[0] - f.thxnrm * s[1];
         const double uew = f.salpha * s[0] + f.calpha * s[1];
         fprintf(fd, " %24.17e", uew);
         fprintf(fd, " %24.17e", uns);
         fprintf(fd, " %24.17e", -s[2]);
      }
      fprintf(fd, "\n");
   }
   if (fclose(fd) != 0) return set_error(E_INVAL, "error writing USGS file " + name);
   return OK;
}

}  // namespace

int write_station(const std::string& path, const Station& st, const GeoFrame& geo, double t0, double dt,
                  int nsamples, const double* samples)
{
   if (nsamples < 0 || (nsamples > 0 && !samples)) return set_error(E_INVAL, "write_station: no samples");
   if (st.mode != 0 && st.mode != 1) return set_error(E_INVAL, "write_station: mode must be 0 or 1");
   const StationFrame f = frame_of(st, geo);
   // file prefix: the path unless it is ".", then the file name
   const std::string prefix = (path != "." ? path : std::string()) + st.file + ".";
   int rc = OK;
   if (st.sac) {
      // the recorded floats (TimeSeries::recordData keeps (float) u)
      std::vector<float> c[3];
      for (int q = 0; q < 3; q++) {
         c[q].resize(nsamples);
         for (int i = 0; i < nsamples; i++) c[q][i] = (float)samples[3 * (long)i + q];
      }
      const float b = (float)t0, d = (float)dt;
      const bool vel = st.mode == 1;
      if (st.xyz) {
         const float azx = (float)f.x_azimuth, azy = (float)(f.x_azimuth + 90.0);
         const char* ext[3] = {vel ? "xv" : "x", vel ? "yv" : "y", vel ? "zv" : "z"};
         const char* var[3] = {vel ? "Vx" : "X", vel ? "Vy" : "Y", vel ? "Vz" : "Z"};
         if ((rc = write_sac(prefix + ext[0], st, f, c[0], b, d, var[0], 90.0f, azx)) ||
             (rc = write_sac(prefix + ext[1], st, f, c[1], b, d, var[1], 90.0f, azy)) ||
             (rc = write_sac(prefix + ext[2], st, f, c[2], b, d, var[2], 180.0f, 0.0f)))
            return rc;
      } else {
         // east-west, north-south, up from the recorded floats (TimeSeries.C:699-723)
         std::vector<float> g[3];
         for (int q = 0; q < 3; q++) g[q].resize(nsamples);
         for (int i = 0; i < nsamples; i++) {
            g[1][i] = (float)(f.thynrm * c[0][i] - f.thxnrm * c[1][i]);
            g[0][i] = (float)(f.salpha * c[0][i] + f.calpha * c[1][i]);
            g[2][i] = -c[2][i];
         }
         const char* ext[3] = {vel ? "ev" : "e", vel ? "nv" : "n", vel ? "uv" : "u"};
         const char* var[3] = {vel ? "Vew" : "EW", vel ? "Vns" : "NS", vel ? "Vup" : "UP"};
         if ((rc = write_sac(prefix + ext[0], st, f, g[0], b, d, var[0], 90.0f, 90.0f)) ||
             (rc = write_sac(prefix + ext[1], st, f, g[1], b, d, var[1], 90.0f, 0.0f)) ||
             (rc = write_sac(prefix + ext[2], st, f, g[2], b, d, var[2], 0.0f, 0.0f)))
            return rc;
      }
   }
   if (st.usgs) rc = write_usgs(prefix + "txt", st, f, t0, dt, nsamples, samples);
   return rc;
}

}  // namespace sw4amd
