// timeseries.h -- receiver station output: the USGS text format and binary
// SAC files of TimeSeries::writeFile (TimeSeries.C:416-1060, sacsubc.C).
// Host code; the samples come from the GPU receivers (EW::record).
#pragma once
#include <string>
#include <vector>

namespace sw4amd {

// Geographic frame of the grid (EW.C:192-197: origin, azimuth of the x axis,
// meters per degree) and the mapping of EW::computeGeographicCoord
// (EW.C:5794-5816).
struct GeoFrame {
   double lat0 = 37.0, lon0 = -118.0, az = 0.0, mpd = 111319.5;
   bool const_mpl = false;  // m_ConstMetersPerLongitude
   double mpl = 0.0;        // meters per longitude when const_mpl
   void geographic(double x, double y, double& lon, double& lat) const;
};

// One station: what the deck's rec line sets (EW::processReceiver,
// EW.C:1600-1850) and where it landed.
struct Station {
   std::string file;        // file= (name prefix of the output files)
   std::string sta;         // sta= (station name; file when not given)
   int mode = 0;            // 0 displacement, 1 velocity
   bool xyz = true;         // false: east-west, north-south, up (nsew=1)
   bool usgs = false, sac = true;
   double x = 0, y = 0, z = 0;     // requested location
   double gx = 0, gy = 0, gz = 0;  // the grid point it records
   int utc[7] = {1970, 1, 1, 0, 0, 0, 0};  // reference time: year, month, day, hour, minute, second, millisecond
};

// Writes the station's files into path (a directory name ending in '/', or
// "." for the working directory) as TimeSeries::writeFile does: SAC files
// <file>.x/.y/.z (displacement) or .xv/.yv/.zv (velocity), .e/.n/.u or
// .ev/.nv/.uv with nsew, and/or <file>.txt in the USGS format.  samples:
// nsamples x 3 values (x, y, z components) at times t0 + i dt.  Returns 0 or
// a status code with the message set.
int write_station(const std::string& path, const Station& st, const GeoFrame& geo, double t0, double dt,
                  int nsamples, const double* samples);

// The current UTC time as the reference stamps its records
// (EW::get_utc: year, month, day, hour, minute, second, millisecond 0).
void utc_now(int utc[7]);

}  // namespace sw4amd
