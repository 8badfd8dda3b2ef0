// kernels_step.h -- launchers of the fused step and helper kernels
// (kernels_step.hip).
#pragma once
#include <functional>

#include "kernels_common.h"
#include "kernels_fast.h"

namespace sw4amd {

struct ExactParams {
   double x0, y0, z0, rho, alpha, beta, time, fr;
   double m[6];  // xx xy xz yy yz zz
   int tfkind;   // 1 Gaussian, 2 C6SmoothBump
   double h, zmin, eps;
};

struct NormWindow {
   int imin, imax, jmin, jmax, kmin, kmax;
   double x0, y0, z0, r2, h, zmin;
};

int launch_pred_fused(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                      const double* u, const double* um, const double* mu, const double* la,
                      const double* rho, const SgCoef& sg, double* up, double h, double dt,
                      hipStream_t s, const std::function<int()>& after_interior, int variant = 1,
                      int kchunk = 32, int zfill = 0, double* accf = nullptr);
// Doubles of the acceleration slab launch_corr_fused needs as scratch (acc).
long accel_slab_size(const Box& b);
int launch_corr_fused(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp,
                      const double* up, const double* u, const double* um, const double* mu,
                      const double* la, const double* rho, const SgCoef& sg, double beta,
                      double* upo, double* acc, double h, double dt, hipStream_t s,
                      const std::function<int()>& after_interior, const std::function<int()>& after_damping,
                      int variant = 1, int kchunk = 32, int zfill = 0, const double* accf = nullptr,
                      const DampTile* dtiles = nullptr, int ndtiles = 0);
// Interior planes of the fused sweeps (below and above the SBP rows).
void fused_planes(const Box& b, int nk, const int* onesided, int* k1, int* k2);
// acc = (up - 2u + um)/dt^2 at the grid points of the sources (after the
// sources have been added to up).
int launch_accel_points(const Box& b, const SourceTable& st, const double* up, const double* u, const double* um,
                        double dt, double* acc, hipStream_t s);
// Rows next to a free surface and the ghost shell of the fused sweeps
// (variant >= 1), after the interior sweep.  zfill (bit s): side s takes
// zero ghost values in the shell's pass (Dirichlet or supergrid side with no
// data, no periodic side); the boundary conditions then skip it.
int launch_pred_tail(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp, const double* u,
                     const double* um, const double* mu, const double* la, const double* rho, const SgCoef& sg,
                     double* up, double h, double dt, hipStream_t s, int zfill = 0, double* accf = nullptr);
int launch_corr_tail(const Box& b, int nk, const int* onesided, const SbpTables* d_sbp, const double* up,
                     const double* u, const double* um, const double* mu, const double* la, const double* rho,
                     const SgCoef& sg, double beta, double* upo, double h, double dt, hipStream_t s,
                     int zfill = 0,
                     const double* accf = nullptr, bool split_damp = false);
// tdev: when not null the time is read from device memory (graph replay)
int launch_src_apply(const Box& b, const SourceTable& st, double t, int deriv2, double coef,
                     const double* rho, double* up, hipStream_t s, const double* tdev = nullptr);
int launch_maxeig(long n, const double* mu, const double* la, const double* rho, double* partial,
                  int npartial, double* out, hipStream_t s);
int launch_exact(const Box& b, const ExactParams& P, double* out, hipStream_t s);
int launch_solerr(const Box& b, const NormWindow& w, const double* uex, const double* u,
                  double* partial, int npartial, double* out, hipStream_t s);
long shell_count(const Box& b);
// sdev: when not null, out is the start of the sample buffer and the sample
// index is read from device memory (samples >= maxs are not recorded)
int launch_record(const Box& b, int nrec, const int* rec, const double* up, const double* um, double dt,
                  double* out, hipStream_t s, const int* sdev = nullptr, int maxs = 0);
// t[q] += dt[q]; smp[q] += 1 while below maxs[q]
int launch_clock_tick(int n, double* t, int* smp, const double* dt, const int* maxs, hipStream_t s);

}  // namespace sw4amd
