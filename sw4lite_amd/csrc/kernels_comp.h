// kernels_comp.h -- the FP32 form of the fused step with a compensated time
// update (kernels_comp.hip).
//
// State, component-blocked float arrays over the whole box:
//   U, C  the displacement u^n = U + C (C is the low part, |C| <= ulp(U)/2),
//   D     u^n - u^{n-1},
//   A     the predicted acceleration (Up - 2U + Um)/dt^2,
//   Dn    u^{n+1} - u^n (written by the corrector; D of the next step).
// The predictor writes A = L(U)/rho (+ F/rho), so Up = u + D + dt^2 A; the
// corrector forms Dn = D + dt^2 A + dt^4/(12 rho) L(A) - damping(D) and adds
// Dn to (U, C) with an error-free two-sum.  Stencils and damping run in FP32;
// the ghost-point boundary conditions and the source time functions run in
// FP64 on values combined from the float state.
#pragma once
#include <functional>

#include "kernels_common.h"

namespace sw4amd {

// Float copies of the 1-D supergrid arrays (SgCoef).
struct SgCoefF {
   const float *dcx, *dcy, *dcz;
   const float *strx, *stry, *strz;
   const float *cox, *coy, *coz;
};

struct CompState {
   float* U;
   float* C;
   float* D;   // u^n - u^{n-1}
   float* Dn;  // u^{n+1} - u^n
   float* A;
   const float* mu;
   const float* la;
   const float* rho;
};

// double (u, um) -> (U, C, D); (U, C, D) -> double (u, um) (um may be null)
int launch_comp_init(long n3, const double* u, const double* um, float* U, float* C, float* D, hipStream_t s);
int launch_comp_flush(long n3, const float* U, const float* C, const float* D, double* u, double* um,
                      hipStream_t s);
int launch_to_float(long n, const double* src, float* dst, hipStream_t s);

// Predictor: A on every point (interior sweep, SBP rows, zero on the ghost
// shell), then the sources A += F/rho and the ghost-point boundary conditions
// of Up.  after_interior runs after the interior sweep is queued.
// zfill: sides (bit s) whose zero ghost values (Dirichlet or supergrid, no
// data) are set in the shell's pass; launch_comp_bc then skips them.  Valid
// when no side is periodic.
int launch_comp_pred(const Box& b, int nk, const int* onesided, const SbpTablesT<float>* d_sbpf, const CompState& st,
                     const SgCoefF& sgf, double h, hipStream_t s, const std::function<int()>& after_interior,
                     int kchunk, double dt = 0.0, int zfill = 0, bool async = true);
// Corrector: Dn and the compensated (U, C) += Dn on every point.  async: the
// interior sweeps bring their planes to LDS by asynchronous loads where the
// rows allow it (ni % 4 == 0, aligned arrays); the same bits either way.
int launch_comp_corr(const Box& b, int nk, const int* onesided, const SbpTablesT<float>* d_sbpf, const CompState& st,
                     const SgCoefF& sgf, double beta, double h, double dt, hipStream_t s,
                     const std::function<int()>& after_interior, int kchunk, int zfill = 0, bool async = true);
// Sources: predictor A += F(t)/rho; corrector Dn += coef/rho F_tt(t) and
// (U, C) += the same increment.
int launch_comp_src(bool corr, const Box& b, const SourceTable& st, double t, double coef, const CompState& cs,
                    hipStream_t s);
// Ghost-point boundary conditions (bcfortsg with zero boundary data,
// ew-cfromfort.C:480-747) in FP64 on Up = U + C + D + dt^2 A (predictor) or
// on u^{n+1} = U + C (corrector); mu, la, strx, stry are the FP64 arrays.
int launch_comp_bc(bool corr, const Box& b, const int* wind, int nx, int ny, int nz, const int* bccnd,
                   const CompState& cs, double h, double dt, const SbpTables* d_sbp, const double* mu,
                   const double* la, const double* strx, const double* stry, hipStream_t s, int skip = 0);
// Receivers: displacement U + C of u^{n+1}, or velocity (Dn + D)/(2 dt).
int launch_comp_record(const Box& b, int nrec, const int* rec, const CompState& cs, double dt, double* out,
                       hipStream_t s);

}  // namespace sw4amd
