// kernels_fast.h -- LDS-staged sweeps (kernels_fast.hip).
#pragma once
#include "kernels_common.h"

namespace sw4amd {

// Interior planes k1..k2 of the predictor (corr = false: out = Up from u, um)
// or corrector sweep (corr = true: out = Up' from upin = Up, u, um), chunks of
// kc planes per workgroup.
// variant 1: plain k-march, 2 waves/SIMD; 2: register prefetch of the next
// plane, 2 waves/SIMD; 3: prefetch, 1 wave/SIMD; 4: as 3 with the corrector's
// supergrid damping (beta != 0) applied in the same sweep; 5: corrector as 4,
// predictor as 1; 6: both as 1, the damping by launch_damp_lds.
// The six SBP rows next to side 4 (top) or 5 (bottom) of one sweep from an
// LDS slab (corr: with the acceleration and the supergrid damping).
int launch_sbp_rows(bool corr, const Box& b, int side, int nk, const SbpTables* d_sbp, const double* upin,
                    const double* u, const double* um, const double* mu, const double* la, const double* rho,
                    const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s);
// up -= supergrid damping(u, um) on planes k1..k2, k-marching LDS sweep.
int launch_damp_lds(const Box& b, int k1, int k2, int kc, double* up, const double* u, const double* um,
                    const double* rho, const SgCoef& sg, double beta, hipStream_t s);
int launch_sweep_lds(bool corr, int variant, const Box& b, int k1, int k2, int kc, const double* upin,
                     const double* u, const double* um, const double* mu, const double* la, const double* rho,
                     const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s);

}  // namespace sw4amd
