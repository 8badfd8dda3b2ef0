// kernels_fast.h -- LDS-staged sweeps (kernels_fast.hip).
#pragma once
#include <vector>

#include "kernels_common.h"

namespace sw4amd {

struct TileMap {
   int ntx, nty, ntiles;  // tiles along i and j, tiles in all
};

// One grid of a batched sweep: the arguments of its single-grid launch and
// its first block id in the batched launch.
struct DampTile;
struct SweepShot {
   Box b;
   TileMap map;
   const DampTile* tiles;  // damping: the tile list (map.ntiles entries), or null for the whole map
   int kbeg, kend, kc, grid0;
   const double* upin;
   const double* u;
   const double* um;
   const double* mu;
   const double* la;
   const double* rho;
   SgCoef sg;
   double* out;
   double cof, dt2i, cupd, beta;
   // asynchronous-load predictor: where to store the acceleration (or null);
   // asynchronous-load corrector: the stored acceleration it stages
   double* acc;
};
// Fill sh for planes k1..k2 of one grid (damping: the 32 x 8 damping sweep,
// out = the array it updates, over the tiles of damp_tiles when tiles is
// given; else the predictor (corr false) or corrector sweep of variant 6);
// returns the next free block id.
int fill_sweep_shot(bool damping, const Box& b, int k1, int k2, int kc, const double* upin, const double* u,
                    const double* um, const double* mu, const double* la, const double* rho, const SgCoef& sg,
                    double* out, double h, double dt, double beta, bool corr, int grid0, SweepShot* sh,
                    const DampTile* tiles = nullptr, int ntiles = 0, double* acc = nullptr);
// kind 0 predictor, 1 corrector, 2 damping sweeps of nshots grids in one
// launch of nblocks workgroups (d_shots in device memory).
int launch_sweep_batch(int kind, const SweepShot* d_shots, int nshots, int nblocks, hipStream_t s);
// The same with the asynchronous-load bodies, for grids where
// async_loads_fit (even row length, 16-byte aligned arrays); kind 1 is the
// corrector staging the acceleration its predictor stored (SweepShot::acc).
int launch_sweep_batch_async(int kind, const SweepShot* d_shots, int nshots, int nblocks, hipStream_t s);
bool async_loads_fit(const Box& b, const double* const* arrays, int n);

// Interior planes k1..k2 of the predictor (corr = false: out = Up from u, um)
// or corrector sweep (corr = true: out = Up' from upin = Up, u, um), chunks of
// kc planes per workgroup.
// variant 1: plain k-march, 2 waves/SIMD; 2: register prefetch of the next
// plane, 2 waves/SIMD; 3: prefetch, 1 wave/SIMD; 4: as 3 with the corrector's
// supergrid damping (beta != 0) applied in the same sweep; 5: corrector as 4,
// predictor as 1; 6: both as 1, the damping by launch_damp_lds; 7: as 6, the
// predictor by asynchronous global-to-LDS loads where the rows allow it; 8:
// as 7, with acc the acceleration stored and staged (below); 9: as 8, the
// damping sweep restricted to the supergrid layers (damp_tiles).
// The six SBP rows next to side 4 (top) or 5 (bottom) of one sweep from an
// LDS slab (corr: with the acceleration and the supergrid damping).
// The corrector's damping on the six SBP rows of a side (4 top, 5 bottom),
// for rows k_sbp_rows wrote without it (launch_sbp_rows with beta 0); the
// same bits as the rows kernel with its damping.
int launch_rows_damp(const Box& b, int side, int nk, double* up, const double* u, const double* um,
                     const double* rho, const SgCoef& sg, double beta, hipStream_t s);
int launch_sbp_rows(bool corr, const Box& b, int side, int nk, const SbpTables* d_sbp, const double* upin,
                    const double* u, const double* um, const double* mu, const double* la, const double* rho,
                    const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s,
                    const double* acc = nullptr);
// Column tile (tix, tjy) of the 32 x 8 sweeps and its planes k0..k1.
struct DampTile {
   int tix, tjy, k0, k1;
};
// The tiles of the damping sweep over planes k1..k2 (chunks of kc planes)
// whose points can carry a nonzero damping term, from the host damping
// profiles dcx, dcy, dcz (box-sized, from the box start): addsgd4 at a point
// is a sum of terms each with a factor dc at the point or a neighbour, and dc
// is exactly zero outside the supergrid layers, where the update subtracts
// zero and leaves every bit of up.  Tiles whose columns meet an x or y layer
// keep all their planes, the others only the planes of the z layers.
std::vector<DampTile> damp_tiles(const Box& b, int k1, int k2, int kc, const double* dcx, const double* dcy,
                                 const double* dcz);
// up -= supergrid damping(u, um) on planes k1..k2, k-marching LDS sweep;
// with tiles (device array of damp_tiles for the same k1, k2, kc) only those
// tiles run (asynchronous-load sweep; otherwise the whole box).
int launch_damp_lds(const Box& b, int k1, int k2, int kc, double* up, const double* u, const double* um,
                    const double* rho, const SgCoef& sg, double beta, hipStream_t s,
                    const DampTile* tiles = nullptr, int ntiles = 0);
// Variant 8 with acc (3 npts doubles, even row length, aligned arrays): the
// predictor also stores the acceleration (up - 2u + um)/dt^2 at the points
// it computes, the corrector stages acc instead of forming it; between the
// two, launch_accel_fix and launch_accel_points complete acc from the final
// Up.  Without acc, variant 8 is variant 7.
int launch_sweep_lds(bool corr, int variant, const Box& b, int k1, int k2, int kc, const double* upin,
                     const double* u, const double* um, const double* mu, const double* la, const double* rho,
                     const SgCoef& sg, double* out, double h, double dt, double beta, hipStream_t s,
                     double* acc = nullptr);
// acc = (up - 2u + um)/dt^2 on the planes k1-2, k1-1, k2+1, k2+2 (all_planes:
// every plane of the box outside k1..k2) and (frame) the two-point frame of
// columns of planes k1..k2: where the sweep of planes k1..k2 does not form it.  The predictor's ghost shell stores the frame's
// values; they stay final, and frame may be false, when the boundary
// conditions change no ghost value of the four sides (all zero-filled).
int launch_accel_fix(const Box& b, int k1, int k2, bool frame, bool all_planes, const double* up, const double* u,
                     const double* um, double dt, double* acc, hipStream_t s);
// out = cof L(w)/S (S = strx stry strz) on planes k1..k2 of the box's
// interior columns: the predictor sweep without the time update, for the
// transposed operator.
int launch_sweep_lds_T(const Box& b, int k1, int k2, int kc, const double* w, const double* mu, const double* la,
                       const SgCoef& sg, double* out, double h, hipStream_t s);

}  // namespace sw4amd
