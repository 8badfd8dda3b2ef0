// batch.h -- several shots (EW objects, grids of any size) advanced together:
// the interior sweeps of all shots run as one launch each (the batched sweep
// kernels of kernels_fast.hip), the per-grid boundary work follows on side
// streams forked from the batch's stream, and the time loop can be captured
// once as a hipGraph and replayed.
#pragma once
#include <vector>

#include "ew.h"
#include "kernels_fast.h"

namespace sw4amd {

class Batch {
public:
   explicit Batch(const std::vector<EW*>& shots) : m_shots(shots) {}
   ~Batch();
   Batch(const Batch&) = delete;
   Batch& operator=(const Batch&) = delete;

   int init();
   int set_kchunk(int kc);
   // Time one step per candidate k-block (planes per workgroup of the
   // sweeps) and keep the fastest; the shots are reset to t = 0 afterwards.
   int tune(int* best, double* ms_per_candidate, int ncand, const int* cand);
   // Advance every shot by n steps; graph: replay a captured three-step
   // period (the buffer rotation) where it fits.
   int step(int n, bool graph);
   hipStream_t stream() const { return m_s; }
   // Gradients of every shot's receiver misfit (EW::gradient of each shot,
   // gtr[q], grho[q], gmu[q], gla[q] as there) with up to max_concurrent
   // shots in flight at once, each on a stream of its own driven by a host
   // thread of its own, so that the kernels of small grids share the device;
   // a shot starts when its work memory (EW::gradient_plan) fits beside the
   // shots running.  The shots run largest first and take their work
   // memory from one pool (WorkPool), which the smaller shots reuse and
   // which stays allocated for the next call.
   int gradient(int nsteps, int ckpt, const double* const* gtr, double* const* grho, double* const* gmu,
                double* const* gla, bool on_device, int max_concurrent, bool fp32 = false);
   // frees the gradients' pool
   int release_workspace();

private:
   int refresh();  // base pointers and clocks from the shots, tables
   int build_tables();
   int one_step(int phase);
   int sync_clocks(int nsteps);
   int run_period_graph();

   std::vector<EW*> m_shots;
   hipStream_t m_s = nullptr;
   SweepShot* m_dtab = nullptr;     // [phase 0..2][kind 0..2][shot]
   int m_nblk[3][3] = {{0}};        // workgroups of the register-staged launch of each kind
   // predictor and damping: the first m_nasync shots of a table take the
   // asynchronous-load kernels (their own launch of m_nblk_async workgroups)
   int m_nasync[3][3] = {{0}};
   int m_nblk_async[3][3] = {{0}};
   double* m_clk_t = nullptr;       // per shot: time, dt
   double* m_clk_dt = nullptr;
   int* m_clk_s = nullptr;          // per shot: sample count, maximum
   int* m_clk_max = nullptr;
   std::vector<double*> m_base;     // per shot: Um, U, Up2 of phase 0
   int m_phase = 0;
   int m_kchunk = 64;
   hipGraphExec_t m_exec = nullptr;
   std::vector<void*> m_allocs;
   std::vector<DampTile*> m_dtiles;  // per shot: its damping tile list (kernel variant 9), or null
   std::vector<int> m_ndtiles;
   std::vector<double*> m_acc;       // per shot: its acceleration field (kernel variant 8 and later), or null
   std::vector<char> m_useacc;       // [phase][shot]: the sweeps store and stage the acceleration
   std::vector<hipStream_t> m_saved;
   WorkPool m_wpool;  // the shots' gradient work memory
   // the per-grid boundary work of the shots runs on NSIDE streams (default
   // priority) forked from and joined back into m_s, so the small kernels of
   // different grids overlap; the fork/join is captured into the graph too
   static constexpr int NSIDE = 4;
   hipStream_t m_side[NSIDE] = {nullptr, nullptr, nullptr, nullptr};
   hipEvent_t m_fork = nullptr;
   hipEvent_t m_join[NSIDE] = {nullptr, nullptr, nullptr, nullptr};
   int fork();
   int join();
};

}  // namespace sw4amd
