// ew.h -- C++ host facade with the reference's EW time-step methods, backed
// by device-resident Sarray data on one MI355X.  The method names and their
// argument meaning follow EW.h:93-127 so the reference time loop can be kept
// as is; the bodies launch gfx950 kernels instead of OpenMP loops.
#pragma once
#include <mutex>
#include <vector>

#include "kernels_adj.h"
#include "kernels_common.h"
#include "kernels_comp.h"
#include "kernels_curv.h"
#include "kernels_step.h"
#include "setup.h"

namespace sw4amd {

// Sarray with the reference's index arithmetic (Sarray.h:54-187) whose data
// lives in device memory.  nc = 1 (material) or 3 (displacement).
struct Sarray {
   int m_nc = 0, m_ib = 0, m_ie = -1, m_jb = 0, m_je = -1, m_kb = 0, m_ke = -1;
   bool m_corder = true;
   double* dev = nullptr;  // owned by EW
   long npts() const { return (long)(m_ie - m_ib + 1) * (m_je - m_jb + 1) * (m_ke - m_kb + 1); }
   long size() const { return npts() * m_nc; }
   // offset of (c,i,j,k), c = 1..m_nc as in the reference
   long index(int c, int i, int j, int k) const
   {
      const long ni = m_ie - m_ib + 1, nij = ni * (m_je - m_jb + 1);
      const long p = (i - m_ib) + ni * (j - m_jb) + nij * (k - m_kb);
      return m_corder ? p + npts() * (c - 1) : (long)m_nc * p + (c - 1);
   }
};

struct EWOptions {
   GridSetup grid;
   int corder = 1;
   double beta = 0.02;  // m_supergrid_damping_coefficient (EW.C:174)
   double cfl = 1.3;    // mCFL (EW.C:163)
};

class Batch;
class GridStack;

// Device buffers shared by the gradients of several EW objects run one
// after another or at once (Batch::gradient): a buffer given back is
// reused by the next request it is large enough for, so the shots' work
// memory is allocated once for the largest shots and never freed between
// them.  Thread-safe.
class WorkPool {
public:
   ~WorkPool() { release(); }
   // the smallest free buffer of at least bytes, else a new one (after
   // freeing the idle buffers when the allocation fails; null when it
   // still does)
   void* take(size_t bytes);
   void give(void* p);
   double held() const;  // bytes allocated
   void release();       // frees every buffer; none may be in use

private:
   struct Buf {
      size_t bytes;
      void* p;
      bool used;
   };
   mutable std::mutex m_mtx;
   std::vector<Buf> m_bufs;
};

class EW {
   friend class Batch;
   friend class GridStack;

public:
   explicit EW(const EWOptions& opt);
   ~EW();
   EW(const EW&) = delete;
   EW& operator=(const EW&) = delete;

   int init();  // device allocation; returns status
   void set_stream(hipStream_t s) { m_stream = s; }

   int set_material(const double* rho, const double* mu, const double* la, bool on_device);
   int set_material_const(double rho, double cs, double cp);
   int add_moment_source(const MomentSource& s);
   // Curvilinear grid (EW::generate_grid + setup_metric, EW.C:6164-6300):
   // z = grid depth over the box, increasing with k; before setup.
   int set_topography(const double* z, bool on_device);
   // setupRun + computeDT; dt0 > 0 replaces the grid's own stable step
   // (two stacked grids step with the smaller of their two)
   int setup(double tmax, int nsteps, double dt0 = 0);
   // EW::computeDT (EW.C:5041-5146) for this grid alone: CFL h / sqrt(max
   // (4 mu + lambda) / rho) on a Cartesian grid, CFL / sqrt(max eigenvalue)
   // on a curvilinear one.  Needs the material (and the topography).
   int stable_dt(double* dt);
   // Grid-point sources made outside (curvilinear sources of stacked grids),
   // added to those of the moment sources at setup
   int add_grid_point_source(const GridPointSource& p);

   // --- reference step methods (EW.h:93-127), single Cartesian grid ---
   int Force(double t, Sarray& F, bool tt);
   int evalRHS(Sarray& U, Sarray& Mu, Sarray& Lambda, Sarray& Lu);
   int evalPredictor(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Rho, Sarray& Lu, Sarray& F);
   int evalCorrector(Sarray& Up, Sarray& Rho, Sarray& Lu, Sarray& F);
   int evalDpDmInTime(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Uacc);
   int addSuperGridDamping(Sarray& Up, Sarray& U, Sarray& Um, Sarray& Rho);
   // skip (bit s): side s already holds its zero ghost values (zfill())
   int enforceBC(Sarray& U, Sarray& Mu, Sarray& Lambda, double t, int skip = 0);
   // Sides whose zero ghost values the fused step's shell pass sets
   // (Dirichlet or supergrid; none with a periodic side or on a curvilinear grid)
   int zfill() const;
   void cycleSolutionArrays();

   bool curvilinear() const { return m_curv; }

   // --- time loop ---
   int step_reference(int n);  // stage order of EW::timesteploop (EW.C:2527-2842)
   // two fused sweeps per step; with stage_ms, kStages accumulated stage
   // times (ms): predictor interior, predictor boundary rows + ghost shell,
   // predictor sources + BC, corrector interior, corrector interior damping
   // (LDS variant; 0 otherwise), corrector boundary rows + ghost shell,
   // corrector sources + BC
   static constexpr int kStages = 7;
   int step_fused(int n, double* stage_ms = nullptr);
   // The fused step in FP32 with the compensated time update
   // (kernels_comp.h), same stage times.  The float state is built from U, Um
   // on the first such step; comp_end writes it back to U, Um and every FP64
   // access to the solution calls it first.
   int step_comp(int n, double* stage_ms = nullptr);
   int comp_end();
   bool comp_active() const { return m_comp; }
   int point_source_error(double& linf, double& l2, double& solinf);

   // --- discrete adjoint of the FP64 step (adjoint.cpp) ---
   // Zero initial data at t = 0, receivers restarted with the t = 0 sample.
   int reset();
   // U, Um (component-blocked host arrays over the box, corder 1) of the
   // current state, waiting for the stream
   int host_solution(double* u, double* um);
   // Continue from a checkpoint: U, Um (host arrays as above) at time t after
   // `cycle` steps; the boundary conditions of both time levels are applied
   // as after CheckPoint::read_checkpoint (EW.C:2403-2414) and the receivers
   // restart with this state as their first sample.
   int restart_from(const double* u, const double* um, double t, int cycle);
   // Gradient of a receiver misfit with respect to rho, mu, lambda (whole
   // box): forward nsteps fused steps from zero data with checkpoints every
   // ckpt steps, then the reverse sweep.  gtr: host array
   // [nsteps+1][nreceivers][3], the misfit's derivative with respect to each
   // recorded sample.  Outputs are overwritten (device pointers if on_device).
   // fp32: the forward run and its recomputation in the FP32 form with the
   // compensated update (step_comp), checkpoints of its float state (36
   // bytes a point instead of 48); the reverse steps in FP64 on that
   // trajectory.  The gradient of the FP32 run's misfit.
   // store: keep the predicted field and the acceleration of every step of
   // a recomputed segment (two states a step), so that the reverse steps do
   // not recompute their predictor (FP64 only).  -1: when they fit in free
   // device memory and, for an interval chosen here, cost fewer recomputed
   // steps; 0: never; 1: always (the interval chosen for it).
   int gradient(int nsteps, int ckpt, const double* gtr, double* grho, double* gmu, double* gla, bool on_device,
                bool fp32 = false, int store = -1);
   // checkpoint interval the gradient uses for ckpt (< 1: from free device
   // memory and the workspace already held), the device bytes it needs and
   // whether it keeps the fields above
   int gradient_plan(int nsteps, int ckpt, int* used, double* bytes, bool fp32 = false, int store = -1,
                     int* stored = nullptr);
   int m_adj_store = -1;  // the store argument of the C-ABI gradient calls
   // when set, the gradient's buffers come from this pool (and go back to
   // it) instead of the object's own
   WorkPool* m_shared = nullptr;
   // The gradient's device buffers stay allocated between calls (a pool of
   // exact-size buffers, grown as needed) so that repeated gradients skip
   // allocating hundreds of GB; this frees them.
   int release_workspace();
   double workspace_bytes() const;

   // receivers (TimeSeries, EW.C:5854-5875): mode 0 displacement, 1 velocity
   int add_receiver(double x, double y, double z, int mode, int* index);
   int add_receiver_ijk(int i, int j, int k, int mode, int* index);  // at a given grid point
   int record(const Sarray& Up, const Sarray& Um);  // one sample, after a step
   int get_receivers(int* nsamples, double* out);   // out[sample][rec][3]
   std::vector<int> m_rec;                          // i, j, k, mode per receiver
   int m_nsamples = 0;

   // state access
   Sarray mU, mUm, mUp, mUp2, mRho, mMu, mLambda;
   Sarray mLu, mF, mUacc;  // only for step_reference
   Sarray mMetric, mJ;     // curvilinear grid only
   std::vector<double> m_zhost;
   double* m_dacc = nullptr;  // acceleration on the planes the corrector's SBP rows read
   double* m_accf = nullptr;  // the whole acceleration (kernel variant 8), allocated at its first step
   int accel_field(double** p);
   // kernel variant 9: the damping sweep's tiles (damp_tiles) for planes
   // k1..k2 and the current k-chunk, made and uploaded on first use
   int damp_tile_list(int k1, int k2, const DampTile** tiles, int* n);
   DampTile* m_ddamp = nullptr;
   int m_ndamp = 0, m_damp_k1 = 0, m_damp_k2 = -1, m_damp_kc = 0;
   // kernel variant 9: the tiles of the damping transpose of the adjoint
   // (damp_T_tiles), made and uploaded on first use
   int damp_T_tile_list(const DampTile** tiles, int* n);
   DampTile* m_ddampT = nullptr;
   int m_ndampT = 0;
   double mDt = 0, mT = 0;
   int mNumberOfTimeSteps = 0, mCycle = 0;
   int m_onesided[6] = {0, 0, 0, 0, 0, 0};
   int m_wind[36];
   int m_nbcpts[6];
   Box m_box;
   EWOptions m_opt;
   SupergridArrays m_sg_host;
   std::vector<MomentSource> m_sources;
   std::vector<GridPointSource> m_gps;
   std::vector<GridPointSource> m_gps_extra;
   std::vector<int> m_first;
   hipStream_t m_stream = nullptr;
   int m_variant = 10;  // interior sweep kernels: 0 thread per point, 1-10 LDS-staged k-march
   // planes per workgroup of the LDS sweeps: predictor and corrector, and
   // the damping sweep over the supergrid layers (fastest measured at 512^3
   // on one device: 64 and 32 of 32..256)
   static constexpr int kSweepChunk = 64, kDampChunk = 32;
   int m_kchunk = kSweepChunk;
   int m_kchunk_damp = kDampChunk;

private:
   struct PoolBuf {
      size_t bytes;
      void* p;
      bool used;
   };
   std::vector<PoolBuf> m_pool;  // gradient workspace
   std::vector<void*> m_taken;    // buffers taken from m_shared
   void* pool_take(size_t bytes);
   void pool_return();
   int alloc(Sarray& a, int nc);
   int upload_sources();
   int ensure_reference_arrays();
   SgCoef sg() const;
   SourceTable sources() const;
   // one fused step U, Um -> Up2 (scratch Up, acc: a work field for the stored
   // acceleration of kernel variant 8, or null), no rotation
   int fused_once(double t, double* acc = nullptr);
   int step_vjp(int n, double* u, double* um, double* u1, double* lam_a, double* lam_b, double* lam_c,
                double* grho, double* gmu, double* gla, double* const* tmp, double* plane, const AdjPad& pad,
                double* up_st = nullptr, double* acc_st = nullptr);
   int comp_begin();
   CompState comp_state() const;
   SgCoefF sgf() const;

   double* m_dsg = nullptr;  // 9 supergrid arrays back to back
   SbpTables* m_dsbp = nullptr;
   void* m_dsrc = nullptr;   // packed source table
   double* m_dred = nullptr; // reduction partials
   int* m_drec = nullptr;     // receiver table (device)
   double* m_drecdata = nullptr;
   int m_maxsamples = 0;
   int m_nred = 1024;
   float* m_cbuf = nullptr;  // FP32 form: U, C, D0, D1, A (3 components each), mu, lambda, rho
   float* m_dsgf = nullptr;  // float copies of the supergrid arrays (layout of m_dsg)
   SbpTablesT<float>* m_dsbpf = nullptr;
   int m_cd = 0;             // D0 (0) or D1 (1) holds u^n - u^{n-1}
   bool m_comp = false;      // the float state holds the solution
   bool m_norecord = false;  // steps record no receiver samples (the gradient's recomputation)
   bool m_setup_done = false, m_material_set = false, m_curv = false;
   std::vector<void*> m_allocs;
};

}  // namespace sw4amd
