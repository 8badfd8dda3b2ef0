/*
 * sw4hip.h -- C ABI of the MI355X (gfx950) implementation of the sw4lite
 * elastic-wave time-step hot path.
 *
 * Two levels:
 *  1. Kernel-level entry points with the argument lists of the reference
 *     kernels (index bounds, raw arrays, coefficient tables), operating on
 *     DEVICE arrays in the Sarray layout, stream-ordered and asynchronous.
 *  2. An EW-level object that owns the device state of one Cartesian grid and
 *     runs the reference time-step loop (EW::timesteploop, EW.C:2339-2928),
 *     either stage by stage (the reference order) or with the fused sweeps.
 *
 * Conventions: every function returns SW4HIP_OK (0) or an error code; the
 * message of the last error of the calling thread is sw4hip_last_error().
 * Index boxes include the two ghost points (ifirst = 1-2 on a single grid).
 * corder = 1 selects the component-blocked layout u[i + ni j + nij k +
 * npts c] (Sarray m_corder, Sarray.C:753-778), corder = 0 the interleaved one
 * u[c + 3 (i + ni j + nij k)].  `stream` is a hipStream_t (NULL = default).
 * Host pointers are caller-owned and only read during the call.
 */
#ifndef SW4HIP_H
#define SW4HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
   SW4HIP_OK = 0,
   SW4HIP_EINVAL = 1, /* bad argument */
   SW4HIP_EHIP = 2,   /* HIP runtime error */
   SW4HIP_ENOMEM = 3, /* device allocation failed */
   SW4HIP_ENODEV = 4, /* no usable GPU */
   SW4HIP_ERANGE = 5, /* size beyond a stated limit */
   SW4HIP_ESTATE = 6  /* call out of order (e.g. step before setup) */
};

/* boundaryConditionType of the reference (src/double/sw4.h:37-38) */
enum { SW4HIP_BC_FREE = 0, SW4HIP_BC_DIRICHLET = 1, SW4HIP_BC_SUPERGRID = 2, SW4HIP_BC_PERIODIC = 3 };
/* time functions (timeDep subset, src/double/sw4.h:40-41) */
enum { SW4HIP_TF_RICKER = 0, SW4HIP_TF_GAUSSIAN = 1, SW4HIP_TF_C6SMOOTHBUMP = 2 };

const char* sw4hip_last_error(void);
int sw4hip_version(void);

/* ----------------------------------------------------------------------- */
/* Host-side set-up                                                          */
/* ----------------------------------------------------------------------- */

/* Replaces EW::GetStencilCoefficients (ew-cfromfort.C:1487-1944): acof[384],
 * ghcof[6], bope[48], sbop[5] in the reference's layout. */
int sw4hip_get_stencil_coefficients(double* acof, double* ghcof, double* bope, double* sbop);

/* One Cartesian grid: nx,ny,nz interior points, arrays over [-1,n+2]. */
typedef struct {
   int nx, ny, nz;
   double h;
   double zmin;         /* z of k=1 (0 for a single grid) */
   int corder;          /* Sarray layout, see above */
   int bc[6];           /* SW4HIP_BC_* per side x-,x+,y-,y+,z-,z+ */
   int sg_gp_thickness; /* supergrid width in points (reference default 20) */
   double sg_damping;   /* supergrid damping coefficient (default 0.02) */
   double cfl;          /* CFL number (default 1.3) */
} sw4hip_grid_desc;

/* Replaces EW::setup_supergrid + assign_supergrid_damping_arrays
 * (EW.C:4636-4801, SuperGrid.C:54-194): 1-D arrays of length n+4 per axis. */
int sw4hip_host_supergrid(const sw4hip_grid_desc* g, double* dcx, double* dcy, double* dcz,
                          double* strx, double* stry, double* strz, double* cox, double* coy,
                          double* coz);

/* Replaces EW::setup_boundary_arrays / side_plane (EW.C:3347-3443). */
int sw4hip_host_boundary_windows(const sw4hip_grid_desc* g, int wind[36], int npts[6]);

/* Replaces Source::set_grid_point_sources4 for a Cartesian moment-tensor
 * source (Source.C:1074-2026).  m = {xx,xy,xz,yy,yz,zz}.  Writes at most
 * maxn grid-point sources: ijk[3n], force[3n]; *n = number produced. */
int sw4hip_host_moment_source(const sw4hip_grid_desc* g, double x, double y, double z,
                              const double m[6], int maxn, int* n, int* ijk, double* force);

/* The same for a source on a curvilinear grid whose depth z (host array over
 * the box, i fastest, increasing in k) is given: Source.C:1605-1800 with
 * the grid derivatives at the source interpolated from the grid. */
int sw4hip_host_moment_source_curv(const sw4hip_grid_desc* g, const double* zgrid, double x,
                                   double y, double z, const double m[6], int maxn, int* n,
                                   int* ijk, double* force);
/* Topography input=gaussian (EW::processTopography, EW.C:535-725): a
 * Gaussian hill of elevation amp exp(-((x-xc)/lx)^2 - ((y-yc)/ly)^2) over a
 * curvilinear grid from the surface down to z = zmax, stacked on the
 * Cartesian grid that continues to the domain depth zdomain. */
typedef struct {
   double zmax;       /* zmax= */
   int order;         /* order= (2..7; reference default 4) */
   double zeta_break; /* zetabreak= (default 0.95) */
   double amp, xc, yc, lx, ly; /* gaussianAmp, gaussianXc, gaussianYc, gaussianLx, gaussianLy */
} sw4hip_topography;

/* Replaces EW::defineDimensionsZ + EW::generate_grid (EW.C:2011-2060,
 * 6164-6205): the z sizes of both grids and, when zgrid is not NULL, z of
 * the curvilinear grid over its box [-1,nx+2] x [-1,ny+2] x [-1,nz_curv+2]
 * (i fastest). */
int sw4hip_host_topography_grid(const sw4hip_topography* t, int nx, int ny, double h, double zdomain, int* nz_cart,
                                int* nz_curv, double* zgrid);

/* Replaces the receiver placement of TimeSeries::TimeSeries with
 * EW::computeNearestGridPoint and EW::invert_grid_mapping (TimeSeries.C:
 * 112-183, 307-330; EW.C:5149-5226, 5722-5790): grid 0 Cartesian or 1
 * curvilinear and the grid point (i, j, k); grid -1 where the reference
 * ignores the station (above the surface).  topodepth: z is the depth
 * below the surface. */
int sw4hip_host_topography_receiver(const sw4hip_topography* t, int nx, int ny, double h, double zdomain, double x,
                                    double y, double z, int topodepth, int* grid, int ijk[3]);

/* Replaces Source::adjust_zcoord + Source::set_grid_point_sources4 for a
 * moment tensor in the stacked grids (Source.C:517-553, 1074-1960,
 * 2191-2330): gijk[4n] = (grid, i, j, k), force[3n].  jac: the curvilinear
 * grid's Jacobian over its box (sw4hip_metric).  *n = 0 where the
 * reference ignores the source (above the surface). */
int sw4hip_host_topography_source(const sw4hip_topography* t, int nx, int ny, double h, double zdomain,
                                  const double* jac, double x, double y, double z, int topodepth, const double m[6],
                                  int maxn, int* n, int* gijk, double* force);

/* ----------------------------------------------------------------------- */
/* Kernel-level entry points (device arrays)                                */
/* ----------------------------------------------------------------------- */

/* Replaces rhs4sg (rhs4sg.C:38-41, corder 0) and rhs4sg_rev
 * (rhs4sg_rev.C:43-48, corder 1); prototypes EW.C:141-149.  lu = L(u).
 * onesided[6], acof, bope, ghcof are HOST tables (as in the reference). */
int sw4hip_rhs4sg(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                  int nk, const int* onesided, const double* acof, const double* bope,
                  const double* ghcof, double* lu, const double* u, const double* mu,
                  const double* lambda, double h, const double* strx, const double* stry,
                  const double* strz, void* stream);

/* Replaces EW::Force (EW.C:3085-3123) with its GridPointSource::getFxyz /
 * getFxyz_tt (GridPointSource.C:330-586): F = 0 on the box, then each of the
 * nsrc grid-point sources adds force[3s..3s+2] * g(t - t0[s]) (tt != 0: g'')
 * at (ijk[3s], ijk[3s+1], ijk[3s+2]); g the time function tfkind[s] (0
 * Ricker, 1 Gaussian, 2 C6SmoothBump) of frequency freq[s].  The source
 * table is HOST data (ijk, force, freq, t0, tfkind), summed per grid point
 * in the order the reference's m_identsources gives (sorted by point,
 * EW::sort_grid_point_sources, EW.C:6370).  The library keeps a device copy
 * per distinct table (up to 16 per device; calls with new tables beyond
 * that wait for the device and start a new set).  F: device array. */
int sw4hip_force(int corder, int ib, int ie, int jb, int je, int kb, int ke, double t, int tt, int nsrc,
                 const int* ijk, const double* force, const double* freq, const double* t0, const int* tfkind,
                 double* F, void* stream);

/* Replaces EW::predfort (ew-cfromfort.C:70-113): up = 2u - um + dt2/rho (lu + fo). */
int sw4hip_predfort(int corder, int ib, int ie, int jb, int je, int kb, int ke, double* up,
                    const double* u, const double* um, const double* lu, const double* fo,
                    const double* rho, double dt2, void* stream);

/* Replaces EW::corrfort (ew-cfromfort.C:40-67): up += dt4/(12 rho) (lu + fo). */
int sw4hip_corrfort(int corder, int ib, int ie, int jb, int je, int kb, int ke, double* up,
                    const double* lu, const double* fo, const double* rho, double dt4,
                    void* stream);

/* Replaces EW::dpdmtfort (ew-cfromfort.C:116-141): uacc = dt2i (up - 2u + um). */
int sw4hip_dpdmtfort(int ib, int ie, int jb, int je, int kb, int ke, const double* up,
                     const double* u, const double* um, double* uacc, double dt2i, void* stream);

/* Replaces EW::addsgd4fort / addsgd4fort_indrev (ew-cfromfort.C:748-849,
 * :952-1056).  1-D arrays are device arrays.  Points where dcx, dcy and dcz
 * vanish at the point and both its neighbours are left untouched (the
 * reference subtracts zero there, for finite data). */
int sw4hip_addsgd4(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                   double* up, const double* u, const double* um, const double* rho,
                   const double* dcx, const double* dcy, const double* dcz, const double* strx,
                   const double* stry, const double* strz, const double* cox, const double* coy,
                   const double* coz, double beta, void* stream);

/* Transposes of the kernels above (no counterpart in the reference: they
 * are the adjoints of sw4hip_rhs4sg, sw4hip_addsgd4 and sw4hip_bcfortsg, for
 * the component-blocked layout and the index box [-1, n+2] of one grid).
 * sw4hip_rhs4sg_adjoint: out = L^T z on every point, z read on the rows of L
 *   (the interior, with the SBP rows where onesided).
 * sw4hip_rhs4sg_matgrad: gmu += d/dmu (z . L(u)), glambda += d/dlambda (z . L(u)).
 * sw4hip_addsgd4_adjoint: for up -= D(u - um; rho): dbar = -D^T y on every
 *   point, grho += d/drho (-y . D).
 * sw4hip_bcfortsg_adjoint: y <- BC^T y for zero boundary data (sides in
 *   reverse order; Dirichlet/supergrid ghosts dropped, free-surface ghosts
 *   pushed into the points their closure reads), gmu, glambda += the closure's
 *   derivatives evaluated on fwd, the field the BC acted on.  Periodic sides
 *   are rejected.  Synchronises. */
int sw4hip_rhs4sg_adjoint(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                          const int* onesided, const double* acof, const double* bope, const double* ghcof,
                          double* out, const double* z, const double* mu, const double* lambda, double h,
                          const double* strx, const double* stry, const double* strz, void* stream);
int sw4hip_rhs4sg_matgrad(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast, int nk,
                          const int* onesided, const double* acof, const double* bope, const double* ghcof,
                          const double* z, const double* u, const double* mu, const double* lambda, double h,
                          const double* strx, const double* stry, const double* strz, double* gmu,
                          double* glambda, void* stream);
int sw4hip_addsgd4_adjoint(int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                           const double* y, const double* u, const double* um, const double* rho,
                           const double* dcx, const double* dcy, const double* dcz, const double* strx,
                           const double* stry, const double* strz, const double* cox, const double* coy,
                           const double* coz, double beta, double* dbar, double* grho, void* stream);
int sw4hip_bcfortsg_adjoint(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny,
                            int nz, double* y, const double* fwd, double h, const int* bccnd, const double* sbop,
                            const double* mu, const double* lambda, const double* strx, const double* stry,
                            double* gmu, double* glambda, void* stream);

/* Allocation-free variant of sw4hip_bcfortsg_adjoint (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under stream capture or a serving loop.  Same arguments and results; it
 * calls no hipMalloc/hipFree and does not synchronise (stream-ordered on
 * `stream`).  Its scratch (the SBP weights and one plane of 3 (nx+4)(ny+4)
 * doubles) is a fixed device array of the library, one per GPU, held until
 * the process ends: planes of more than SW4HIP_NOALLOC_MAX_PLANE_POINTS
 * points are rejected (SW4HIP_ERANGE), and calls that overlap on one GPU
 * (other streams, other threads) race on it. */
#define SW4HIP_NOALLOC_MAX_PLANE_POINTS (1030L * 1030L)
int sw4hip_bcfortsg_adjoint_noalloc(int ib, int ie, int jb, int je, int kb, int ke, const int* wind, int nx, int ny,
                                    int nz, double* y, const double* fwd, double h, const int* bccnd,
                                    const double* sbop, const double* mu, const double* lambda, const double* strx,
                                    const double* stry, double* gmu, double* glambda, void* stream);

/* Replaces EW::bcfortsg / bcfortsg_indrev (ew-cfromfort.C:205-747).
 * wind[36], bccnd[6], sbop[5] are HOST tables; bforce[6] is a host array of
 * device pointers (NULL entries mean zero forcing).  t, om, ph, cv of the
 * reference are unused by the routine and omitted. */
int sw4hip_bcfortsg(int corder, int ib, int ie, int jb, int je, int kb, int ke, const int* wind,
                    int nx, int ny, int nz, double* u, double h, const int* bccnd,
                    const double* sbop, const double* mu, const double* lambda,
                    const double* const* bforce, const double* strx, const double* stry,
                    void* stream);

/* ----------------------------------------------------------------------- */
/* Curvilinear (topography) grid, kernel level                              */
/* ----------------------------------------------------------------------- */
/* Replaces EW::metric / metric_rev (curvilinear-c.C:53-172, :175-295): the
 * metric met (4 components (sqrt(z_r), -z_p/sqrt(z_r), -z_q/sqrt(z_r),
 * h/sqrt(z_r)), same layout as u) and jac = h^2 z_r from the grid
 * coordinates x, y, z (device arrays over the box; h = x(ib+1,jb,kb) -
 * x(ib,jb,kb) as in the reference, y unused).  SW4HIP_EINVAL where z_r <= 0
 * (the reference returns -1).  Synchronises the stream. */
int sw4hip_metric(int corder, int ib, int ie, int jb, int je, int kb, int ke, const double* x,
                  const double* y, const double* z, double* met, double* jac, void* stream);
/* Replaces rhs4sgcurv (rhs4sgcurv.C:34-39, corder 0) and rhs4sgcurv_rev
 * (rhs4sgcurv_rev.C:34-40, corder 1); called EW.C:3221-3243.  lu = L(u) on
 * the curvilinear grid; SBP rows 1..6 when onesided[4] == 1, interior rows
 * up to klast-2.  onesided, acof, bope, ghcof are HOST tables. */
int sw4hip_rhs4sgcurv(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                      const double* u, const double* mu, const double* lambda, const double* met,
                      const double* jac, double* lu, const int* onesided, const double* acof,
                      const double* bope, const double* ghcof, const double* strx,
                      const double* stry, void* stream);
/* Replaces EW::addsgd4cfort / addsgd4cfort_indrev (ew-cfromfort.C:1160-1240,
 * :1323-1404): x and y supergrid damping with rho -> rho jac. */
int sw4hip_addsgd4c(int corder, int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                    double* up, const double* u, const double* um, const double* rho,
                    const double* dcx, const double* dcy, const double* strx, const double* stry,
                    const double* jac, const double* cox, const double* coy, double beta,
                    void* stream);
/* Replaces EW::freesurfcurvisg / freesurfcurvisg_rev (curvilinear-c.C:465-618,
 * :621-764): traction-free ghost points on side 5 (k = 1, the top) or 6
 * (k = nz), sides numbered from 1 as in the reference.  forcing: device
 * array of 3 interleaved values per (i,j) of the box plane, or NULL for zero
 * traction; sbop is a HOST table. */
int sw4hip_freesurfcurvisg(int corder, int ib, int ie, int jb, int je, int kb, int ke, int nz,
                           int side, double* u, const double* mu, const double* lambda,
                           const double* met, const double* sbop, const double* forcing,
                           const double* strx, const double* stry, void* stream);
/* Replaces EW::enforceCartTopo (EW.C:3503-3531): the interface of a
 * curvilinear grid stacked on the top Cartesian grid, both over the same
 * (i, j) box [ib, ie] x [jb, je] (ghosts included).  The Cartesian ghost
 * planes kb_cart .. kb_cart+g-1 take the curvilinear planes
 * ke_curv-2g .. ke_curv-g-1; then the curvilinear planes ke_curv-g .. ke_curv
 * (the common plane and the ghosts below it) take the Cartesian planes
 * kb_cart+2g .. kb_cart+g, g = ghost_points (2 in the reference). */
int sw4hip_enforce_cart_topo(int corder, int ib, int ie, int jb, int je, int kb_cart, int ke_cart, double* ucart,
                             int kb_curv, int ke_curv, double* ucurv, int ghost_points, void* stream);
/* ----------------------------------------------------------------------- */
/* EW-level object                                                           */
/* ----------------------------------------------------------------------- */
typedef struct sw4hip_ew sw4hip_ew;

typedef struct {
   double dt;
   int nsteps;
   double t;         /* current time */
   int cycle;        /* steps taken */
   int onesided[6];
   int wind[36];
   long npts;        /* points per component incl. ghosts */
   int nsources;     /* grid-point sources */
   int nunique;      /* distinct source points */
} sw4hip_ew_info;

/* The EW constructor + set-up of one grid (EW::EW / setupRun, EW.C:162-224,
 * 2288-2336) on the current HIP device. */
int sw4hip_ew_create(const sw4hip_grid_desc* g, sw4hip_ew** ew);
int sw4hip_ew_destroy(sw4hip_ew* ew);
/* Stream used by every later call on this object (hipStream_t). */
int sw4hip_ew_set_stream(sw4hip_ew* ew, void* stream);
/* Interior kernels of the fused sweeps: variant 0 one thread per point;
 * variants 1-10 LDS-staged k-marching tiles with kchunk planes per workgroup:
 * 1 plain at 2 waves/SIMD, 2 register prefetch of the next
 * plane, 3 prefetch without a register cap, 4 as 3 with the supergrid
 * damping inside the corrector sweep, 5 corrector as 4 and predictor as 1,
 * 6 as 1 with the damping as its own LDS sweep, 7 as 6 with the
 * predictor's planes brought to LDS by asynchronous loads one plane ahead
 * (even row lengths; otherwise as 6), 8 as 7 with the
 * acceleration (Up - 2U + Um)/dt^2 stored by the predictor (and its ghost
 * shell) and brought to LDS by the corrector the same way (one more field of
 * device memory, allocated at the first step; even row lengths, otherwise as
 * 7; the same bits as 7), 9 as 8 with the damping sweep only over the
 * tiles and planes of the supergrid layers (the same bits as 8), 10 (the
 * default) as 9 with rows of odd length staged by four-byte asynchronous
 * loads and the acceleration stored for them too (the same bits as 9).
 * All give the same results to rounding.  kchunk > 0 sets the planes per
 * workgroup of every sweep; kchunk <= 0 keeps the defaults (64 for the
 * predictor and corrector sweeps, 32 for the damping sweep). */
int sw4hip_ew_set_kernel_variant(sw4hip_ew* ew, int variant, int kchunk);
/* rho, mu, lambda over the whole box incl. ghosts; on_device selects whether
 * the pointers are device (1) or host (0) arrays. */
int sw4hip_ew_set_material(sw4hip_ew* ew, const double* rho, const double* mu,
                           const double* lambda, int on_device);
/* testpointsource material: constant rho, cs, cp (EW.C:893-950). */
int sw4hip_ew_set_material_const(sw4hip_ew* ew, double rho, double cs, double cp);
/* A moment-tensor source (EW::processSource, EW.C:951-1313), m already
 * multiplied by m0. */
int sw4hip_ew_add_moment_source(sw4hip_ew* ew, double x, double y, double z, const double m[6],
                                double t0, double freq, int tfkind);
/* Makes the grid curvilinear (EW::generate_grid + setup_metric,
 * EW.C:6164-6300, for a single grid): z is the grid depth over the box,
 * increasing in k (host or device array per on_device).  Computes the metric
 * and Jacobian on the device.  Must precede setup and receivers.  The time
 * step then uses rhs4sgcurv, addsgd4cfort (x and y only) and the curved
 * free-surface closure, in the reference stage order (sw4hip_ew_step mode 1);
 * dt comes from the curvilinear eigenvalue bound (EW.C:5063-5136). */
int sw4hip_ew_set_topography(sw4hip_ew* ew, const double* z, int on_device);
/* Supergrid arrays, boundary windows, time step (EW::computeDT with
 * t = tmax; nsteps > 0 overrides the step count with dt from CFL), source
 * discretisation.  Must follow the material and the sources. */
int sw4hip_ew_setup(sw4hip_ew* ew, double tmax, int nsteps);
int sw4hip_ew_get_info(sw4hip_ew* ew, sw4hip_ew_info* info);
/* Copies of the set-up products (host buffers). */
int sw4hip_ew_get_supergrid(sw4hip_ew* ew, double* dcx, double* dcy, double* dcz, double* strx,
                            double* stry, double* strz, double* cox, double* coy, double* coz);
int sw4hip_ew_get_sources(sw4hip_ew* ew, int* ijk, double* force, double* freq, double* t0,
                          int* tfkind, int* first);
/* U (time t) and Um (t - dt), 3 components each, in the grid's layout. */
int sw4hip_ew_set_solution(sw4hip_ew* ew, const double* u, const double* um, int on_device);
int sw4hip_ew_get_solution(sw4hip_ew* ew, double* u, double* um, int on_device);
/* Advance nsteps steps.  mode 0: fused sweeps (fast path, corder 1 only);
 * mode 1: the reference's stage order (Force, evalRHS, evalPredictor,
 * enforceBC, Force_tt, evalDpDmInTime, evalRHS, evalCorrector,
 * addSuperGridDamping, enforceBC; EW.C:2527-2842);
 * mode 2: the fused sweeps in FP32 with a compensated time update, the
 * same stage order as EW::timesteploop.  The solution is carried as
 * u = U + C (two floats), D = u - um and the predicted acceleration; the
 * ghost-point closures, source time functions and receivers use FP64.  The
 * float state is built from U, Um on the first mode-2 step and written back
 * to U, Um before any FP64 step or solution access.  The reference's float
 * build (src/float/sw4.h:35) is the uncompensated counterpart. */
int sw4hip_ew_step(sw4hip_ew* ew, int nsteps, int mode);
/* The fused step with per-stage GPU times: hipEvents recorded on the
 * object's stream; stage_ms[7] receives the summed milliseconds of
 * (predictor interior kernel, predictor boundary rows + ghost shell,
 * predictor sources + ghost-point BC, corrector interior kernel, corrector
 * interior supergrid damping, corrector boundary rows + ghost shell,
 * corrector sources + BC).  Synchronises. */
int sw4hip_ew_step_timed(sw4hip_ew* ew, int nsteps, double* stage_ms);
/* Mode 2 (FP32, compensated) with the same per-stage times; the supergrid
 * damping is part of the corrector sweep, so stage 5 reads 0. */
int sw4hip_ew_step_comp_timed(sw4hip_ew* ew, int nsteps, double* stage_ms);
/* Analytic point-source error of the current U (EW::exactSol +
 * normOfDifference, EW.C:3650-3669, 4526-4610). */
int sw4hip_ew_point_source_error(sw4hip_ew* ew, double* linf, double* l2, double* solinf);

/* ---- Discrete adjoint of the FP64 step (no counterpart in the reference,
 * which has no adjoint: these are the exact transposes of the maps of
 * EW::timesteploop, EW.C:2527-2842, for component-blocked Cartesian grids
 * with supergrid, Dirichlet and free-surface sides) ---- */
/* Zero U, Um at t = 0 and restart the receivers with the t = 0 sample. */
int sw4hip_ew_reset(sw4hip_ew* ew);
/* Gradient of a receiver misfit J with respect to rho, mu and lambda on
 * every point of the box (ghost points included).  Runs nsteps fused steps
 * from zero initial data, saving (U, Um) every ckpt steps, then the reverse
 * sweep, recomputing one checkpoint segment at a time (all but the last:
 * nsteps - ckpt forward steps when ckpt divides nsteps).  ckpt <= 0: the
 * largest interval whose states fit in free device memory, nsteps (nothing
 * recomputed) when the whole history fits; sqrt(nsteps) when not even that
 * bound can be met.  grad_traces: HOST array [nsteps+1][nreceivers][3] of dJ/d(sample)
 * in the order of sw4hip_ew_get_receivers.  dt stays the one of setup (set
 * a new material with sw4hip_ew_set_material after setup to keep it).
 * Outputs overwritten; device pointers when on_device. */
int sw4hip_ew_gradient(sw4hip_ew* ew, int nsteps, int ckpt, const double* grad_traces, double* grho,
                       double* gmu, double* glambda, int on_device);
/* The gradient of the FP32 form's misfit: the forward run and its
 * recomputation step in FP32 with the compensated update (sw4hip_ew_step
 * mode 2), checkpointing its float state (U, C, D: 36 bytes a point against
 * 48 for (U, Um) in FP64); the reverse steps run in FP64 on that trajectory.
 * Arguments as sw4hip_ew_gradient.  No counterpart in the reference. */
int sw4hip_ew_gradient_comp(sw4hip_ew* ew, int nsteps, int ckpt, const double* grad_traces, double* grho,
                            double* gmu, double* glambda, int on_device);
/* The checkpoint interval sw4hip_ew_gradient would use for (nsteps, ckpt)
 * and the device bytes it needs (device_bytes may be NULL). */
int sw4hip_ew_gradient_plan(sw4hip_ew* ew, int nsteps, int ckpt, int* used_ckpt, double* device_bytes);
/* Whether sw4hip_ew_gradient keeps, for every step of a recomputed segment,
 * its predicted field and acceleration (two more states a step), so that
 * the reverse steps do not recompute their predictor; the same gradients
 * to the last bit.  store -1 (the default): when they fit in free device
 * memory and, with ckpt <= 0, when that recomputes fewer steps; 0 never;
 * 1 always.  The FP32 form's gradient never keeps them.  No counterpart in
 * the reference. */
int sw4hip_ew_set_gradient_store(sw4hip_ew* ew, int store);
/* sw4hip_ew_gradient keeps its device buffers (states, work fields) between
 * calls and reuses them; this frees them (sw4hip_ew_destroy does too).  No
 * counterpart in the reference. */
int sw4hip_ew_release_workspace(sw4hip_ew* ew);
/* A receiver at the grid point nearest to (x,y,z) (EW::computeNearestGridPoint,
 * EW.C:5149-5180); mode 0 records displacement, 1 velocity (up-um)/(2dt)
 * (EW::extractRecordData, EW.C:5854-5875).  One sample at t=0 and one per
 * step, up to the step count given to setup.  Must precede setup. */
int sw4hip_ew_add_receiver(sw4hip_ew* ew, double x, double y, double z, int mode, int* index);
/* Samples recorded so far: out[sample][receiver][3] (host buffer of
 * nsamples*nreceivers*3 doubles, or NULL to query *nsamples). */
int sw4hip_ew_get_receivers(sw4hip_ew* ew, int* nsamples, double* out);
/* Device pointers of the state: which = 0 U, 1 Um, 2 rho, 3 mu, 4 lambda,
 * 5 metric, 6 Jacobian (curvilinear grid). */
int sw4hip_ew_device_array(sw4hip_ew* ew, int which, double** ptr);

/* ---- Batched shots (no counterpart in the reference, which runs one grid
 * per process): several set-up EW objects of any grid size advanced together
 * with the fused FP64 step.  The three interior sweeps of all shots run as one
 * launch each; the boundary work of every grid follows on the batch's stream;
 * sources and receivers read per-shot device clocks, so a captured period of
 * three steps (the buffer rotation) replays as one hipGraph.  The shots keep
 * their own state: after a batch step each EW reports its new time, samples
 * and solution as if stepped alone (same kernels, same results). ---- */
typedef struct sw4hip_batch sw4hip_batch;
int sw4hip_batch_create(sw4hip_ew* const* shots, int nshots, sw4hip_batch** batch);
int sw4hip_batch_destroy(sw4hip_batch* batch);
/* Planes per workgroup of the batched sweeps (default 64). */
int sw4hip_batch_set_kchunk(sw4hip_batch* batch, int kchunk);
/* Autotuner: time one batch step for each candidate k-block (cand NULL:
 * 16, 24, 32, 40, 48, 64, 96), keep the fastest (*best, per-candidate ms in ms[] when
 * not NULL), then reset every shot to zero data at t = 0. */
int sw4hip_batch_tune(sw4hip_batch* batch, int ncand, const int* cand, int* best, double* ms);
/* Gradients of every shot's receiver misfit with respect to rho, mu,
 * lambda (sw4hip_ew_gradient of each shot: grad_traces[q], grho[q], gmu[q],
 * glambda[q] as there; every shot run from zero data).  Up to
 * max_concurrent shots run at once, each on a stream of its own, so that
 * the kernels of small grids share the device; a shot starts when its work
 * memory fits beside the running ones.  The largest shots run first; the
 * shots take their work memory from one pool of the batch, which the later
 * (smaller) shots reuse and which stays allocated for the next call
 * (sw4hip_batch_release_workspace, sw4hip_batch_destroy free it).
 * ckpt < 1: about sqrt(nsteps) steps between checkpoints. */
int sw4hip_batch_gradient(sw4hip_batch* batch, int nsteps, int ckpt, const double* const* grad_traces,
                          double* const* grho, double* const* gmu, double* const* glambda, int on_device,
                          int max_concurrent);
/* The same with sw4hip_ew_gradient_comp for every shot. */
int sw4hip_batch_gradient_comp(sw4hip_batch* batch, int nsteps, int ckpt, const double* const* grad_traces,
                               double* const* grho, double* const* gmu, double* const* glambda, int on_device,
                               int max_concurrent);
/* Frees the work memory sw4hip_batch_gradient keeps.  No counterpart in
 * the reference. */
int sw4hip_batch_release_workspace(sw4hip_batch* batch);
/* Advance every shot by nsteps; use_graph: replay the captured period. */
int sw4hip_batch_step(sw4hip_batch* batch, int nsteps, int use_graph);
int sw4hip_ew_sync(sw4hip_ew* ew);

/* ---- Input decks, receiver files, checkpoints ---- */

/* Replaces EW::EW(deck) + EW::timesteploop (EW.C:2171-2336, 2339-2928;
 * main.C): the commands grid, time, fileio, supergrid, developer,
 * testpointsource, source (moment tensors; Ricker, Gaussian, C6SmoothBump),
 * block, rec, checkpoint, restart and topography (input=gaussian: a
 * curvilinear grid under the hill stacked on the Cartesian grid below zmax,
 * coupled by EW::enforceCartTopo; no checkpoints or point-source test with
 * it).  Anything else (point forces, other time functions, other topography
 * inputs) is refused with SW4HIP_EINVAL and a message naming the line.  Runs
 * on the current GPU: mode 0 fused sweeps, 1 the reference stage order, 2
 * FP32 with the compensated update; decks with topography step in the
 * reference stage order in modes 0 and 1.  Writes the receiver files,
 * checkpoints and (point source test) PointSourceErr.txt into the deck's
 * fileio path, or into `path` when it is not NULL. */
typedef struct {
   int nx, ny, nz;          /* nz: the Cartesian grid's */
   double h, dt, t;
   int nsteps, cycle, nstations, ncheckpoints, point_source_test;
   double linf, l2, solinf; /* point source test: errors at the end */
   int ngrids, nz_curv;     /* with topography: 2, and the curvilinear grid's nz */
} sw4hip_deck_report;
int sw4hip_run_deck(const char* deck_file, const char* path, int mode, sw4hip_deck_report* report);
/* Parse deck text only (no GPU): grid sizes and step settings in report
 * (nsteps = steps= of the deck, -1 when it gives t=). */
int sw4hip_parse_deck(const char* deck_text, sw4hip_deck_report* report);

/* Replaces TimeSeries::writeFile (TimeSeries.C:416-1060, sacsubc.C): the
 * station's binary SAC files (.x .y .z, velocity .xv .yv .zv, nsew .e .n .u /
 * .ev .nv .uv) and/or its USGS text file (.txt) in path ("." for none).
 * samples: nsamples x 3 (x, y, z) at t0 + i dt.  Geographic frame: the
 * reference's defaults (origin 37 N 118 W, azimuth 0, 111319.5 m/degree). */
typedef struct {
   const char* file; /* file name prefix */
   const char* sta;  /* station name (NULL: file) */
   int mode;         /* 0 displacement, 1 velocity */
   int nsew;         /* 1: east-west, north-south, up components */
   int usgs, sac;    /* formats to write */
   double x, y, z;   /* requested location */
   double gx, gy, gz; /* the grid point recorded */
   int utc[7];       /* reference time: year, month, day, hour, minute, second, ms */
} sw4hip_station;
int sw4hip_write_station(const sw4hip_station* station, const char* path, double t0, double dt, int nsamples,
                         const double* samples);

/* Replaces CheckPoint::write_checkpoint / read_checkpoint
 * (CheckPoint.C:251-480) for one grid: header (precision 8, 1 grid, time,
 * cycle, 1 nx 1 ny 1 nz), then Um and U on the interior points, the three
 * components of a point together, i fastest.  um, u: host arrays over the
 * box [-1, n+2], component-blocked (corder 1); read fills the interior. */
int sw4hip_write_checkpoint(const char* file, double time, int cycle, int nx, int ny, int nz, const double* um,
                            const double* u);
int sw4hip_read_checkpoint(const char* file, int nx, int ny, int nz, double* time, int* cycle, double* um,
                           double* u);

#ifdef __cplusplus
}
#endif
#endif
