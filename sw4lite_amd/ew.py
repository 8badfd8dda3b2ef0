"""Python mirror of the reference ``EW`` time-step interface over the C ABI.

:class:`EW` owns one Cartesian grid on the current GPU (through
``sw4hip_ew_*``) and exposes the reference's driver-level operations:
material and source set-up (EW::setupRun, EW.C:2288-2336), the time-step loop
(EW::timesteploop, EW.C:2339-2928) and the point-source error
(EW::normOfDifference, EW.C:4526-4610).  The kernel-level functions below
(:func:`rhs4sg`, :func:`predfort`, ...) take the reference kernels' argument
lists with device tensors.  Device memory comes from PyTorch; all arithmetic
runs in the HIP library.
"""
import ctypes as C

import numpy as np

from . import lib as L


def _ptr(a):
    """Device or host address of a torch tensor / numpy array."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr())
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data_as(C.c_void_p)
    raise TypeError("expected a torch tensor or numpy array")


def _is_dev(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _fields(n, *arrays):
    """Check arrays handed to an EW method that copies n float64 values from
    or into each: float64, exactly n elements, contiguous, and all of them on
    one side (all device tensors or all host arrays).  Returns whether they
    are on the device.  Raises ValueError otherwise."""
    on_dev = None
    for a in arrays:
        if hasattr(a, "data_ptr"):
            import torch
            if a.dtype != torch.float64:
                raise ValueError("expected a float64 tensor, got %s" % a.dtype)
            if a.numel() != n:
                raise ValueError("expected %d elements, got %d" % (n, a.numel()))
            if not a.is_contiguous():
                raise ValueError("tensor must be contiguous")
            dev = bool(a.is_cuda)
        elif isinstance(a, np.ndarray):
            if a.dtype != np.float64:
                raise ValueError("expected a float64 array, got %s" % a.dtype)
            if a.size != n:
                raise ValueError("expected %d elements, got %d" % (n, a.size))
            if not a.flags["C_CONTIGUOUS"]:
                raise ValueError("array must be C-contiguous")
            dev = False
        else:
            raise TypeError("expected a torch tensor or numpy array")
        if on_dev is None:
            on_dev = dev
        elif on_dev != dev:
            raise ValueError("arrays must all be device tensors or all host arrays")
    if on_dev:
        devs = {a.device for a in arrays}
        if len(devs) != 1:
            raise ValueError("device tensors must be on one device, got %s" % sorted(map(str, devs)))
    return bool(on_dev)


def _stream(stream):
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def stencil_coefficients():
    """acof[384], ghcof[6], bope[48], sbop[5] (EW::GetStencilCoefficients)."""
    lib = L.load()
    acof, ghcof, bope, sbop = np.zeros(384), np.zeros(6), np.zeros(48), np.zeros(5)
    L.check(lib.sw4hip_get_stencil_coefficients(_ptr(acof), _ptr(ghcof), _ptr(bope), _ptr(sbop)))
    return acof, ghcof, bope, sbop


def host_supergrid(desc):
    lib = L.load()
    n = (desc.nx + 4, desc.ny + 4, desc.nz + 4)
    out = {k: np.zeros(n["xyz".index(k[-1])]) for k in
           ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz")}
    L.check(lib.sw4hip_host_supergrid(C.byref(desc), *[_ptr(out[k]) for k in
                                                        ("dcx", "dcy", "dcz", "strx", "stry", "strz",
                                                         "cox", "coy", "coz")]))
    return out


def host_boundary_windows(desc):
    lib = L.load()
    wind = np.zeros(36, dtype=np.int32)
    npts = np.zeros(6, dtype=np.int32)
    L.check(lib.sw4hip_host_boundary_windows(C.byref(desc), _ptr(wind), _ptr(npts)))
    return wind, npts


def host_moment_source(desc, x, y, z, m):
    lib = L.load()
    maxn = 216
    ijk = np.zeros(3 * maxn, dtype=np.int32)
    force = np.zeros(3 * maxn)
    n = C.c_int(0)
    mm = np.asarray(m, dtype=np.float64)
    L.check(lib.sw4hip_host_moment_source(C.byref(desc), x, y, z, _ptr(mm), maxn, C.byref(n),
                                          ijk.ctypes.data_as(L.IP), _ptr(force)))
    return ijk[:3 * n.value].reshape(-1, 3), force[:3 * n.value].reshape(-1, 3)


def host_moment_source_curv(desc, zgrid, x, y, z, m):
    """Grid-point sources of a moment source on the curvilinear grid of depth
    zgrid (host array over the box)."""
    lib = L.load()
    maxn = 216
    ijk = np.zeros(3 * maxn, dtype=np.int32)
    force = np.zeros(3 * maxn)
    n = C.c_int(0)
    mm = np.asarray(m, dtype=np.float64)
    zg = np.ascontiguousarray(zgrid, dtype=np.float64)
    L.check(lib.sw4hip_host_moment_source_curv(C.byref(desc), _ptr(zg), x, y, z, _ptr(mm), maxn, C.byref(n),
                                               ijk.ctypes.data_as(L.IP), _ptr(force)))
    return ijk[:3 * n.value].reshape(-1, 3), force[:3 * n.value].reshape(-1, 3)


# --------------------------------------------------------------------------
# kernel level (device tensors, reference argument lists)
# --------------------------------------------------------------------------
def rhs4sg(corder, box, nk, onesided, acof, bope, ghcof, lu, u, mu, la, h, strx, stry, strz, stream=None):
    lib = L.load()
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    L.check(lib.sw4hip_rhs4sg(corder, *box, nk, _ptr(os_), _ptr(np.ascontiguousarray(acof)),
                              _ptr(np.ascontiguousarray(bope)), _ptr(np.ascontiguousarray(ghcof)),
                              _ptr(lu), _ptr(u), _ptr(mu), _ptr(la), h, _ptr(strx), _ptr(stry), _ptr(strz),
                              _stream(stream)))


def force(corder, box, t, tt, ijk, frc, freq, t0, tfkind, F, stream=None):
    """EW::Force for one grid: F (device) = sum of the grid-point sources
    (host arrays ijk [n,3], frc [n,3], freq, t0, tfkind [n]) at time t (tt:
    second time derivative)."""
    ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1)
    frc = np.ascontiguousarray(frc, dtype=np.float64).reshape(-1)
    n = ijk.size // 3
    fq = np.ascontiguousarray(freq, dtype=np.float64)
    t0_ = np.ascontiguousarray(t0, dtype=np.float64)
    tf = np.ascontiguousarray(tfkind, dtype=np.int32)
    if not (frc.size == 3 * n and fq.size == n and t0_.size == n and tf.size == n):
        raise ValueError("source table arrays disagree in length")
    L.check(L.load().sw4hip_force(corder, *box, float(t), int(bool(tt)), n, _ptr(ijk), _ptr(frc), _ptr(fq), _ptr(t0_),
                                  _ptr(tf), _ptr(F), _stream(stream)))


def predfort(corder, box, up, u, um, lu, fo, rho, dt2, stream=None):
    L.check(L.load().sw4hip_predfort(corder, *box, _ptr(up), _ptr(u), _ptr(um), _ptr(lu), _ptr(fo), _ptr(rho),
                                     dt2, _stream(stream)))


def corrfort(corder, box, up, lu, fo, rho, dt4, stream=None):
    L.check(L.load().sw4hip_corrfort(corder, *box, _ptr(up), _ptr(lu), _ptr(fo), _ptr(rho), dt4, _stream(stream)))


def dpdmtfort(box, up, u, um, uacc, dt2i, stream=None):
    L.check(L.load().sw4hip_dpdmtfort(*box, _ptr(up), _ptr(u), _ptr(um), _ptr(uacc), dt2i, _stream(stream)))


def addsgd4(corder, box, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta, stream=None):
    L.check(L.load().sw4hip_addsgd4(corder, *box, _ptr(up), _ptr(u), _ptr(um), _ptr(rho), _ptr(dcx), _ptr(dcy),
                                    _ptr(dcz), _ptr(strx), _ptr(stry), _ptr(strz), _ptr(cox), _ptr(coy), _ptr(coz),
                                    beta, _stream(stream)))


def metric(corder, box, x, y, z, met, jac, stream=None):
    """EW::metric / metric_rev: met (4 components) and jac from the grid
    coordinates; raises where z is not increasing in k."""
    L.check(L.load().sw4hip_metric(corder, *box, _ptr(x), _ptr(y), _ptr(z), _ptr(met), _ptr(jac), _stream(stream)))


def rhs4sgcurv(corder, box, u, mu, la, met, jac, lu, onesided, acof, bope, ghcof, strx, stry, stream=None):
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    L.check(L.load().sw4hip_rhs4sgcurv(corder, *box, _ptr(u), _ptr(mu), _ptr(la), _ptr(met), _ptr(jac), _ptr(lu),
                                       _ptr(os_), _ptr(np.ascontiguousarray(acof)), _ptr(np.ascontiguousarray(bope)),
                                       _ptr(np.ascontiguousarray(ghcof)), _ptr(strx), _ptr(stry), _stream(stream)))


def addsgd4c(corder, box, up, u, um, rho, dcx, dcy, strx, stry, jac, cox, coy, beta, stream=None):
    L.check(L.load().sw4hip_addsgd4c(corder, *box, _ptr(up), _ptr(u), _ptr(um), _ptr(rho), _ptr(dcx), _ptr(dcy),
                                     _ptr(strx), _ptr(stry), _ptr(jac), _ptr(cox), _ptr(coy), beta,
                                     _stream(stream)))


def freesurfcurvisg(corder, box, nz, side, u, mu, la, met, sbop, forcing, strx, stry, stream=None):
    """side 5 (k = 1) or 6 (k = nz), numbered from 1 as in the reference;
    forcing None means zero traction."""
    sb = np.ascontiguousarray(sbop, dtype=np.float64)
    L.check(L.load().sw4hip_freesurfcurvisg(corder, *box, nz, side, _ptr(u), _ptr(mu), _ptr(la), _ptr(met), _ptr(sb),
                                            _ptr(forcing), _ptr(strx), _ptr(stry), _stream(stream)))


def enforce_cart_topo(corder, ij, cart_k, ucart, curv_k, ucurv, ghost_points=2, stream=None):
    """Grid interface of a curvilinear grid stacked on a Cartesian one
    (EW::enforceCartTopo): ij = (ib, ie, jb, je) common to both boxes,
    cart_k = (kb, ke) of the Cartesian box, curv_k = (kb, ke) of the
    curvilinear box."""
    L.check(L.load().sw4hip_enforce_cart_topo(corder, *ij, *cart_k, _ptr(ucart), *curv_k, _ptr(ucurv), ghost_points,
                                              _stream(stream)))


def bcfortsg(corder, box, wind, nx, ny, nz, u, h, bccnd, sbop, mu, la, bforce, strx, stry, stream=None):
    w = np.ascontiguousarray(wind, dtype=np.int32)
    bc = np.ascontiguousarray(bccnd, dtype=np.int32)
    sb = np.ascontiguousarray(sbop, dtype=np.float64)
    bf = (C.c_void_p * 6)(*[(_ptr(b).value if b is not None else None) for b in bforce])
    L.check(L.load().sw4hip_bcfortsg(corder, *box, _ptr(w), nx, ny, nz, _ptr(u), h, _ptr(bc), _ptr(sb), _ptr(mu),
                                     _ptr(la), C.cast(bf, C.c_void_p), _ptr(strx), _ptr(stry), _stream(stream)))


# transposed kernels of the discrete adjoint (component-blocked layout, box [-1, n+2])
def rhs4sg_adjoint(box, nk, onesided, acof, bope, ghcof, out, z, mu, la, h, strx, stry, strz, stream=None):
    """out = L^T z on every point (z read on the rows of rhs4sg)."""
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    L.check(L.load().sw4hip_rhs4sg_adjoint(*box, nk, _ptr(os_), _ptr(np.ascontiguousarray(acof)),
                                           _ptr(np.ascontiguousarray(bope)), _ptr(np.ascontiguousarray(ghcof)),
                                           _ptr(out), _ptr(z), _ptr(mu), _ptr(la), h, _ptr(strx), _ptr(stry),
                                           _ptr(strz), _stream(stream)))


def rhs4sg_matgrad(box, nk, onesided, acof, bope, ghcof, z, u, mu, la, h, strx, stry, strz, gmu, gla, stream=None):
    """gmu += d/dmu (z . L(u)), gla += d/dlambda (z . L(u))."""
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    L.check(L.load().sw4hip_rhs4sg_matgrad(*box, nk, _ptr(os_), _ptr(np.ascontiguousarray(acof)),
                                           _ptr(np.ascontiguousarray(bope)), _ptr(np.ascontiguousarray(ghcof)),
                                           _ptr(z), _ptr(u), _ptr(mu), _ptr(la), h, _ptr(strx), _ptr(stry),
                                           _ptr(strz), _ptr(gmu), _ptr(gla), _stream(stream)))


def addsgd4_adjoint(box, y, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta, dbar, grho,
                    stream=None):
    """For up -= D(u - um): dbar = -D^T y, grho += d/drho (-y . D)."""
    L.check(L.load().sw4hip_addsgd4_adjoint(*box, _ptr(y), _ptr(u), _ptr(um), _ptr(rho), _ptr(dcx), _ptr(dcy),
                                            _ptr(dcz), _ptr(strx), _ptr(stry), _ptr(strz), _ptr(cox), _ptr(coy),
                                            _ptr(coz), beta, _ptr(dbar), _ptr(grho), _stream(stream)))


def bcfortsg_adjoint(box, wind, nx, ny, nz, y, fwd, h, bccnd, sbop, mu, la, strx, stry, gmu, gla, stream=None):
    """y <- BC^T y (zero boundary data); gmu, gla += the free-surface closure's
    derivatives evaluated on fwd."""
    w = np.ascontiguousarray(wind, dtype=np.int32)
    bc = np.ascontiguousarray(bccnd, dtype=np.int32)
    sb = np.ascontiguousarray(sbop, dtype=np.float64)
    L.check(L.load().sw4hip_bcfortsg_adjoint(*box, _ptr(w), nx, ny, nz, _ptr(y), _ptr(fwd), h, _ptr(bc), _ptr(sb),
                                             _ptr(mu), _ptr(la), _ptr(strx), _ptr(stry), _ptr(gmu), _ptr(gla),
                                             _stream(stream)))


# --------------------------------------------------------------------------
# EW object
# --------------------------------------------------------------------------
class EW:
    """One Cartesian grid with the reference's time-stepping driver.

    Arrays span [-1, n+2] on every axis (two ghost points).  bc lists the
    boundary type of the six sides x-, x+, y-, y+, z-, z+.
    """

    def __init__(self, nx, ny, nz, h, bc, corder=1, sg_gp_thickness=20, sg_damping=0.02, cfl=1.3,
                 zmin=0.0, stream=None):
        self.lib = L.load()
        self.desc = L.grid_desc(nx, ny, nz, h, bc, corder=corder, zmin=zmin, sg_gp_thickness=sg_gp_thickness,
                                sg_damping=sg_damping, cfl=cfl)
        self.h = C.c_void_p()
        L.check(self.lib.sw4hip_ew_create(C.byref(self.desc), C.byref(self.h)))
        self.shape = (nz + 4, ny + 4, nx + 4)  # (k, j, i), i fastest
        self.corder = corder
        if stream is not None:
            L.check(self.lib.sw4hip_ew_set_stream(self.h, C.c_void_p(stream)))

    # set-up -------------------------------------------------------------
    def set_kernel_variant(self, variant, kchunk=0):
        """0: one thread per point; 1-10: LDS-staged k-marching tiles (1 plain
        at 2 waves/SIMD, 2 with register prefetch, 3 prefetch without a
        register cap, 4 as 3 with the corrector's damping in the same sweep, 5
        corrector as 4 and predictor as 1, 6 as 1 with an LDS damping sweep,
        7 as 6 with the predictor's planes loaded asynchronously into LDS
        where the row length is even, 8 as 7 with the acceleration stored by
        the predictor and loaded asynchronously by the corrector (one more
        field of device memory; even row lengths, otherwise as 7), 9 as 8
        with the damping sweep only over the tiles of the supergrid layers,
        where the damping profiles are nonzero (same bits), 10 as 9 with
        rows of odd length staged by four-byte asynchronous loads and the
        acceleration stored for them too (the default; same bits)).
        kchunk > 0: planes per workgroup of every sweep; 0: the library's
        defaults (64 for the predictor and corrector, 32 for the damping)."""
        L.check(self.lib.sw4hip_ew_set_kernel_variant(self.h, int(variant), int(kchunk)))

    def npoints(self):
        """Points of the box (two ghost points per side): the length of a
        scalar field, a third of a displacement field."""
        return int(np.prod(self.shape))

    def set_material(self, rho, mu, la):
        on_dev = _fields(self.npoints(), rho, mu, la)
        L.check(self.lib.sw4hip_ew_set_material(self.h, _ptr(rho), _ptr(mu), _ptr(la), int(on_dev)))

    def set_material_const(self, rho, cs, cp):
        L.check(self.lib.sw4hip_ew_set_material_const(self.h, rho, cs, cp))

    def add_moment_source(self, x, y, z, m, t0, freq, tfkind):
        mm = np.asarray(m, dtype=np.float64)
        L.check(self.lib.sw4hip_ew_add_moment_source(self.h, x, y, z, _ptr(mm), t0, freq, tfkind))

    def add_receiver(self, x, y, z, mode=0):
        """Receiver at the nearest grid point; mode 0 displacement, 1 velocity."""
        idx = C.c_int(0)
        L.check(self.lib.sw4hip_ew_add_receiver(self.h, x, y, z, int(mode), C.byref(idx)))
        self._nrec = getattr(self, "_nrec", 0) + 1
        return idx.value

    def receivers(self):
        """Recorded samples, array [nsamples, nreceivers, 3]."""
        n = C.c_int(0)
        L.check(self.lib.sw4hip_ew_get_receivers(self.h, C.byref(n), None))
        nrec = getattr(self, "_nrec", 0)
        out = np.zeros((n.value, nrec, 3))
        if n.value and nrec:
            L.check(self.lib.sw4hip_ew_get_receivers(self.h, C.byref(n), _ptr(out)))
        return out

    def setup(self, tmax=0.0, nsteps=0):
        L.check(self.lib.sw4hip_ew_setup(self.h, tmax, nsteps))

    def info(self):
        inf = L.EWInfo()
        L.check(self.lib.sw4hip_ew_get_info(self.h, C.byref(inf)))
        return inf

    def supergrid(self):
        n = {"x": self.desc.nx + 4, "y": self.desc.ny + 4, "z": self.desc.nz + 4}
        keys = ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz")
        out = {k: np.zeros(n[k[-1]]) for k in keys}
        L.check(self.lib.sw4hip_ew_get_supergrid(self.h, *[_ptr(out[k]) for k in keys]))
        return out

    def sources(self):
        inf = self.info()
        n, nu = inf.nsources, inf.nunique
        ijk = np.zeros(3 * n, dtype=np.int32)
        force = np.zeros(3 * n)
        freq = np.zeros(n)
        t0 = np.zeros(n)
        tf = np.zeros(n, dtype=np.int32)
        first = np.zeros(nu + 1, dtype=np.int32)
        L.check(self.lib.sw4hip_ew_get_sources(self.h, _ptr(ijk), _ptr(force), _ptr(freq), _ptr(t0), _ptr(tf),
                                               _ptr(first)))
        return {"ijk": ijk.reshape(-1, 3), "force": force.reshape(-1, 3), "freq": freq, "t0": t0,
                "tfkind": tf, "first": first}

    # state ---------------------------------------------------------------
    def set_solution(self, u, um):
        on_dev = _fields(3 * self.npoints(), u, um)
        L.check(self.lib.sw4hip_ew_set_solution(self.h, _ptr(u), _ptr(um), int(on_dev)))

    def get_solution(self):
        n = 3 * int(np.prod(self.shape))
        u, um = np.empty(n), np.empty(n)
        L.check(self.lib.sw4hip_ew_get_solution(self.h, _ptr(u), _ptr(um), 0))
        return u, um

    def device_array(self, which):
        p = C.c_void_p()
        L.check(self.lib.sw4hip_ew_device_array(self.h, which, C.byref(p)))
        return p.value

    # time stepping -------------------------------------------------------
    def set_topography(self, z):
        """Curvilinear grid of depth z over the box (numpy array or device
        tensor); before setup and receivers."""
        on_dev = _fields(self.npoints(), z)
        L.check(self.lib.sw4hip_ew_set_topography(self.h, _ptr(z), int(on_dev)))

    def timesteploop(self, nsteps=None, mode=0):
        """Advance nsteps (default: all remaining) steps; mode 0 fused sweeps,
        mode 1 the reference stage order, mode 2 the fused sweeps in FP32 with
        the compensated time update."""
        inf = self.info()
        if nsteps is None:
            nsteps = inf.nsteps - inf.cycle
        L.check(self.lib.sw4hip_ew_step(self.h, int(nsteps), int(mode)))

    def timesteploop_timed(self, nsteps, fp32=False):
        """Fused steps (FP64, or FP32 compensated) with per-stage GPU times
        (ms, summed over the steps)."""
        ms = np.zeros(7)
        fn = self.lib.sw4hip_ew_step_comp_timed if fp32 else self.lib.sw4hip_ew_step_timed
        L.check(fn(self.h, int(nsteps), _ptr(ms)))
        return ms

    def sync(self):
        L.check(self.lib.sw4hip_ew_sync(self.h))

    # discrete adjoint ------------------------------------------------------
    def reset(self):
        """Zero initial data at t = 0; receivers restart with the t = 0 sample."""
        L.check(self.lib.sw4hip_ew_reset(self.h))

    def gradient(self, grad_traces, nsteps=None, ckpt=0, out=None, fp32=False):
        """Gradient of a receiver misfit with respect to (rho, mu, lambda) by
        the discrete adjoint of the fused FP64 step.  grad_traces: the
        misfit's derivative with respect to every recorded sample, array
        [nsteps+1, nreceivers, 3].  Runs the forward from zero data itself.
        out: three device tensors to fill (else numpy arrays are returned).
        fp32: the forward and its recomputation in the FP32 form (step mode
        2) with float checkpoints, the reverse in FP64: the gradient of the
        FP32 run's misfit (sw4hip_ew_gradient_comp)."""
        g = np.ascontiguousarray(np.asarray(grad_traces, dtype=np.float64))
        if nsteps is None:
            nsteps = g.shape[0] - 1
        if g.ndim != 3 or g.shape[0] < nsteps + 1 or g.shape[2] != 3:
            raise ValueError("grad_traces must be [nsteps+1, nreceivers, 3], got %s" % (g.shape,))
        if out is None:
            n = self.npoints()
            res = [np.empty(n), np.empty(n), np.empty(n)]
            on_dev = 0
        else:
            res = list(out)
            on_dev = int(_fields(self.npoints(), *res))
        fn = self.lib.sw4hip_ew_gradient_comp if fp32 else self.lib.sw4hip_ew_gradient
        L.check(fn(self.h, int(nsteps), int(ckpt), _ptr(g), _ptr(res[0]), _ptr(res[1]), _ptr(res[2]), on_dev))
        return tuple(res)

    def gradient_plan(self, nsteps, ckpt=0):
        """(checkpoint interval, device bytes) that gradient(nsteps, ckpt)
        would use; ckpt <= 0 chooses the interval from free device memory."""
        used, nbytes = C.c_int(), C.c_double()
        L.check(self.lib.sw4hip_ew_gradient_plan(self.h, int(nsteps), int(ckpt), C.byref(used), C.byref(nbytes)))
        return used.value, nbytes.value

    def set_gradient_store(self, store):
        """Whether gradient() keeps the predicted field and acceleration of
        every step of a recomputed segment (two more states a step) so that
        the reverse steps skip recomputing their predictor: -1 when it fits
        and pays (default), 0 never, 1 always.  Same gradients either way."""
        L.check(self.lib.sw4hip_ew_set_gradient_store(self.h, int(store)))

    def release_workspace(self):
        """Free the device buffers that gradient() keeps between calls."""
        L.check(self.lib.sw4hip_ew_release_workspace(self.h))

    def point_source_error(self):
        li, l2, xi = C.c_double(), C.c_double(), C.c_double()
        L.check(self.lib.sw4hip_ew_point_source_error(self.h, C.byref(li), C.byref(l2), C.byref(xi)))
        return li.value, l2.value, xi.value

    def close(self):
        if self.h:
            self.lib.sw4hip_ew_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several set-up EW shots of any grid size advanced together (batched
    sweeps, optional hipGraph replay of the time loop)."""

    def __init__(self, shots):
        self.lib = L.load()
        self.shots = list(shots)
        arr = (C.c_void_p * len(self.shots))(*[s.h for s in self.shots])
        h = C.c_void_p()
        L.check(self.lib.sw4hip_batch_create(C.cast(arr, C.c_void_p), len(self.shots), C.byref(h)))
        self.h = h

    def set_kchunk(self, kc):
        L.check(self.lib.sw4hip_batch_set_kchunk(self.h, int(kc)))

    def tune(self, candidates=(16, 24, 32, 40, 48, 64, 96)):
        """Autotune the k-block; returns (best, {candidate: ms per step}).
        Resets every shot to zero data at t = 0."""
        cand = np.asarray(candidates, dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int(0)
        L.check(self.lib.sw4hip_batch_tune(self.h, len(cand), _ptr(cand), C.byref(best), _ptr(ms)))
        return best.value, dict(zip(cand.tolist(), ms.tolist()))

    def step(self, nsteps, graph=True):
        L.check(self.lib.sw4hip_batch_step(self.h, int(nsteps), int(bool(graph))))

    def gradient(self, grad_traces, nsteps, ckpt=0, out=None, max_concurrent=4, fp32=False):
        """Gradients of every shot's receiver misfit (EW.gradient per shot,
        run concurrently on streams of their own).  grad_traces: one array
        [nsteps+1, nreceivers, 3] per shot.  out: per shot three float64
        device tensors of the shot's point count (else numpy arrays are
        returned).  fp32: as EW.gradient.  Returns a list of (grho, gmu, gla)."""
        n = len(self.shots)
        if len(grad_traces) != n:
            raise ValueError("one trace array per shot")
        gs = [np.ascontiguousarray(np.asarray(g, dtype=np.float64)) for g in grad_traces]
        for g in gs:
            if g.ndim != 3 or g.shape[0] < nsteps + 1 or g.shape[2] != 3:
                raise ValueError("grad_traces must be [nsteps+1, nreceivers, 3]")
        if out is None:
            res = [(np.empty(s.npoints()), np.empty(s.npoints()), np.empty(s.npoints())) for s in self.shots]
            on_dev = 0
        else:
            res = [tuple(o) for o in out]
            on_dev = int(all(_fields(s.npoints(), *o) for s, o in zip(self.shots, res)))
            if not on_dev:
                raise ValueError("out must hold device tensors")
        arr = lambda xs: (C.c_void_p * n)(*[_ptr(x).value for x in xs])  # noqa: E731
        fn = self.lib.sw4hip_batch_gradient_comp if fp32 else self.lib.sw4hip_batch_gradient
        L.check(fn(self.h, int(nsteps), int(ckpt), C.cast(arr(gs), C.c_void_p),
                   C.cast(arr([r[0] for r in res]), C.c_void_p), C.cast(arr([r[1] for r in res]), C.c_void_p),
                   C.cast(arr([r[2] for r in res]), C.c_void_p), on_dev, int(max_concurrent)))
        return res

    def release_workspace(self):
        """Free the device work memory that gradient() keeps between calls."""
        L.check(self.lib.sw4hip_batch_release_workspace(self.h))

    def close(self):
        if self.h:
            self.lib.sw4hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



# --------------------------------------------------------------------------
# input decks, receiver files, checkpoints
# --------------------------------------------------------------------------
def _report(r):
    return {f: getattr(r, f) for f, _ in L.DeckReport._fields_}


def parse_deck(text):
    """Grid and step settings of deck text (no GPU); raises on commands the
    front end does not run."""
    r = L.DeckReport()
    L.check(L.load().sw4hip_parse_deck(text.encode(), C.byref(r)))
    return _report(r)


def run_deck(deck_file, path=None, mode=0):
    """Run a reference input deck on the current GPU (sw4hip_run_deck):
    receiver files, checkpoints and PointSourceErr.txt go to the deck's
    fileio path or to `path`.  Returns the run's report as a dict."""
    r = L.DeckReport()
    L.check(L.load().sw4hip_run_deck(str(deck_file).encode(), None if path is None else str(path).encode(),
                                     int(mode), C.byref(r)))
    return _report(r)


def write_station(path, file, samples, dt, t0=0.0, sta=None, mode=0, nsew=False, usgs=True, sac=True,
                  loc=(0.0, 0.0, 0.0), grid_point=None, utc=(1970, 1, 1, 0, 0, 0, 0)):
    """Receiver files of one station (TimeSeries::writeFile formats);
    samples: array [n, 3] of x, y, z components."""
    s = np.ascontiguousarray(np.asarray(samples, dtype=np.float64).reshape(-1, 3))
    st = L.Station()
    st.file = file.encode()
    st.sta = None if sta is None else sta.encode()
    st.mode, st.nsew, st.usgs, st.sac = int(mode), int(bool(nsew)), int(bool(usgs)), int(bool(sac))
    st.x, st.y, st.z = loc
    st.gx, st.gy, st.gz = loc if grid_point is None else grid_point
    for q in range(7):
        st.utc[q] = int(utc[q])
    L.check(L.load().sw4hip_write_station(C.byref(st), str(path).encode(), float(t0), float(dt), s.shape[0], _ptr(s)))


def write_checkpoint(file, time, cycle, nx, ny, nz, um, u):
    """CheckPoint file of (Um, U), host arrays over the box [-1, n+2],
    component-blocked."""
    n = 3 * (nx + 4) * (ny + 4) * (nz + 4)
    _fields(n, um, u)
    L.check(L.load().sw4hip_write_checkpoint(str(file).encode(), float(time), int(cycle), nx, ny, nz, _ptr(um),
                                             _ptr(u)))


def read_checkpoint(file, nx, ny, nz):
    """(time, cycle, um, u) of a CheckPoint file; um, u over the box
    (ghost points zero)."""
    n = 3 * (nx + 4) * (ny + 4) * (nz + 4)
    um, u = np.zeros(n), np.zeros(n)
    t, cyc = C.c_double(), C.c_int()
    L.check(L.load().sw4hip_read_checkpoint(str(file).encode(), nx, ny, nz, C.byref(t), C.byref(cyc), _ptr(um),
                                            _ptr(u)))
    return t.value, cyc.value, um, u
