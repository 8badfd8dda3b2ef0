"""The reference time-step loop (EW::timesteploop, EW.C:2527-2842) driven on
the CPU with the oracle kernels.  Set-up products (supergrid arrays, ghost
windows, grid-point sources) come from the library's host-side set-up calls,
which need no GPU.

TEST INFRASTRUCTURE ONLY (tests/ and the cpu_baseline leg of bench.py).
With kernels="reference" the two stencil stages (rhs4sg_rev, addsgd4) run
the reference's own code compiled by oracle/Makefile into oracle/_ref."""
import numpy as np

from oracle import oracle as O
from sw4lite_amd import ew
from sw4lite_amd import lib as L


def grid_box(nx, ny, nz):
    return (-1, nx + 2, -1, ny + 2, -1, nz + 2)


def npts(box):
    ib, ie, jb, je, kb, ke = box
    return (ie - ib + 1) * (je - jb + 1) * (ke - kb + 1)


def curvilinear_eigmax(met, jac, rho, mu, la, corder, offdiag=1.0):
    """max over the points of -w0 / (4 rho), w0 the most negative eigenvalue
    of the principal-part matrix of the curvilinear operator (EW::computeDT,
    EW.C:5063-5136; the reference's LAPACK dspev here numpy's eigvalsh).
    met: 4 components over the box (layout corder), jac, rho, mu, la over the
    box.  dt = CFL / sqrt(this).  offdiag scales the off-diagonal entries
    (1 in the scheme; tests use other values to show that they matter)."""
    n = jac.size
    m = met.reshape(4, n) if corder == 1 else met.reshape(n, 4).T
    l2, ji = la + 2 * mu, 1 / jac
    s1, s2, s3, s4 = (m[c] ** 2 for c in range(4))
    A = np.empty((n, 3, 3))
    A[:, 0, 0] = -4 * (s1 * l2 + s1 * mu + s2 * l2 + s3 * mu + s4 * mu) * ji
    A[:, 0, 1] = A[:, 1, 0] = -4 * offdiag * m[1] * m[2] * (mu + la) * ji
    A[:, 0, 2] = A[:, 2, 0] = -4 * offdiag * m[1] * m[3] * (mu + la) * ji
    A[:, 1, 1] = -4 * (s1 * mu + s1 * l2 + s2 * mu + s3 * l2 + s4 * mu) * ji
    A[:, 1, 2] = A[:, 2, 1] = -4 * offdiag * m[2] * m[3] * (mu + la) * ji
    A[:, 2, 2] = -4 * (2 * s1 * mu + s2 * mu + s3 * mu + s4 * l2) * ji
    w0 = np.linalg.eigvalsh(A)[:, 0]
    return np.max(-w0 / (4 * rho))


class OracleRun:
    def __init__(self, nx, ny, nz, h, bc, rho, mu, la, sources, corder=1, sg_gp=20, beta=0.02, cfl=1.3,
                 tmax=0.0, nsteps=0, kernels="oracle", zgrid=None):
        """rho, mu, la: arrays over the box (npts) or scalars.  sources:
        list of dicts x,y,z,m(6),t0,freq,tfkind.  zgrid: grid depth over the
        box for a curvilinear grid (rhs4sgcurv, addsgd4cfort, freesurfcurvisg;
        EW.C:3218-3243, 3479-3497, 3560-3582)."""
        self.stencil_lib = None
        if kernels == "reference":
            self.stencil_lib = O.reference()
            if self.stencil_lib is None:
                raise RuntimeError("oracle/_ref/libsw4ref.so is not built")
        self.desc = L.grid_desc(nx, ny, nz, h, bc, corder=corder, sg_gp_thickness=sg_gp, sg_damping=beta, cfl=cfl)
        self.nx, self.ny, self.nz, self.h, self.bc, self.corder, self.beta = nx, ny, nz, h, list(bc), corder, beta
        self.box = grid_box(nx, ny, nz)
        n = self.n = npts(self.box)
        self.rho = np.broadcast_to(np.asarray(rho, float), (n,)).copy()
        self.mu = np.broadcast_to(np.asarray(mu, float), (n,)).copy()
        self.la = np.broadcast_to(np.asarray(la, float), (n,)).copy()
        self.sg = ew.host_supergrid(self.desc)
        self.wind, _ = ew.host_boundary_windows(self.desc)
        self.onesided = [0, 0, 0, 0, int(bc[4] == L.BC_FREE), int(bc[5] == L.BC_FREE)]
        self.acof, self.ghcof, self.bope, self.sbop = ew.stencil_coefficients()
        self.zgrid = None if zgrid is None else np.ascontiguousarray(zgrid, dtype=np.float64)
        if self.zgrid is not None:
            self.met, self.jac = np.zeros(4 * n), np.zeros(n)
            if O.metric(1, self.box, h, self.zgrid, self.met, self.jac) != 0:
                raise ValueError("folded curvilinear grid")
            if corder == 0:
                self.met = np.ascontiguousarray(self.met.reshape(4, n).T.reshape(-1))
            self.onesided[5] = 0
        # EW::computeDT (EW.C:5041-5146); curvilinear: most negative eigenvalue
        # of the principal-part matrix per point (EW.C:5063-5136), here from
        # numpy's symmetric eigensolver
        if self.zgrid is None:
            eig = np.max((4 * self.mu + self.la) / self.rho)
            dt = min(1e10, cfl * h * (1 / np.sqrt(eig)))
        else:
            eig = curvilinear_eigmax(self.met, self.jac, self.rho, self.mu, self.la, corder)
            dt = min(1e10, cfl * (1 / np.sqrt(eig)))
        if tmax > 0:
            nst = int(tmax / dt + 0.5)
            nst = max(nst, 1)
            dt = tmax / nst
        else:
            nst = nsteps
        self.dt, self.nsteps = dt, nst
        # grid-point sources, sorted by grid index (stable)
        ijk, frc, fq, t0, tf = [], [], [], [], []
        for s in sources:
            if self.zgrid is None:
                a, b = ew.host_moment_source(self.desc, s["x"], s["y"], s["z"], s["m"])
            else:
                a, b = ew.host_moment_source_curv(self.desc, self.zgrid, s["x"], s["y"], s["z"], s["m"])
            ijk.append(a)
            frc.append(b)
            fq += [s["freq"]] * len(a)
            t0 += [s["t0"]] * len(a)
            tf += [s["tfkind"]] * len(a)
        if ijk:
            ijk = np.concatenate(ijk)
            frc = np.concatenate(frc)
            ni, nij = nx + 4, (nx + 4) * (ny + 4)
            key = (ijk[:, 0] + 1) + ni * (ijk[:, 1] + 1) + nij * (ijk[:, 2] + 1)
            order = np.argsort(key, kind="stable")
            self.src = (ijk[order], frc[order], np.array(fq)[order], np.array(t0)[order], np.array(tf)[order])
        else:
            self.src = None
        # every array touched in parallel before stepping (EW.C:2390-2398)
        self.U = O.filled(3 * n)
        self.Um = O.filled(3 * n)
        self.Up = O.filled(3 * n)
        self.Lu = O.filled(3 * n)
        self.F = O.filled(3 * n)
        self.Uacc = O.filled(3 * n)
        self.t = 0.0
        self.cycle = 0
        ni, nj = nx + 4, ny + 4
        self.bforce = []
        wind = self.wind.reshape(6, 6)
        for s in range(6):
            w = wind[s]
            m = max(0, (w[1] - w[0] + 1) * (w[3] - w[2] + 1) * (w[5] - w[4] + 1))
            if bc[s] == L.BC_FREE:
                m = ni * nj
            self.bforce.append(np.zeros(3 * max(m, 1)))

    def force(self, t, tt):
        if self.src is None:
            self.F[:] = 0
            return
        ijk, frc, fq, t0, tf = self.src
        O.force(self.corder, self.box, self.F, ijk.ravel(), frc.ravel(), fq, t0, tf, t, int(tt))

    def rhs(self, u):
        sg = self.sg
        if self.zgrid is not None:
            O.rhs4sgcurv(self.corder, self.box, u, self.mu, self.la, self.met, self.jac, self.Lu, self.onesided,
                         self.acof, self.bope, self.ghcof, sg["strx"], sg["stry"], lib=self.stencil_lib)
            return
        O.rhs4sg(self.corder, self.box, self.nz, self.onesided, self.acof, self.bope, self.ghcof, self.Lu, u, self.mu,
                 self.la, self.h, sg["strx"], sg["stry"], sg["strz"], lib=self.stencil_lib)

    def enforce_bc(self, u):
        if self.zgrid is None:
            O.bcfortsg(self.corder, self.box, self.wind, self.nx, self.ny, self.nz, u, self.h, self.bc, self.sbop,
                       self.mu, self.la, self.bforce, self.sg["strx"], self.sg["stry"])
            return
        # EW::enforceBC (EW.C:3446-3501): bcfortsg, then the curved closure
        # on the top, which replaces the Cartesian one at the same points
        bc = [8 if b == L.BC_FREE else b for b in self.bc]
        O.bcfortsg(self.corder, self.box, self.wind, self.nx, self.ny, self.nz, u, self.h, bc, self.sbop,
                   self.mu, self.la, self.bforce, self.sg["strx"], self.sg["stry"])
        if self.bc[4] == L.BC_FREE:
            O.freesurfcurvisg(self.corder, self.box, self.nz, 5, u, self.mu, self.la, self.met, self.sbop,
                              np.zeros(3 * (self.nx + 4) * (self.ny + 4)), self.sg["strx"], self.sg["stry"])

    def damp(self):
        sg = self.sg
        if self.zgrid is not None:
            O.addsgd4c(self.corder, self.box, self.Up, self.U, self.Um, self.rho, sg["dcx"], sg["dcy"], sg["strx"],
                       sg["stry"], self.jac, sg["cox"], sg["coy"], self.beta)
            return
        O.addsgd4(self.corder, self.box, self.Up, self.U, self.Um, self.rho, sg["dcx"], sg["dcy"], sg["dcz"],
                  sg["strx"], sg["stry"], sg["strz"], sg["cox"], sg["coy"], sg["coz"], self.beta,
                  lib=self.stencil_lib)

    def step(self, n=1):
        sg = self.sg
        for _ in range(n):
            t, dt = self.t, self.dt
            self.force(t, False)
            self.rhs(self.U)
            O.predfort(self.corder, self.box, self.Up, self.U, self.Um, self.Lu, self.F, self.rho, dt * dt)
            self.enforce_bc(self.Up)
            self.force(t, True)
            O.dpdmtfort(self.box, self.Up, self.U, self.Um, self.Uacc, 1.0 / (dt * dt))
            self.rhs(self.Uacc)
            O.corrfort(self.corder, self.box, self.Up, self.Lu, self.F, self.rho, dt * dt * dt * dt)
            self.damp()
            self.enforce_bc(self.Up)
            self.t += dt
            self.cycle += 1
            self.Um, self.U, self.Up = self.U, self.Up, self.Um

    def point_source_error(self, src):
        """EW::exactSol + normOfDifference for source src (EW.C:3650-3669, 4526-4610)."""
        n = self.n
        nx, ny, nz, h = self.nx, self.ny, self.nz, self.h
        im = (self.box[0] + self.box[1]) // 2
        jm = (self.box[2] + self.box[3]) // 2
        km = (self.box[4] + self.box[5]) // 2
        p = (im - self.box[0]) + (nx + 4) * ((jm - self.box[2]) + (ny + 4) * (km - self.box[4]))
        rho, mu, la = self.rho[p], self.mu[p], self.la[p]
        beta, alpha = np.sqrt(mu / rho), np.sqrt((2 * mu + la) / rho)
        fr = src["freq"]
        time = (self.t - src["t0"]) * src["freq"]
        if src["tfkind"] == L.TF_GAUSSIAN:
            fr = 1 / fr
            time = time * fr
        uex = np.zeros(3 * n)
        O.exact_moment(self.box, h, 0.0, src["x"], src["y"], src["z"], rho, alpha, beta, time, fr, src["m"],
                       src["tfkind"], uex, self.corder)
        sgp = self.desc.sg_gp_thickness
        bc = self.bc
        win = [max(1, sgp + 1) if bc[0] == L.BC_SUPERGRID else 1, min(nx, nx - sgp) if bc[1] == L.BC_SUPERGRID else nx,
               max(1, sgp + 1) if bc[2] == L.BC_SUPERGRID else 1, min(ny, ny - sgp) if bc[3] == L.BC_SUPERGRID else ny,
               max(1, sgp + 1) if bc[4] == L.BC_SUPERGRID else 1, min(nz, nz - sgp) if bc[5] == L.BC_SUPERGRID else nz]
        li, l2, xli = O.solerr3(self.box, h, uex, self.U, 0.0, src["x"], src["y"], src["z"], 4 * h, win, self.corder)
        return li, np.sqrt(l2), xli, uex
