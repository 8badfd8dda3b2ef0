"""ctypes wrapper of oracle/_ref/libsw4ew.so: the reference program (the EW
class and everything it calls, main.C aside) compiled from its own sources,
with the C-linkage driver oracle/ref_ew.cpp.

TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/ scripts in the
container that holds the reference sources, to check the oracle.  The GPU
box has no such library; every user skips when ``available()`` is False.

An object is made by running an input deck through the reference's own
constructor (EW::EW, EW.C:162-224); its member kernels are then called on
caller arrays.  The reference prints its progress on stdout; ``RefEW``
silences file descriptor 1 while it runs the deck.
"""
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EW_SO = os.path.join(HERE, "_ref", "libsw4ew.so")

_lib = None
D = C.c_double
P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def available():
    return os.path.exists(EW_SO)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libsw4ew.so is not built (make -C oracle ref)")
        _lib = C.CDLL(EW_SO)
        _lib.refew_create.restype = P
        _lib.refew_array.restype = C.c_long
        _lib.refew_timefunc.restype = D
        _lib.refew_timefunc.argtypes = [C.c_int, C.c_int, D, D, D]
        _lib.refew_point_sources.restype = C.c_int
    return _lib


class _Quiet:
    """Redirects file descriptor 1 (the reference's cout) to /dev/null."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *a):
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


class RefEW:
    """An EW object of the reference built from deck text.  The deck's
    ``fileio path=`` is set to a temporary directory (``self.path``) unless
    the text gives one; output files of the run (PointSourceErr.txt,
    receiver files, checkpoints) land there."""

    def __init__(self, deck, quiet=True, workdir=None):
        self.tmp = None
        if workdir is None:
            self.tmp = tempfile.TemporaryDirectory(prefix="refew_")
            workdir = self.tmp.name
        self.path = os.path.join(workdir, "out")
        lines = [ln for ln in deck.strip().splitlines()]
        if not any(ln.strip().startswith("fileio") for ln in lines):
            lines.insert(0, "fileio path=%s verbose=0" % self.path)
        fn = os.path.join(workdir, "deck.in")
        with open(fn, "w") as f:
            f.write("\n".join(lines) + "\n")
        L = lib()
        if quiet:
            with _Quiet():
                self.h = L.refew_create(fn.encode())
        else:
            self.h = L.refew_create(fn.encode())
        dt, n, g, box = D(), C.c_int(), C.c_int(), (C.c_int * 24)()
        L.refew_info(P(self.h), C.byref(dt), C.byref(n), C.byref(g), box)
        self.dt, self.nsteps, self.ngrids = dt.value, n.value, g.value
        self.boxes = [tuple(box[6 * q:6 * q + 6]) for q in range(min(4, self.ngrids))]

    def close(self):
        if self.h:
            lib().refew_destroy(P(self.h))
            self.h = None
        if self.tmp is not None:
            self.tmp.cleanup()
            self.tmp = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- state
    def set_corder(self, corder):
        lib().refew_set_corder(P(self.h), int(corder))

    def grid(self, g=0):
        h, zmin, nz = D(), D(), C.c_int()
        lib().refew_grid(P(self.h), g, C.byref(h), C.byref(zmin), C.byref(nz))
        return h.value, zmin.value, nz.value

    def array(self, which, g=0):
        """0 U, 1 Um, 2 rho, 3 mu, 4 lambda, 5 x, 6 y, 7 z, 8 metric, 9 J."""
        n = lib().refew_array(P(self.h), g, which, None)
        out = np.empty(n)
        lib().refew_array(P(self.h), g, which, _p(out))
        return out

    def boundary(self, g=0):
        """(wind[36], npts[6], bctype[6], onesided[6]) of grid g."""
        w, n, t, o = (np.zeros(36, dtype=np.int32), np.zeros(6, dtype=np.int32), np.zeros(6, dtype=np.int32),
                      np.zeros(6, dtype=np.int32))
        lib().refew_boundary(P(self.h), g, _p(w), _p(n), _p(t), _p(o))
        return w, n, t, o

    def supergrid(self, g=0):
        ib, ie, jb, je, kb, ke = self.boxes[g]
        ni, nj, nk = ie - ib + 1, je - jb + 1, ke - kb + 1
        out = np.empty(3 * (ni + nj + nk))
        lib().refew_supergrid(P(self.h), g, _p(out))
        names = ["dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz"]
        lens = [ni, nj, nk] * 3
        res, o = {}, 0
        for nm, ln in zip(names, lens):
            res[nm] = out[o:o + ln].copy()
            o += ln
        return res

    def point_sources(self, t=0.0):
        """(ijkg [n,4], forces without time function [n,3], F(t) [n,3],
        F_tt(t) [n,3])."""
        n = lib().refew_point_sources(P(self.h), D(t), None, None, None, None)
        ijkg = np.zeros((n, 4), dtype=np.int32)
        f, ft, ftt = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        lib().refew_point_sources(P(self.h), D(t), _p(ijkg), _p(f), _p(ft), _p(ftt))
        return ijkg, f, ft, ftt

    def exact_point_source(self, t, g=0):
        n = 3 * (lambda b: (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1))(self.boxes[g])
        out = np.zeros(n)
        lib().refew_exact_point_source(P(self.h), g, D(t), _p(out))
        return out

    def point_source_error(self, t):
        """(Linf, L2, max|u_exact|) of the current U at time t."""
        li, l2, xi = D(), D(), D()
        lib().refew_point_source_error(P(self.h), D(t), C.byref(li), C.byref(l2), C.byref(xi))
        return li.value, l2.value, xi.value

    # ----------------------------------------------------- member kernels
    def predfort(self, corder, box, up, u, um, lu, fo, rho, dt2):
        self.set_corder(corder)
        lib().refew_predfort(P(self.h), *box, _p(up), _p(u), _p(um), _p(lu), _p(fo), _p(rho), D(dt2))

    def corrfort(self, corder, box, up, lu, fo, rho, dt4):
        self.set_corder(corder)
        lib().refew_corrfort(P(self.h), *box, _p(up), _p(lu), _p(fo), _p(rho), D(dt4))

    def dpdmtfort(self, box, up, u, um, uacc, dt2i):
        lib().refew_dpdmtfort(P(self.h), *box, _p(up), _p(u), _p(um), _p(uacc), D(dt2i))

    def solerr3(self, box, h, uex, u, zmin, x0, y0, z0, radius, win):
        li, l2, xli = D(), D(), D()
        lib().refew_solerr3(P(self.h), *box, D(h), _p(uex), _p(u), C.byref(li), C.byref(l2), C.byref(xli), D(zmin),
                            D(x0), D(y0), D(z0), D(radius), *win)
        return li.value, l2.value, xli.value

    def bcfortsg(self, corder, box, wind, nx, ny, nz, u, h, bccnd, sbop, mu, la, bforce, strx, stry):
        w = np.ascontiguousarray(wind, dtype=np.int32)
        bc = np.ascontiguousarray(bccnd, dtype=np.int32)
        bf = (P * 6)(*[(b.ctypes.data if b is not None else None) for b in bforce])
        lib().refew_bcfortsg(P(self.h), corder, *box, _p(w), nx, ny, nz, _p(u), D(h), _p(bc),
                             _p(np.ascontiguousarray(sbop, dtype=np.float64)), _p(mu), _p(la), bf, _p(strx),
                             _p(stry))

    def addsgd4(self, corder, box, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta):
        lib().refew_addsgd4(P(self.h), corder, *box, _p(up), _p(u), _p(um), _p(rho), _p(dcx), _p(dcy), _p(dcz),
                            _p(strx), _p(stry), _p(strz), _p(cox), _p(coy), _p(coz), D(beta))

    def addsgd4c(self, corder, box, up, u, um, rho, dcx, dcy, strx, stry, jac, cox, coy, beta):
        lib().refew_addsgd4c(P(self.h), corder, *box, _p(up), _p(u), _p(um), _p(rho), _p(dcx), _p(dcy), _p(strx),
                             _p(stry), _p(jac), _p(cox), _p(coy), D(beta))

    def stencil_coefficients(self):
        acof, ghcof, bope, sbop = np.zeros(384), np.zeros(6), np.zeros(48), np.zeros(5)
        lib().refew_stencil_coefficients(P(self.h), _p(acof), _p(ghcof), _p(bope), _p(sbop))
        return acof, ghcof, bope, sbop

    def freesurfcurvisg(self, corder, box, nz, side, u, mu, la, met, sbop, forcing, strx, stry):
        lib().refew_freesurfcurvisg(P(self.h), corder, *box, nz, side, _p(u), _p(mu), _p(la), _p(met),
                                    _p(np.ascontiguousarray(sbop, dtype=np.float64)), _p(forcing), _p(strx),
                                    _p(stry))

    def metricexgh(self, corder, box, nz, x, y, z, met, jac, order, sb, zmax, amp, xc, yc, xl, yl):
        lib().refew_metricexgh(P(self.h), corder, *box, nz, _p(x), _p(y), _p(z), _p(met), _p(jac), order, D(sb),
                               D(zmax), D(amp), D(xc), D(yc), D(xl), D(yl))


def timefunc(kind, deriv2, freq, t, par0=-700.0):
    """The reference's time function (0 Ricker, 1 Gaussian, 2 C6SmoothBump)."""
    return lib().refew_timefunc(kind, deriv2, freq, t, par0)


def supergrid_taper(left, x0, right, x1, width, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, d, c = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    lib().refew_supergrid_taper(int(left), D(x0), int(right), D(x1), D(width), len(x), _p(x), _p(s), _p(d), _p(c))
    return s, d, c
