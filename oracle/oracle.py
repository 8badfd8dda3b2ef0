"""ctypes wrapper of the CPU oracle (liboracle.so) and, when it has been
built from the reference sources, of oracle/_ref/libsw4ref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the sw4lite_amd package.
All arrays are numpy float64 / int32, C-contiguous, in the Sarray layouts.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsw4ref.so")

_orc = None
_ref = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build(ref=None):
    """Compile liboracle.so (and the reference library when ref is True or,
    by default, when the reference source tree exists)."""
    targets = ["all"]
    if ref is None:
        ref = os.path.isdir("/root/reference/src")
    if ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _orc = C.CDLL(ORACLE_SO)
        _orc.orc_timefunc.restype = C.c_double
        _orc.orc_timefunc.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        _orc.orc_bcfortsg.restype = C.c_int
    return _orc


def reference():
    """The reference kernels compiled from their own sources, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
    return _ref


D = C.c_double


def rhs4sg(corder, box, nk, onesided, acof, bope, ghcof, lu, u, mu, la, h, strx, stry, strz, lib=None):
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    if lib is None:
        oracle().orc_rhs4sg(*box, nk, _p(os_), _p(acof), _p(bope), _p(ghcof), _p(lu), _p(u), _p(mu), _p(la),
                            D(h), _p(strx), _p(stry), _p(strz), corder)
    else:
        lib.ref_rhs4sg(corder, *box, nk, _p(os_), _p(acof), _p(bope), _p(ghcof), _p(lu), _p(u), _p(mu), _p(la),
                       D(h), _p(strx), _p(stry), _p(strz))


def filled(n, v=0.0):
    """A float64 array of n entries, first touched in parallel."""
    a = np.empty(n)
    oracle().orc_fill(C.c_ssize_t(n), _p(a), D(v))
    return a


def predfort(corder, box, up, u, um, lu, fo, rho, dt2):
    oracle().orc_predfort(*box, _p(up), _p(u), _p(um), _p(lu), _p(fo), _p(rho), D(dt2), corder)


def corrfort(corder, box, up, lu, fo, rho, dt4):
    oracle().orc_corrfort(*box, _p(up), _p(lu), _p(fo), _p(rho), D(dt4), corder)


def dpdmtfort(box, up, u, um, uacc, dt2i):
    oracle().orc_dpdmtfort(*box, _p(up), _p(u), _p(um), _p(uacc), D(dt2i))


def addsgd4(corder, box, up, u, um, rho, dcx, dcy, dcz, strx, stry, strz, cox, coy, coz, beta, lib=None):
    args = [_p(up), _p(u), _p(um), _p(rho), _p(dcx), _p(dcy), _p(dcz), _p(strx), _p(stry), _p(strz), _p(cox),
            _p(coy), _p(coz), D(beta)]
    if lib is None:
        oracle().orc_addsgd4(*box, *args, corder)
    else:
        lib.ref_addsgd4(corder, *box, *args)


def bcfortsg(corder, box, wind, nx, ny, nz, u, h, bccnd, sbop, mu, la, bforce, strx, stry):
    w = np.ascontiguousarray(wind, dtype=np.int32)
    bc = np.ascontiguousarray(bccnd, dtype=np.int32)
    bf = (C.c_void_p * 6)(*[(b.ctypes.data if b is not None else None) for b in bforce])
    rc = oracle().orc_bcfortsg(*box, _p(w), nx, ny, nz, _p(u), D(h), _p(bc), _p(np.ascontiguousarray(sbop)),
                               _p(mu), _p(la), bf, _p(strx), _p(stry), corder)
    if rc != 0:
        raise ValueError("orc_bcfortsg rejected the boundary types")


def timefunc(kind, deriv2, freq, t):
    return oracle().orc_timefunc(kind, deriv2, D(freq), D(t), D(-700.0))


def force(corder, box, F, ijk, frc, freq, t0, tfkind, t, deriv2):
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    frc = np.ascontiguousarray(frc, dtype=np.float64)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    t0 = np.ascontiguousarray(t0, dtype=np.float64)
    tf = np.ascontiguousarray(tfkind, dtype=np.int32)
    oracle().orc_force(*box, _p(F), len(freq), _p(ijk), _p(frc), _p(freq), _p(t0), _p(tf), D(t), deriv2, corder)


def exact_moment(box, h, zmin, x0, y0, z0, rho, alpha, beta, time, f, m6, kind, out, corder):
    m6 = np.ascontiguousarray(m6, dtype=np.float64)
    oracle().orc_exact_moment(*box, D(h), D(zmin), D(x0), D(y0), D(z0), D(rho), D(alpha), D(beta), D(time), D(f),
                              _p(m6), kind, _p(out), corder)


def solerr3(box, h, uex, u, zmin, x0, y0, z0, radius, win, corder):
    li, l2, xli = D(), D(), D()
    oracle().orc_solerr3(*box, D(h), _p(uex), _p(u), C.byref(li), C.byref(l2), C.byref(xli), D(zmin), D(x0),
                         D(y0), D(z0), D(radius), *win, corder)
    return li.value, l2.value, xli.value


# ---------------------------------------------------------------- curvilinear
def metric(corder, box, h, z, met, jac, lib=None):
    """met (4 components), jac from the grid depth z; returns the kernel's
    status (0, or -1 where z is not increasing in k).  With lib (the reference
    build) x and y are made from h as the reference reads h off x."""
    if lib is None:
        return oracle().orc_metric(*box, D(h), _p(z), _p(met), _p(jac), corder)
    ib, ie, jb, je, kb, ke = box
    ni, nj, nk = ie - ib + 1, je - jb + 1, ke - kb + 1
    x = np.broadcast_to((np.arange(ib, ie + 1) - 1) * h, (nk, nj, ni)).copy()
    y = np.broadcast_to(((np.arange(jb, je + 1) - 1) * h)[:, None], (nk, nj, ni)).copy()
    lib.ref_metric.restype = C.c_int
    return lib.ref_metric(corder, *box, _p(x), _p(y), _p(z), _p(met), _p(jac))


def rhs4sgcurv(corder, box, u, mu, la, met, jac, lu, onesided, acof, bope, ghcof, strx, stry, lib=None):
    os_ = np.ascontiguousarray(onesided, dtype=np.int32)
    args = [_p(u), _p(mu), _p(la), _p(met), _p(jac), _p(lu), _p(os_), _p(acof), _p(bope), _p(ghcof), _p(strx),
            _p(stry)]
    if lib is None:
        oracle().orc_rhs4sgcurv(*box, *args, corder)
    else:
        lib.ref_rhs4sgcurv(corder, *box, *args)


def addsgd4c(corder, box, up, u, um, rho, dcx, dcy, strx, stry, jac, cox, coy, beta):
    oracle().orc_addsgd4c(*box, _p(up), _p(u), _p(um), _p(rho), _p(dcx), _p(dcy), _p(strx), _p(stry), _p(jac),
                          _p(cox), _p(coy), D(beta), corder)


def freesurfcurvisg(corder, box, nz, side, u, mu, la, met, sbop, forcing, strx, stry):
    oracle().orc_freesurfcurvisg(*box, nz, side, _p(u), _p(mu), _p(la), _p(met), _p(np.ascontiguousarray(sbop)),
                                 _p(forcing), _p(strx), _p(stry), corder)
