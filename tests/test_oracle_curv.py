"""Pins the CPU oracle's curvilinear kernels:
  * metric and rhs4sgcurv against the reference's own sources compiled by
    oracle/Makefile (tests/testil/metricC.C, src/rhs4sgcurv.C and
    rhs4sgcurv_rev.C), when that build exists in this container;
  * against committed golden vectors made from them (tests/golden/curv.npz);
  * on a flat grid, against the Cartesian kernels (rhs4sg, addsgd4, the
    Cartesian free-surface closure), which are pinned against the reference;
  * the free-surface closure by the traction it leaves on a curved grid,
    computed here with the forward 3x3 normal-stress matrix;
  * the metric by its closed form on a grid that is linear in k.
"""
import json
import os

import numpy as np
import pytest

from common import ROOT, grid_box, hills_depth, npts, random_inputs, relerr, to_layout
from oracle import oracle as O

GOLD = os.path.join(ROOT, "tests", "golden")


def metric_of(box, h, z, corder, lib=None):
    n = npts(box)
    met, jac = np.zeros(4 * n), np.zeros(n)
    assert O.metric(corder, box, h, z, met, jac, lib=lib) == 0
    return met, jac


def met_layout(met_blocked, box, corder):
    if corder == 1:
        return met_blocked
    n = npts(box)
    return np.ascontiguousarray(met_blocked.reshape(4, n).T.reshape(-1))


def ref_or_skip():
    lib = O.reference()
    if lib is None:
        pytest.skip("reference kernels not built here (oracle/_ref)")
    return lib


@pytest.mark.parametrize("corder", [0, 1])
def test_metric_matches_reference(corder):
    lib = ref_or_skip()
    box, z = hills_depth(14, 12, 11, 0.1, 3)
    got = metric_of(box, 0.1, z, corder)
    want = metric_of(box, 0.1, z, corder, lib=lib)
    assert relerr(got[0], want[0]) < 1e-15 and relerr(got[1], want[1]) < 1e-15


def test_metric_closed_form_on_flat_grid():
    """z = (k-1) h + z0: z_r = h, z_p = z_q = 0 everywhere, also on the
    one-sided boundary closures (exact for linear functions)."""
    nx, ny, nz, h = 9, 8, 10, 0.25
    box = grid_box(nx, ny, nz)
    ib, ie, jb, je, kb, ke = box
    k = np.arange(kb, ke + 1)
    z = np.broadcast_to(((k - 1) * h + 3.0)[:, None, None], (ke - kb + 1, je - jb + 1, ie - ib + 1))
    z = np.ascontiguousarray(z.reshape(-1))
    met, jac = metric_of(box, h, z, 1)
    n = npts(box)
    m = met.reshape(4, n)
    assert np.allclose(m[0], np.sqrt(h), rtol=1e-14) and np.allclose(m[3], np.sqrt(h), rtol=1e-14)
    assert np.max(np.abs(m[1])) < 1e-13 and np.max(np.abs(m[2])) < 1e-13
    assert np.allclose(jac, h ** 3, rtol=1e-14)


def test_metric_rejects_folded_grid():
    box, z = hills_depth(9, 9, 9, 0.1, 1)
    n = npts(box)
    ib, ie, jb, je, kb, ke = box
    z = z.reshape(ke - kb + 1, je - jb + 1, ie - ib + 1)[::-1].copy().reshape(-1)
    assert O.metric(1, box, 0.1, z, np.zeros(4 * n), np.zeros(n)) == -1


def curv_inputs(seed, nx, ny, nz, h=0.1):
    box, f = random_inputs(seed, nx, ny, nz)
    _, z = hills_depth(nx, ny, nz, h, seed + 100)
    met, jac = metric_of(box, h, z, 1)
    return box, f, met, jac


def run_curv(corder, box, f, met, jac, top, coefs, lib=None):
    acof, ghcof, bope, _ = coefs
    u = to_layout(f["u"], box, corder)
    lu = np.zeros_like(u)
    O.rhs4sgcurv(corder, box, u, f["mu"], f["la"], met_layout(met, box, corder), jac, lu, [0, 0, 0, 0, top, 0],
                 acof, bope, ghcof, f["strx"], f["stry"], lib=lib)
    return lu


@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("top", [0, 1])
def test_rhs4sgcurv_matches_reference(coefs, corder, top):
    lib = ref_or_skip()
    box, f, met, jac = curv_inputs(21 + top, 11, 10, 14)
    got = run_curv(corder, box, f, met, jac, top, coefs)
    want = run_curv(corder, box, f, met, jac, top, coefs, lib=lib)
    assert np.max(np.abs(want)) > 0
    assert relerr(got, want) < 1e-13


@pytest.mark.parametrize("top", [0, 1])
def test_rhs4sgcurv_matches_golden(coefs, top):
    meta = json.load(open(os.path.join(GOLD, "curv.json")))
    gold = np.load(os.path.join(GOLD, "curv.npz"))
    case = meta["cases"]["rhs4sgcurv_rev_top%d" % top]
    nx, ny, nz = meta["grid"]
    box, f, met, jac = curv_inputs(case["seed"], nx, ny, nz, h=meta["h"])
    assert abs(float(np.sum(met)) - case["met_sum"]) <= 1e-12 * abs(case["met_sum"])
    lu = run_curv(1, box, f, met, jac, top, coefs)
    assert relerr(lu, gold["rhs4sgcurv_rev_top%d" % top]) < 1e-13


def flat_metric(box, h):
    ib, ie, jb, je, kb, ke = box
    k = np.arange(kb, ke + 1)
    z = np.broadcast_to(((k - 1) * h)[:, None, None], (ke - kb + 1, je - jb + 1, ie - ib + 1))
    return metric_of(box, h, np.ascontiguousarray(z.reshape(-1)), 1)


@pytest.mark.parametrize("top", [0, 1])
def test_rhs4sgcurv_flat_equals_cartesian(coefs, top):
    """On z = (k-1) h the curvilinear operator is the Cartesian rhs4sg with
    strz = 1 (rows k = 1..6 closed with the same SBP operator)."""
    nx, ny, nz, h = 10, 11, 15, 0.2
    box, f = random_inputs(5, nx, ny, nz)
    met, jac = flat_metric(box, h)
    lu = run_curv(1, box, f, met, jac, top, coefs)
    acof, ghcof, bope, _ = coefs
    want = np.zeros_like(lu)
    O.rhs4sg(1, box, nz, [0, 0, 0, 0, top, 0], acof, bope, ghcof, want, f["u"], f["mu"], f["la"], h, f["strx"],
             f["stry"], np.ones(nz + 4))
    # rhs4sgcurv covers k up to klast-2 = nz; rhs4sg the same without a bottom closure
    assert relerr(lu, want) < 1e-12


def test_rhs4sgcurv_linear_and_kills_constants(coefs):
    box, f, met, jac = curv_inputs(8, 10, 9, 13)
    g = dict(f)
    g["u"] = np.full_like(f["u"], 0.7)
    assert np.max(np.abs(run_curv(1, box, g, met, jac, 1, coefs))) < 1e-10
    a, b = 1.7, -0.3
    g["u"] = a * f["u"] + b * f["up"]
    lhs = run_curv(1, box, g, met, jac, 1, coefs)
    g["u"] = f["up"]
    rhs = a * run_curv(1, box, f, met, jac, 1, coefs) + b * run_curv(1, box, g, met, jac, 1, coefs)
    assert relerr(lhs, rhs) < 1e-13


@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4c_reduces_to_cartesian(corder):
    """jac = 1: the curvilinear damping is the Cartesian addsgd4 without its
    z part (dcz = 0, coz = 1); a constant jac drops out."""
    nx, ny, nz = 9, 10, 8
    box, f = random_inputs(13, nx, ny, nz)
    n = npts(box)
    up0 = to_layout(f["up"], box, corder)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    got = up0.copy()
    O.addsgd4c(corder, box, got, u, um, f["rho"], f["dcx"], f["dcy"], f["strx"], f["stry"], np.ones(n), f["cox"],
               f["coy"], 0.02)
    want = up0.copy()
    O.addsgd4(corder, box, want, u, um, f["rho"], f["dcx"], f["dcy"], np.zeros(nz + 4), f["strx"], f["stry"],
              f["strz"], f["cox"], f["coy"], np.ones(nz + 4), 0.02)
    assert relerr(got - up0, want - up0) < 1e-13
    got3 = up0.copy()
    O.addsgd4c(corder, box, got3, u, um, f["rho"], f["dcx"], f["dcy"], f["strx"], f["stry"], np.full(n, 3.0),
               f["cox"], f["coy"], 0.02)
    assert relerr(got3 - up0, want - up0) < 1e-13


def test_freesurf_flat_equals_cartesian(coefs):
    """Flat top, zero traction: the curvilinear closure gives the Cartesian
    free-surface ghost values (bcfortsg, ew-cfromfort.C:415-474)."""
    nx, ny, nz, h = 9, 10, 9, 0.2
    box, f = random_inputs(17, nx, ny, nz)
    met, _ = flat_metric(box, h)
    sbop = coefs[3]
    ni, nj = nx + 4, ny + 4
    got = f["u"].copy()
    O.freesurfcurvisg(1, box, nz, 5, got, f["mu"], f["la"], met, sbop, np.zeros(3 * ni * nj), f["strx"], f["stry"])
    want = f["u"].copy()
    wind = np.zeros(36, dtype=np.int32)  # only the free-surface side acts
    bforce = [None] * 6
    bforce[4] = np.zeros(3 * ni * nj)
    O.bcfortsg(1, box, wind, nx, ny, nz, want, h, [5, 5, 5, 5, 0, 5], sbop, f["mu"], f["la"], bforce, f["strx"],
               f["stry"])
    assert relerr(got, want) < 1e-13


@pytest.mark.parametrize("side", [5, 6])
def test_freesurf_curved_leaves_given_traction(coefs, side):
    """After the closure, N (sum_w s_w u_w) + g = 0 on the boundary plane,
    with N = mu ac I + (mu + lambda) v v^T the normal-stress matrix and g the
    tangential traction minus the forcing (the closure solves with N^-1)."""
    nx, ny, nz, h = 12, 11, 10, 0.1
    box, f, met, _ = curv_inputs(31, nx, ny, nz, h=h)
    sbop = coefs[3]
    ni, nj = nx + 4, ny + 4
    forcing = np.random.default_rng(4).uniform(-1, 1, 3 * ni * nj)
    u = f["u"].copy()
    O.freesurfcurvisg(1, box, nz, side, u, f["mu"], f["la"], met, sbop, forcing, f["strx"], f["stry"])
    n = npts(box)
    ib, ie, jb, je, kb, ke = box
    shp = (ke - kb + 1, je - jb + 1, ie - ib + 1)
    U = u.reshape(3, *shp)
    M = met.reshape(4, *shp)
    mu, la = f["mu"].reshape(shp), f["la"].reshape(shp)
    k = 1 if side == 5 else nz
    kl = 1 if side == 5 else -1
    kk = k - kb
    c1, c2 = 2.0 / 3, -1.0 / 12
    sl = np.s_[2:-2]
    sx = f["strx"][2:-2][None, :]
    sy = f["stry"][2:-2][:, None]

    def dp(c):
        a = U[c, kk]
        return c2 * (a[sl, 4:] - a[sl, :-4]) + c1 * (a[sl, 3:-1] - a[sl, 1:-3])

    def dq(c):
        a = U[c, kk]
        return c2 * (a[4:, sl] - a[:-4, sl]) + c1 * (a[3:-1, sl] - a[1:-3, sl])

    m1, m2, m3, m4 = (M[c, kk][sl, sl] for c in range(4))
    m0, l0 = mu[kk][sl, sl], la[kk][sl, sl]
    K = 2 * m0 + l0
    # tangential traction (pr and qr parts of the stress on the plane)
    g = [K * m2 * m1 * dp(0) * sx / sy + m0 * m3 * m1 * dp(1) + m0 * m4 * m1 * dp(2) / sy
         + m0 * m3 * m1 * dq(0) * sy / sx + l0 * m2 * m1 * dq(1),
         l0 * m3 * m1 * dp(0) + m0 * m2 * m1 * dp(1) * sx / sy + m0 * m2 * m1 * dq(0)
         + K * m3 * m1 * dq(1) * sy / sx + m0 * m4 * m1 * dq(2) / sx,
         l0 * m4 * m1 * dp(0) / sy + m0 * m2 * m1 * dp(2) * sx / sy + m0 * m3 * m1 * dq(2) * sy / sx
         + l0 * m4 * m1 * dq(1) / sx]
    F = forcing.reshape(nj, ni, 3)
    g = [g[c] - F[sl, sl, c] for c in range(3)]
    v = [np.sqrt(sx / sy) * m2, np.sqrt(sy / sx) * m3, m4 / np.sqrt(sx * sy)]
    ac = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
    du = [sum(sbop[w] * U[c, kk + kl * (w - 1)][sl, sl] for w in range(5)) for c in range(3)]
    vdu = v[0] * du[0] + v[1] * du[1] + v[2] * du[2]
    res = [m0 * ac * du[c] + (m0 + l0) * v[c] * vdu + g[c] for c in range(3)]
    scale = max(np.max(np.abs(gc)) for gc in g)
    assert max(np.max(np.abs(r)) for r in res) < 1e-12 * scale


@pytest.mark.parametrize("corder", [0, 1])
def test_cart_topo_interface_places_physical_depths(corder):
    """enforce_cart_topo (EW.C:3503-3531) on a field that is one linear
    function of physical depth on both grids: the curvilinear grid's last
    interior plane k = nzc and the Cartesian plane k = 1 are the same depth,
    and the spacing is h across the interface.  After the coupling every
    ghost plane of both grids holds the field at its own depth, whatever it
    held before: the plane offsets connect points at the same depth."""
    from oracle.coupling import enforce_cart_topo
    ni, nj, nzc, nzk, h, zint = 7, 6, 9, 8, 0.25, 3.0
    ib, ie, jb, je = -1, ni - 2, -1, nj - 2
    kc = np.arange(-1, nzc + 3)           # curvilinear planes, nzc at zint
    kk = np.arange(-1, nzk + 3)           # Cartesian planes, 1 at zint
    zc = zint + (kc - nzc) * h
    zk = zint + (kk - 1) * h
    rng = np.random.default_rng(5)
    a, b = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)

    def field(z):
        f = a[:, None] * z[None, :] + b[:, None]  # [c, k]
        return np.broadcast_to(f[:, :, None, None], (3, z.size, nj, ni)).copy()

    def pack(f):
        return f.ravel().copy() if corder == 1 else f.transpose(1, 2, 3, 0).ravel().copy()

    fc, fk = field(zc), field(zk)
    gc, gk = fc.copy(), fk.copy()
    gc[:, -3:] = rng.uniform(-9, 9, gc[:, -3:].shape)   # curvilinear ghosts and common plane: garbage
    gk[:, :2] = rng.uniform(-9, 9, gk[:, :2].shape)     # Cartesian ghosts above the interface: garbage
    uc, uk = pack(gc), pack(gk)
    enforce_cart_topo(corder, (ib, ie, jb, je), (-1, nzk + 2), uk, (-1, nzc + 2), uc, 2)
    assert np.allclose(uc, pack(fc), rtol=0, atol=1e-13)
    assert np.allclose(uk, pack(fk), rtol=0, atol=1e-13)
