"""GPU parity of the curvilinear-grid entry points (sw4hip_metric,
sw4hip_rhs4sgcurv, sw4hip_addsgd4c, sw4hip_freesurfcurvisg) against the CPU
oracle on the same seeded inputs: a grid under Gaussian hills, random
displacements and material, both Sarray layouts.  Bar: max|gpu - oracle| /
max|oracle| <= 1e-12 (FP64)."""
import numpy as np
import pytest

from common import hills_depth, npts, random_inputs, relerr, to_layout
from oracle import oracle as O
from sw4lite_amd import lib as L

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    L.load()
    return _t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch as _t
    _t.cuda.synchronize()
    return t.cpu().numpy()


def met_layout(m, box, corder):
    if corder == 1:
        return m
    return np.ascontiguousarray(m.reshape(4, npts(box)).T.reshape(-1))


def coords(box, h):
    ib, ie, jb, je, kb, ke = box
    shp = (ke - kb + 1, je - jb + 1, ie - ib + 1)
    x = np.broadcast_to((np.arange(ib, ie + 1) - 1) * h, shp).copy().reshape(-1)
    y = np.broadcast_to(((np.arange(jb, je + 1) - 1) * h)[:, None], shp).copy().reshape(-1)
    return x, y


def curved(seed, nx, ny, nz, h):
    box, f = random_inputs(seed, nx, ny, nz)
    _, z = hills_depth(nx, ny, nz, h, seed + 100)
    n = npts(box)
    met, jac = np.zeros(4 * n), np.zeros(n)
    assert O.metric(1, box, h, z, met, jac) == 0
    return box, f, z, met, jac


@pytest.mark.parametrize("corder", [0, 1])
def test_metric(torch, corder):
    from sw4lite_amd import ew
    h = 0.1
    box, f, z, met, jac = curved(60, 33, 21, 17, h)
    x, y = coords(box, h)
    n = npts(box)
    gm, gj = dev(torch, np.zeros(4 * n)), dev(torch, np.zeros(n))
    ew.metric(corder, box, dev(torch, x), dev(torch, y), dev(torch, z), gm, gj)
    e1, e2 = relerr(host(gm), met_layout(met, box, corder)), relerr(host(gj), jac)
    print("metric corder", corder, "relerr", e1, e2, flush=True)
    assert e1 < TOL and e2 < TOL


def test_metric_rejects_folded_grid(torch):
    from sw4lite_amd import ew
    h = 0.1
    box, f, z, met, jac = curved(61, 12, 12, 12, h)
    ib, ie, jb, je, kb, ke = box
    z = z.reshape(ke - kb + 1, je - jb + 1, ie - ib + 1)[::-1].copy().reshape(-1)
    x, y = coords(box, h)
    n = npts(box)
    with pytest.raises(L.Sw4HipError):
        ew.metric(1, box, dev(torch, x), dev(torch, y), dev(torch, z), dev(torch, np.zeros(4 * n)),
                  dev(torch, np.zeros(n)))


@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("top", [0, 1])
@pytest.mark.parametrize("shape", [(37, 21, 19), (38, 21, 40)])
def test_rhs4sgcurv(torch, coefs, corder, top, shape):
    """Against the oracle on an odd and an even row length, with and without
    the top SBP rows."""
    from sw4lite_amd import ew
    acof, ghcof, bope, _ = coefs
    box, f, z, met, jac = curved(70 + top, *shape, 0.1)
    u = to_layout(f["u"], box, corder)
    m = met_layout(met, box, corder)
    os_ = [0, 0, 0, 0, top, 0]
    want = np.zeros_like(u)
    O.rhs4sgcurv(corder, box, u, f["mu"], f["la"], m, jac, want, os_, acof, bope, ghcof, f["strx"], f["stry"])
    lu = dev(torch, np.zeros_like(u))
    ew.rhs4sgcurv(corder, box, dev(torch, u), dev(torch, f["mu"]), dev(torch, f["la"]), dev(torch, m),
                  dev(torch, jac), lu, os_, acof, bope, ghcof, dev(torch, f["strx"]), dev(torch, f["stry"]))
    err = relerr(host(lu), want)
    print("rhs4sgcurv corder", corder, "top", top, "shape", shape, "relerr", err, flush=True)
    assert err < TOL


@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4c(torch, corder):
    from sw4lite_amd import ew
    box, f, z, met, jac = curved(80, 29, 23, 13, 0.1)
    up0 = to_layout(f["up"], box, corder)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    want = up0.copy()
    O.addsgd4c(corder, box, want, u, um, f["rho"], f["dcx"], f["dcy"], f["strx"], f["stry"], jac, f["cox"],
               f["coy"], 0.02)
    up = dev(torch, up0)
    ew.addsgd4c(corder, box, up, dev(torch, u), dev(torch, um), dev(torch, f["rho"]), dev(torch, f["dcx"]),
                dev(torch, f["dcy"]), dev(torch, f["strx"]), dev(torch, f["stry"]), dev(torch, jac),
                dev(torch, f["cox"]), dev(torch, f["coy"]), 0.02)
    err = relerr(host(up) - up0, want - up0)
    print("addsgd4c corder", corder, "relerr", err, flush=True)
    assert err < TOL


@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("side,forced", [(5, False), (5, True), (6, True)])
def test_freesurfcurvisg(torch, coefs, corder, side, forced):
    from sw4lite_amd import ew
    nx, ny, nz = 31, 19, 14
    box, f, z, met, jac = curved(90, nx, ny, nz, 0.1)
    sbop = coefs[3]
    u0 = to_layout(f["u"], box, corder)
    m = met_layout(met, box, corder)
    forcing = np.random.default_rng(5).uniform(-1, 1, 3 * (nx + 4) * (ny + 4)) if forced else None
    want = u0.copy()
    O.freesurfcurvisg(corder, box, nz, side, want, f["mu"], f["la"], m, sbop,
                      forcing if forced else np.zeros(3 * (nx + 4) * (ny + 4)), f["strx"], f["stry"])
    u = dev(torch, u0)
    ew.freesurfcurvisg(corder, box, nz, side, u, dev(torch, f["mu"]), dev(torch, f["la"]), dev(torch, m), sbop,
                       dev(torch, forcing) if forced else None, dev(torch, f["strx"]), dev(torch, f["stry"]))
    got = host(u)
    err = relerr(got, want)
    print("freesurfcurvisg corder", corder, "side", side, "forced", forced, "relerr", err, flush=True)
    assert err < TOL
    assert np.array_equal(got != u0, want != u0)


@pytest.mark.parametrize("corder,bc", [(1, [2, 2, 2, 2, 0, 2]), (0, [2, 2, 2, 2, 0, 1]), (1, [1, 2, 2, 1, 2, 2])])
def test_curvilinear_time_loop(torch, corder, bc):
    """The EW time loop on a grid under Gaussian hills (reference stage order
    with rhs4sgcurv, addsgd4cfort and the curved free-surface closure) against
    the oracle's loop: same dt, sources, and displacements after 8 steps."""
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    nx, ny, nz, h = 33, 29, 24, 0.05
    _, z = hills_depth(nx, ny, nz, h, 12)
    box, f = random_inputs(13, nx, ny, nz)
    rho = 1.5 + 0.2 * f["rho"]
    mu, la = 1.0 + 0.1 * f["mu"], 1.2 + 0.1 * f["la"]
    src = [dict(x=0.81, y=0.74, z=0.52, m=[0.3, 0.8, -0.45, -0.6, 0.25, 0.9], t0=0.12, freq=18.0,
                tfkind=L.TF_GAUSSIAN)]
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, src, corder=corder, sg_gp=6, nsteps=8, zgrid=z)
    e = EW(nx, ny, nz, h, bc, corder=corder, sg_gp_thickness=6)
    e.set_material(rho, mu, la)
    e.set_topography(z)
    for s in src:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    e.add_receiver(0.8, 0.7, 0.0, 0)
    e.setup(nsteps=8)
    inf = e.info()
    assert abs(inf.dt - ref.dt) < 1e-13 * ref.dt
    e.timesteploop(8, mode=1)
    ref.step(8)
    u, um = e.get_solution()
    rec = e.receivers()
    e.close()
    err = relerr(u, ref.U)
    print("curvilinear loop corder", corder, "bc", bc, "relerr", err, "dt", inf.dt, flush=True)
    assert np.max(np.abs(ref.U)) > 0
    assert err < TOL and relerr(um, ref.Um) < TOL
    assert rec.shape == (9, 1, 3) and np.max(np.abs(rec[-1])) > 0


@pytest.mark.parametrize("corder", [0, 1])
def test_enforce_cart_topo_matches_oracle(corder):
    """sw4hip_enforce_cart_topo against the numpy restatement of
    EW::enforceCartTopo on random fields: a copy, so bit for bit."""
    import torch
    from oracle.coupling import enforce_cart_topo
    from sw4lite_amd import ew
    ni, nj, nzc, nzk = 37, 21, 15, 12
    ij = (-1, ni - 2, -1, nj - 2)
    rng = np.random.default_rng(17 + corder)
    uc = rng.standard_normal(3 * ni * nj * (nzc + 4))
    uk = rng.standard_normal(3 * ni * nj * (nzk + 4))
    want_c, want_k = uc.copy(), uk.copy()
    enforce_cart_topo(corder, ij, (-1, nzk + 2), want_k, (-1, nzc + 2), want_c, 2)
    dc, dk = torch.from_numpy(uc).cuda(), torch.from_numpy(uk).cuda()
    ew.enforce_cart_topo(corder, ij, (-1, nzk + 2), dk, (-1, nzc + 2), dc, 2)
    torch.cuda.synchronize()
    assert np.array_equal(dc.cpu().numpy(), want_c) and np.array_equal(dk.cpu().numpy(), want_k)
