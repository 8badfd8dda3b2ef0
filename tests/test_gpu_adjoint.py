"""GPU: the discrete adjoint of the FP64 step.

Kernel level, exact identities on seeded random inputs (the transposes must
satisfy them to rounding, ~1e-15 relative):
  <z, L u> = <L^T z, u>                  rhs4sg, every one-sided combination
  <z, L(u; dmu, 0)> = <G_mu(z, u), dmu>   L is linear in mu and lambda
  <y, -D(d)> = <-D^T y, d>                supergrid damping
  <y, BC(u)> = <BC^T y, u>                ghost-point closure, zero data
and central differences for the nonlinear material dependences (damping in
rho, free-surface closure in mu, lambda).  End to end: the gradient of a
receiver misfit against central finite differences of the FP64 forward
solver along random directions; the bar is 1e-4 relative (the agreement the
adjoint is specified to), the printed values are what the run measured."""
import numpy as np
import pytest

from common import random_inputs, relerr
from sw4lite_amd import lib as L
from test_gpu_parity import SOURCES, layered_material

pytestmark = pytest.mark.gpu
DOT_TOL = 1e-13
FD_TOL = 1e-4
# FP32 form: central differences of the FP32 run's misfit with a relative
# step of 1e-3 carry its rounding (about 1e-7 of the misfit) over a change of
# about 1e-3 of it, and the float material's rounding: 1e-4 of the slope
FD32_TOL = 2e-3


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    L.load()
    return _t


@pytest.fixture(scope="module")
def coefs():
    from sw4lite_amd import ew
    return ew.stencil_coefficients()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch as _t
    _t.cuda.synchronize()
    return t.cpu().numpy()


def interior_mask(box):
    ib, ie, jb, je, kb, ke = box
    m = np.zeros((ke - kb + 1, je - jb + 1, ie - ib + 1))
    m[2:-2, 2:-2, 2:-2] = 1.0
    return np.tile(m.ravel(), 3)


def dot_rel(a, b):
    """|a - b| relative to the size of the terms."""
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


# the tall shapes span several k-blocks of the LDS k-march kernels (32 or 64
# planes per workgroup) with a partial last block; even nx (even row length)
# takes the transposed sweep that loads its planes asynchronously into LDS
@pytest.mark.parametrize("os45,shape", [((0, 0), (21, 18, 17)), ((1, 0), (21, 18, 17)), ((0, 1), (21, 18, 17)),
                                        ((1, 1), (21, 18, 17)), ((1, 0), (37, 11, 75)), ((1, 0), (22, 18, 17)),
                                        ((1, 1), (36, 19, 17)), ((0, 1), (40, 13, 75))])
def test_rhs4sg_transpose(torch, coefs, os45, shape):
    from sw4lite_amd import ew
    acof, ghcof, bope, _ = coefs
    nx, ny, nz = shape
    box, f = random_inputs(300, nx, ny, nz)
    rng = np.random.default_rng(301)
    z = rng.uniform(-1, 1, f["u"].size)
    zr = z * interior_mask(box)
    os_ = [0, 0, 0, 0, *os45]
    d = {k: dev(torch, f[k]) for k in ("mu", "la", "strx", "stry", "strz")}
    lu = dev(torch, np.zeros_like(f["u"]))
    ew.rhs4sg(1, box, nz, os_, acof, bope, ghcof, lu, dev(torch, f["u"]), d["mu"], d["la"], 0.3, d["strx"],
              d["stry"], d["strz"])
    lt = dev(torch, np.full_like(f["u"], 7.0))  # overwritten on every point
    ew.rhs4sg_adjoint(box, nz, os_, acof, bope, ghcof, lt, dev(torch, z), d["mu"], d["la"], 0.3, d["strx"], d["stry"],
                      d["strz"])
    a = float(np.dot(zr, host(lu)))
    b = float(np.dot(host(lt), f["u"]))
    print("rhs4sg transpose onesided", os45, "<z,Lu>", a, "<L^T z,u>", b, "rel", dot_rel(a, b), flush=True)
    assert dot_rel(a, b) < DOT_TOL
    # L^T of a single row entry is a column of L: compare to probing L
    assert np.all(np.isfinite(host(lt)))


@pytest.mark.parametrize("os45,shape", [((0, 0), (19, 17, 16)), ((1, 0), (19, 17, 16)), ((1, 1), (19, 17, 16)),
                                        ((1, 1), (35, 9, 69))])
def test_rhs4sg_material_derivative(torch, coefs, os45, shape):
    from sw4lite_amd import ew
    acof, ghcof, bope, _ = coefs
    nx, ny, nz = shape
    box, f = random_inputs(310, nx, ny, nz)
    rng = np.random.default_rng(311)
    z = rng.uniform(-1, 1, f["u"].size)
    zr = z * interior_mask(box)
    dm, dl = rng.uniform(-1, 1, f["mu"].size), rng.uniform(-1, 1, f["la"].size)
    zero = np.zeros_like(f["mu"])
    os_ = [0, 0, 0, 0, *os45]
    s = {k: dev(torch, f[k]) for k in ("strx", "stry", "strz")}
    u = dev(torch, f["u"])

    def apply(mu, la):
        lu = dev(torch, np.zeros_like(f["u"]))
        ew.rhs4sg(1, box, nz, os_, acof, bope, ghcof, lu, u, dev(torch, mu), dev(torch, la), 0.3, s["strx"],
                  s["stry"], s["strz"])
        return float(np.dot(zr, host(lu)))

    gmu, gla = dev(torch, np.zeros_like(zero)), dev(torch, np.zeros_like(zero))
    ew.rhs4sg_matgrad(box, nz, os_, acof, bope, ghcof, dev(torch, z), u, dev(torch, f["mu"]), dev(torch, f["la"]),
                      0.3, s["strx"], s["stry"], s["strz"], gmu, gla)
    gm, gl = host(gmu), host(gla)
    a_m, b_m = apply(dm, zero), float(np.dot(gm, dm))
    a_l, b_l = apply(zero, dl), float(np.dot(gl, dl))
    print("rhs4sg material derivative onesided", os45, "mu", dot_rel(a_m, b_m), "lambda", dot_rel(a_l, b_l), flush=True)
    assert dot_rel(a_m, b_m) < DOT_TOL and dot_rel(a_l, b_l) < DOT_TOL


@pytest.mark.parametrize("shape", [(23, 17, 19), (37, 11, 70)])
def test_addsgd4_transpose(torch, shape):
    from sw4lite_amd import ew
    nx, ny, nz = shape
    box, f = random_inputs(320, nx, ny, nz)
    rng = np.random.default_rng(321)
    y = rng.uniform(-1, 1, f["u"].size)
    one = [f[k] for k in ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz")]
    done = [dev(torch, a) for a in one]
    u, um = dev(torch, f["u"]), dev(torch, f["um"])

    def damp(rho):
        up = dev(torch, np.zeros_like(f["u"]))
        ew.addsgd4(1, box, up, u, um, dev(torch, rho), *done, 0.02)
        return -host(up)  # D(u - um)

    dbar = dev(torch, np.full_like(f["u"], 7.0))  # overwritten on every point
    grho = dev(torch, np.zeros_like(f["rho"]))
    ew.addsgd4_adjoint(box, dev(torch, y), u, um, dev(torch, f["rho"]), *done, 0.02, dbar, grho)
    d = f["u"] - f["um"]
    a = float(np.dot(y, -damp(f["rho"])))
    b = float(np.dot(host(dbar), d))
    print("addsgd4 transpose", a, b, dot_rel(a, b), flush=True)
    assert dot_rel(a, b) < DOT_TOL
    # rho: central differences of -y.D(d; rho) along a random direction
    drho = rng.uniform(-1, 1, f["rho"].size) * f["rho"]
    eps = 1e-6
    fd = (np.dot(y, -damp(f["rho"] + eps * drho)) - np.dot(y, -damp(f["rho"] - eps * drho))) / (2 * eps)
    g = float(np.dot(host(grho), drho))
    print("addsgd4 rho derivative", fd, g, dot_rel(fd, g), flush=True)
    assert dot_rel(fd, g) < 1e-7


@pytest.mark.parametrize("bc", [[2, 2, 2, 2, 0, 2], [1, 1, 2, 2, 2, 0], [2, 2, 2, 2, 0, 0], [1, 1, 1, 1, 1, 1]])
def test_bcfortsg_transpose(torch, coefs, bc):
    from sw4lite_amd import ew
    nx, ny, nz = 16, 15, 14
    h = 0.1
    box, f = random_inputs(330, nx, ny, nz)
    sbop = coefs[3]
    desc = L.grid_desc(nx, ny, nz, h, bc)
    wind, _ = ew.host_boundary_windows(desc)
    rng = np.random.default_rng(331)
    y = rng.uniform(-1, 1, f["u"].size)
    sx, sy = dev(torch, f["strx"]), dev(torch, f["stry"])

    def bcf(mu, la):
        v = dev(torch, f["u"])
        ew.bcfortsg(1, box, wind, nx, ny, nz, v, h, bc, sbop, dev(torch, mu), dev(torch, la), [None] * 6, sx, sy)
        return host(v)

    fwd = bcf(f["mu"], f["la"])
    yt = dev(torch, y)
    gmu, gla = dev(torch, np.zeros_like(f["mu"])), dev(torch, np.zeros_like(f["la"]))
    ew.bcfortsg_adjoint(box, wind, nx, ny, nz, yt, dev(torch, fwd), h, bc, sbop, dev(torch, f["mu"]),
                        dev(torch, f["la"]), sx, sy, gmu, gla)
    a = float(np.dot(y, fwd))
    b = float(np.dot(host(yt), f["u"]))
    print("bcfortsg transpose bc", bc, a, b, dot_rel(a, b), flush=True)
    assert dot_rel(a, b) < DOT_TOL
    if 0 in bc:
        eps = 1e-6
        for name, g in (("mu", host(gmu)), ("la", host(gla))):
            dm = rng.uniform(-1, 1, f["mu"].size) * f[name]
            pm = {k: f[k].copy() for k in ("mu", "la")}
            mm = {k: f[k].copy() for k in ("mu", "la")}
            pm[name] += eps * dm
            mm[name] -= eps * dm
            fd = (np.dot(y, bcf(pm["mu"], pm["la"])) - np.dot(y, bcf(mm["mu"], mm["la"]))) / (2 * eps)
            gg = float(np.dot(g, dm))
            print("bcfortsg", name, "derivative", fd, gg, dot_rel(fd, gg), flush=True)
            assert dot_rel(fd, gg) < 1e-6


RECS = [(1.0, 1.05, 0.0, 0), (1.53, 0.77, 0.41, 1), (0.3, 1.6, 1.1, 0), (1.2, 1.1, 0.0, 1), (0.7, 0.6, 0.9, 0)]


def make_ew(nx, ny, nz, h, bc, nsteps, seed=5):
    from sw4lite_amd import EW
    rho, mu, la = layered_material(nx, ny, nz, h, seed)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=5)
    e.set_material(rho, mu, la)
    for s in SOURCES[:2]:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    for x, y, z, m in RECS:
        e.add_receiver(x, y, z, m)
    e.setup(nsteps=nsteps)
    return e, (rho, mu, la)


def misfit(e, mat, nsteps, dobs, mode=0):
    e.set_material(*mat)
    e.reset()
    e.timesteploop(nsteps, mode=mode)
    tr = e.receivers()[: nsteps + 1]
    return 0.5 * float(np.sum((tr - dobs) ** 2)), tr


# nx = 25: rows of odd length (the four-byte asynchronous loads, the stored
# acceleration, the transposed sweep on the odd padded rows)
@pytest.mark.parametrize("bc,nx", [([2, 2, 2, 2, 0, 2], 24), ([2, 2, 2, 2, 2, 2], 24), ([1, 2, 2, 1, 0, 0], 24),
                                   ([2, 2, 2, 2, 0, 2], 25)])
def test_gradient_matches_finite_differences(torch, bc, nx):
    ny, nz, h, nsteps = 22, 20, 0.1, 30
    e, mat = make_ew(nx, ny, nz, h, bc, nsteps)
    rng = np.random.default_rng(41)
    J0, tr = misfit(e, mat, nsteps, 0.0)
    dobs = 0.7 * tr + 0.1 * np.max(np.abs(tr)) * rng.standard_normal(tr.shape)
    J0, tr = misfit(e, mat, nsteps, dobs)
    g = e.gradient(tr - dobs, nsteps, ckpt=7)
    eps = 1e-6
    worst = 0.0
    for which in (0, 1, 2, None):
        d = [np.zeros_like(m) for m in mat]
        for q in range(3):
            if which is None or which == q:
                d[q] = rng.uniform(-1, 1, mat[q].size) * mat[q]
        jp, _ = misfit(e, [m + eps * dd for m, dd in zip(mat, d)], nsteps, dobs)
        jm, _ = misfit(e, [m - eps * dd for m, dd in zip(mat, d)], nsteps, dobs)
        fd = (jp - jm) / (2 * eps)
        ad = sum(float(np.dot(gq, dq)) for gq, dq in zip(g, d))
        err = dot_rel(fd, ad)
        worst = max(worst, err)
        print("gradient bc", bc, "direction", which, "fd", fd, "adjoint", ad, "rel", err, flush=True)
    e.close()
    assert worst < FD_TOL


def test_gradient_independent_of_checkpoints(torch):
    nx, ny, nz, h, nsteps = 20, 18, 16, 0.1, 17
    e, mat = make_ew(nx, ny, nz, h, [2, 2, 2, 2, 0, 2], nsteps)
    rng = np.random.default_rng(43)
    gtr = rng.standard_normal((nsteps + 1, len(RECS), 3))
    # 0: chosen from free device memory, the whole history for this grid
    assert e.gradient_plan(nsteps, 0)[0] == nsteps
    assert e.gradient_plan(nsteps, 5)[0] == 5 and e.gradient_plan(nsteps, 99)[0] == nsteps
    assert e.gradient_plan(nsteps, 5)[1] < e.gradient_plan(nsteps, 0)[1]
    res = [e.gradient(gtr, nsteps, ckpt=c) for c in (1, 5, 17, 0)]
    again = e.gradient(gtr, nsteps, ckpt=0)  # reuses the kept workspace
    e.release_workspace()
    after = e.gradient(gtr, nsteps, ckpt=0)  # allocates it again
    e.close()
    for r in res[1:]:
        for a, b in zip(r, res[0]):
            assert relerr(a, b) < 1e-12
    for a, b, c in zip(again, after, res[3]):
        assert np.array_equal(a, c) and np.array_equal(b, c)


@pytest.mark.parametrize("bc", [[2, 2, 2, 2, 0, 2], [1, 1, 2, 2, 2, 0]])
def test_gradient_stored_acceleration_bitwise(torch, bc):
    """Kernel variant 8 lets the recomputed predictor of every reverse step
    store its acceleration (sweep, ghost shell, then the planes outside the
    sweep, the side frame where the boundary conditions change it and the
    source points) instead of a separate pass over the whole box: the
    gradients must be variant 6's to the last bit."""
    nx, ny, nz, h, nsteps = 20, 18, 16, 0.1, 12
    rng = np.random.default_rng(47)
    gtr = rng.standard_normal((nsteps + 1, len(RECS), 3))
    res = []
    for variant in (6, 8):
        e, mat = make_ew(nx, ny, nz, h, bc, nsteps)
        e.set_kernel_variant(variant, 7)
        res.append(e.gradient(gtr, nsteps, ckpt=4))
        e.close()
    ndiff = [int(np.count_nonzero(a != b)) for a, b in zip(*res)]
    print("gradient, stored acceleration: bc", bc, "points differing from variant 6 (rho, mu, lambda):", ndiff,
          flush=True)
    assert np.max(np.abs(res[0][0])) > 0 and ndiff == [0, 0, 0]


@pytest.mark.parametrize("variant,bc,nx", [(9, [2, 2, 2, 2, 0, 2], 20), (9, [1, 1, 2, 2, 2, 0], 20),
                                           (6, [2, 2, 2, 2, 0, 2], 20), (10, [2, 2, 2, 2, 0, 2], 21)])
def test_gradient_kept_fields_bitwise(torch, variant, bc, nx):
    """Kept from the recomputation of each segment, the predicted field and
    the acceleration of every step spare the reverse steps their own
    predictor (variant 6 keeps the field only and forms the acceleration in
    the reverse step): the gradients must be those of the recomputing
    reverse steps to the last bit, for an interval that leaves three
    recomputed segments and for the whole history (nx = 21: odd rows)."""
    ny, nz, h, nsteps = 18, 16, 0.1, 13
    rng = np.random.default_rng(53)
    gtr = rng.standard_normal((nsteps + 1, len(RECS), 3))
    e, mat = make_ew(nx, ny, nz, h, bc, nsteps)
    e.set_kernel_variant(variant, 7)
    for ckpt in (4, 0):
        res = []
        for store in (0, 1):
            e.set_gradient_store(store)
            res.append(e.gradient(gtr, nsteps, ckpt=ckpt))
        ndiff = [int(np.count_nonzero(a != b)) for a, b in zip(*res)]
        print("gradient, kept fields: variant", variant, "bc", bc, "ckpt", ckpt, "points differing:", ndiff,
              flush=True)
        assert np.max(np.abs(res[0][0])) > 0 and ndiff == [0, 0, 0]
    # the kept fields' memory: two states per step of the interval
    e.set_gradient_store(0)
    b0 = e.gradient_plan(nsteps, 4)[1]
    e.set_gradient_store(1)
    b1 = e.gradient_plan(nsteps, 4)[1]
    e.close()
    assert b1 - b0 == pytest.approx(2 * 4 * 3 * 8 * (nx + 4) * (ny + 4) * (nz + 4))


def test_autograd_function(torch):
    """torch.autograd through sw4lite_amd.adjoint.seismograms: the backward
    pass equals the C-ABI gradient, and a gradient step lowers the misfit."""
    from sw4lite_amd.adjoint import seismograms
    nx, ny, nz, h, nsteps = 20, 18, 16, 0.1, 20
    e, mat = make_ew(nx, ny, nz, h, [2, 2, 2, 2, 0, 2], nsteps)
    rho, mu, la = (torch.tensor(m, device="cuda", requires_grad=True) for m in mat)
    with torch.no_grad():
        dobs = 0.8 * seismograms(rho, mu, la, e, nsteps)
    tr = seismograms(rho, mu, la, e, nsteps)
    J = 0.5 * ((tr - dobs) ** 2).sum()
    J.backward()
    g = e.gradient((tr - dobs).detach().cpu().numpy(), nsteps)
    for t, a in zip((rho, mu, la), g):
        assert relerr(t.grad.cpu().numpy(), a) < 1e-12
    # a small step along -grad (relative to the material) lowers the misfit
    with torch.no_grad():
        gn = [t.grad / t.grad.abs().max() for t in (rho, mu, la)]
        step = [m - 1e-4 * t.abs().max() * gg for m, t, gg in zip((rho, mu, la), (rho, mu, la), gn)]
        J1 = 0.5 * ((seismograms(*step, e, nsteps) - dobs) ** 2).sum()
    print("autograd misfit", float(J), "after a descent step", float(J1), flush=True)
    assert float(J1) < float(J)
    e.close()


@pytest.mark.parametrize("max_concurrent", [1, 3])
def test_batch_gradient_matches_single_shots(torch, max_concurrent):
    """sw4hip_batch_gradient over three shots of uneven grid size, run one
    after another or concurrently on streams of their own: every shot's
    gradients must be its single-shot gradients to the last bit, and the
    smallest shot's must match central finite differences of its misfit."""
    from sw4lite_amd import Batch
    nsteps = 14
    grids = [(24, 22, 20), (20, 18, 16), (28, 20, 18)]
    rng = np.random.default_rng(53)
    shots, mats = [], []
    for nx, ny, nz in grids:
        e, mat = make_ew(nx, ny, nz, 0.1, [2, 2, 2, 2, 0, 2], nsteps)
        shots.append(e)
        mats.append(mat)
    gtrs = [rng.standard_normal((nsteps + 1, len(RECS), 3)) for _ in shots]
    single = [e.gradient(g, nsteps, ckpt=4) for e, g in zip(shots, gtrs)]
    for e in shots:
        e.release_workspace()
    b = Batch(shots)
    got = b.gradient(gtrs, nsteps, ckpt=4, max_concurrent=max_concurrent)
    for q, (a, w) in enumerate(zip(got, single)):
        ndiff = [int(np.count_nonzero(x != y)) for x, y in zip(a, w)]
        print("batch gradient, max_concurrent", max_concurrent, "shot", q, grids[q],
              "points differing from the single shot:", ndiff, flush=True)
        assert np.max(np.abs(w[0])) > 0 and ndiff == [0, 0, 0]
    # again from the work memory the first call left in the batch's pool
    # (the largest shots' buffers serve the smaller ones), then after freeing it
    again = b.gradient(gtrs, nsteps, ckpt=4, max_concurrent=max_concurrent)
    b.release_workspace()
    after = b.gradient(gtrs, nsteps, ckpt=4, max_concurrent=max_concurrent)
    for a, c, w in zip(again, after, single):
        assert all(np.array_equal(x, y) and np.array_equal(z, y) for x, z, y in zip(a, c, w))
    # finite differences on the smallest shot, misfit against a perturbed record
    e, mat = shots[1], mats[1]
    J0, tr = misfit(e, mat, nsteps, 0.0)
    dobs = 0.7 * tr
    J0, tr = misfit(e, mat, nsteps, dobs)
    gb = b.gradient([np.zeros((nsteps + 1, len(RECS), 3)), tr - dobs, np.zeros((nsteps + 1, len(RECS), 3))],
                    nsteps, ckpt=5, max_concurrent=max_concurrent)[1]
    d = [rng.uniform(-1, 1, m.size) * m for m in mat]
    eps = 1e-6
    jp, _ = misfit(e, [m + eps * dd for m, dd in zip(mat, d)], nsteps, dobs)
    jm, _ = misfit(e, [m - eps * dd for m, dd in zip(mat, d)], nsteps, dobs)
    fd = (jp - jm) / (2 * eps)
    ad = sum(float(np.dot(gq, dq)) for gq, dq in zip(gb, d))
    print("batch gradient finite differences: fd", fd, "adjoint", ad, "rel", dot_rel(fd, ad), flush=True)
    assert dot_rel(fd, ad) < FD_TOL
    b.close()
    for e in shots:
        e.close()


@pytest.mark.parametrize("bc", [[2, 2, 2, 2, 0, 2], [1, 2, 2, 1, 2, 0]])
def test_gradient_damping_layers_bitwise(torch, bc):
    """Kernel variant 9 runs the damping transpose of every reverse step only
    over the tiles and planes of the supergrid layers (its outputs vanish
    elsewhere): the gradients must be variant 8's to the last bit on a grid
    of several tiles per axis."""
    from sw4lite_amd import EW
    nx, ny, nz, h, nsteps = 96, 40, 41, 0.05, 10
    rho, mu, la = layered_material(nx, ny, nz, h, 9)
    rng = np.random.default_rng(59)
    recs = [(1.0, 0.9, 0.0, 0), (3.2, 1.4, 0.6, 1), (4.5, 0.3, 0.2, 0)]
    gtr = rng.standard_normal((nsteps + 1, len(recs), 3))
    res = []
    for variant in (8, 9):
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
        e.set_kernel_variant(variant, 7)
        e.set_material(rho, mu, la)
        for s in SOURCES[:2]:
            e.add_moment_source(s["x"] * 0.5, s["y"] * 0.5, s["z"] * 0.5, s["m"], s["t0"], s["freq"], s["tfkind"])
        for x, y, z, m in recs:
            e.add_receiver(x, y, z, m)
        e.setup(nsteps=nsteps)
        res.append(e.gradient(gtr, nsteps, ckpt=4))
        e.close()
    ndiff = [int(np.count_nonzero(a != b)) for a, b in zip(*res)]
    print("gradient, damping transpose over the layers: bc", bc, "points differing from variant 8:", ndiff,
          flush=True)
    assert np.max(np.abs(res[0][0])) > 0 and ndiff == [0, 0, 0]


def smooth_direction(rng, mat, shape):
    """A random direction of (rho, mu, lambda), each a relative perturbation
    made of a few random low-order cosine modes over the box."""
    nk, nj, ni = shape
    z, y, x = np.meshgrid(np.linspace(0, 1, nk), np.linspace(0, 1, nj), np.linspace(0, 1, ni), indexing="ij")
    out = []
    for m in mat:
        f = np.full(shape, rng.uniform(-1, 1))
        for _ in range(4):
            a, b, c = rng.integers(0, 3, size=3)
            f += rng.uniform(-1, 1) * np.cos(np.pi * (a * x + rng.uniform())) * np.cos(np.pi * (b * y + rng.uniform())) \
                * np.cos(np.pi * (c * z + rng.uniform()))
        out.append(f.ravel() * m)
    return out


def test_gradient_16_directions_fp64_and_fp32(torch):
    """The gradient against central finite differences of the misfit along
    16 random smooth directions (rho, mu, lambda perturbed together), for the
    FP64 step and for the FP32 form (sw4hip_ew_gradient_comp: the FP32 run
    recomputed from float checkpoints, the reverse in FP64).  The FP32
    gradient also stays near the FP64 gradient of the same traces'
    derivative, though not at the trajectories' 1e-7: the density term
    weighs the corrector's small state differences U^{n+1} - Up (about
    (w dt)^4/12 of the state), in which the float rounding of the
    trajectory is a part in 1e3 (measured 1.4e-2 relative L2 for rho, 3e-3
    for mu and 6e-3 for lambda)."""
    nx, ny, nz, h, nsteps = 24, 22, 20, 0.1, 30
    e, mat = make_ew(nx, ny, nz, h, [2, 2, 2, 2, 0, 2], nsteps)
    rng = np.random.default_rng(61)
    shape = (nz + 4, ny + 4, nx + 4)
    grads = {}
    for mode, fp32, eps, tol in ((0, False, 1e-6, FD_TOL), (2, True, 1e-3, FD32_TOL)):
        J0, tr = misfit(e, mat, nsteps, 0.0, mode)
        dobs = 0.7 * tr + 0.1 * np.max(np.abs(tr)) * rng.standard_normal(tr.shape)
        J0, tr = misfit(e, mat, nsteps, dobs, mode)
        g = e.gradient(tr - dobs, nsteps, ckpt=7, fp32=fp32)
        grads[fp32] = (g, tr - dobs)
        errs = []
        for q in range(16):
            d = smooth_direction(rng, mat, shape)
            jp, _ = misfit(e, [m + eps * dd for m, dd in zip(mat, d)], nsteps, dobs, mode)
            jm, _ = misfit(e, [m - eps * dd for m, dd in zip(mat, d)], nsteps, dobs, mode)
            fd = (jp - jm) / (2 * eps)
            ad = sum(float(np.dot(gq, dq)) for gq, dq in zip(g, d))
            errs.append(dot_rel(fd, ad))
        print("gradient, 16 directions,", "FP32 form" if fp32 else "FP64", ": worst relative difference",
              max(errs), "median", float(np.median(errs)), flush=True)
        assert max(errs) < tol
    # the FP64 gradient of the FP32 run's trace derivative, against the FP32 one
    g32, gtr32 = grads[True]
    g64 = e.gradient(gtr32, nsteps, ckpt=7)
    l2 = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(g32, g64)]
    print("FP32-form gradient against the FP64 gradient of the same trace derivative: relative L2 (rho, mu, "
          "lambda)", l2, flush=True)
    assert max(l2) < 5e-2
    e.close()


def test_batch_gradient_fp32_matches_single_shots(torch):
    """sw4hip_batch_gradient_comp: every shot's gradient must be its own
    sw4hip_ew_gradient_comp to the last bit, with shots in flight together."""
    from sw4lite_amd import Batch
    nsteps = 12
    grids = [(24, 22, 20), (20, 18, 16)]
    rng = np.random.default_rng(67)
    shots = [make_ew(nx, ny, nz, 0.1, [2, 2, 2, 2, 0, 2], nsteps)[0] for nx, ny, nz in grids]
    gtrs = [rng.standard_normal((nsteps + 1, len(RECS), 3)) for _ in shots]
    single = [e.gradient(g, nsteps, ckpt=5, fp32=True) for e, g in zip(shots, gtrs)]
    b = Batch(shots)
    got = b.gradient(gtrs, nsteps, ckpt=5, max_concurrent=2, fp32=True)
    for q, (a, w) in enumerate(zip(got, single)):
        ndiff = [int(np.count_nonzero(x != y)) for x, y in zip(a, w)]
        print("batch FP32-form gradient, shot", q, grids[q], "points differing from the single shot:", ndiff,
              flush=True)
        assert np.max(np.abs(w[0])) > 0 and ndiff == [0, 0, 0]
    b.close()
    for e in shots:
        e.close()
