"""Pins the oracle functions that the reference's free kernels do not cover
against the reference's EW class itself (oracle/_ref/libsw4ew.so: the
reference program compiled from its own sources with the image's MPICH and
MKL, objects made from input decks, member kernels called directly):

  * predfort / corrfort / dpdmtfort (ew-cfromfort.C:40-141), both layouts;
  * bcfortsg / bcfortsg_indrev (ew-cfromfort.C:205-747) for every boundary
    type on every side (free surface, Dirichlet and supergrid with random
    boundary data, periodic), both layouts;
  * addsgd4fort(_indrev) (ew-cfromfort.C:748-1056), the EW members;
  * addsgd4cfort(_indrev) (ew-cfromfort.C:1160-1404) with a Jacobian that
    varies in space;
  * freesurfcurvisg(_rev) (curvilinear-c.C:465-764) on a curved grid, with
    and without traction forcing;
  * GetStencilCoefficients (ew-cfromfort.C:1487-1944) against the library's
    table;
  * the time functions (time_functions.C) and the supergrid tapers
    (SuperGrid.C:54-194) at many arguments;
  * the boundary windows, supergrid arrays, grid-point sources and dt of
    deck runs against the library's host set-up;
  * the analytic moment-tensor solution (EW::get_exact_point_source,
    EW.C:3899-4523) for the Gaussian and C6SmoothBump time functions, and the
    error norms (solerr3fort).

Closed forms that fix the same functions without the reference build (they
run wherever the tests run): the Ricker wavelet's value -1 at its centre,
its zeros at t = +-1/(pi f sqrt 2) and zero integral; the Gaussian's peak
f/sqrt(2 pi) and unit integral; the identity addsgd4c(rho, jac) =
addsgd4 without its z part acting on rho*jac (jac varying in space).
"""
import numpy as np
import pytest

from common import grid_box, npts, random_inputs, relerr, to_layout
from oracle import oracle as O
from oracle import refew as R
from sw4lite_amd import ew
from sw4lite_amd import lib as L

needs_ref = pytest.mark.skipif(not R.available(), reason="reference EW library not built (no reference tree)")

TINY = """grid x=2.0 y=1.8 z=1.6 h=0.1
time steps=1
supergrid gp=4
block vp=2 vs=1 rho=1
developer corder=1
"""


@pytest.fixture(scope="module")
def ref():
    if not R.available():
        pytest.skip("reference EW library not built")
    e = R.RefEW(TINY)
    yield e
    e.close()


# --------------------------------------------------------------- pointwise
@needs_ref
@pytest.mark.parametrize("corder", [0, 1])
def test_predictor_corrector_dpdmt(ref, corder):
    box, f = random_inputs(11, 9, 8, 7)
    _, g = random_inputs(12, 9, 8, 7)
    lu, fo = to_layout(g["u"], box, corder), to_layout(g["um"], box, corder)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    a, b = np.zeros_like(u), np.zeros_like(u)
    O.predfort(corder, box, a, u, um, lu, fo, f["rho"], 0.37)
    ref.predfort(corder, box, b, u, um, lu, fo, f["rho"], 0.37)
    assert np.array_equal(a, b)
    O.corrfort(corder, box, a, lu, fo, f["rho"], 0.021)
    ref.corrfort(corder, box, b, lu, fo, f["rho"], 0.021)
    assert np.array_equal(a, b)
    c, d = np.zeros_like(u), np.zeros_like(u)
    O.dpdmtfort(box, a, u, um, c, 7.3)
    ref.dpdmtfort(box, a, u, um, d, 7.3)
    assert np.array_equal(c, d)


# ------------------------------------------------------------- ghost points
BC_SETS = [
    [2, 2, 2, 2, 0, 2],  # free top, supergrid elsewhere
    [1, 1, 1, 1, 1, 1],  # Dirichlet everywhere
    [3, 3, 3, 3, 0, 0],  # periodic x, y; free top and bottom
    [1, 2, 3, 3, 2, 0],  # every type, free bottom
    [2, 1, 1, 2, 1, 0],
]


@needs_ref
@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("bc", BC_SETS)
def test_bcfortsg(ref, coefs, corder, bc):
    nx, ny, nz, h = 13, 12, 14, 0.1
    box, f = random_inputs(21, nx, ny, nz)
    wind, npt = ew.host_boundary_windows(L.grid_desc(nx, ny, nz, h, bc))
    rng = np.random.default_rng(22)
    bforce = [rng.uniform(-1, 1, 3 * npt[s]) if bc[s] in (1, 2) else (rng.uniform(-1, 1, 3 * (nx + 4) * (ny + 4))
                                                                        if bc[s] == 0 else None) for s in range(6)]
    sbop = coefs[3]
    u0 = to_layout(f["u"], box, corder)
    a, b = u0.copy(), u0.copy()
    O.bcfortsg(corder, box, wind, nx, ny, nz, a, h, bc, sbop, f["mu"], f["la"], bforce, f["strx"], f["stry"])
    ref.bcfortsg(corder, box, wind, nx, ny, nz, b, h, bc, sbop, f["mu"], f["la"], bforce, f["strx"], f["stry"])
    assert not np.array_equal(a, u0)
    assert relerr(a, b) < 1e-15


@needs_ref
def test_boundary_set_up_of_a_deck():
    """Windows and one-sided flags of the default deck (free top, supergrid
    elsewhere) equal the library's host set-up."""
    with R.RefEW(TINY) as e:
        wind, npt, bct, onesided = e.boundary()
        nx, ny, nz = (e.boxes[0][1] - 2, e.boxes[0][3] - 2, e.boxes[0][5] - 2)
        assert list(bct) == [2, 2, 2, 2, 0, 2]
        w2, n2 = ew.host_boundary_windows(L.grid_desc(nx, ny, nz, 0.1, list(bct)))
        assert np.array_equal(np.asarray(w2).ravel(), wind) and np.array_equal(np.asarray(n2), npt)
        assert list(onesided) == [0, 0, 0, 0, 1, 0]


# ---------------------------------------------------------------- damping
@needs_ref
@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4_member(ref, corder):
    box, f = random_inputs(31, 10, 12, 9)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    up0 = to_layout(f["up"], box, corder)
    a, b = up0.copy(), up0.copy()
    args = (f["rho"], f["dcx"], f["dcy"], f["dcz"], f["strx"], f["stry"], f["strz"], f["cox"], f["coy"], f["coz"],
            0.02)
    O.addsgd4(corder, box, a, u, um, *args)
    ref.addsgd4(corder, box, b, u, um, *args)
    assert relerr(a - up0, b - up0) < 1e-14


def varying_jac(box, seed):
    rng = np.random.default_rng(seed)
    ib, ie, jb, je, kb, ke = box
    k, j, i = np.meshgrid(np.arange(kb, ke + 1), np.arange(jb, je + 1), np.arange(ib, ie + 1), indexing="ij")
    jac = 0.5 + 0.3 * np.sin(0.7 * i + 0.2 * j) * np.cos(0.4 * k) + 0.1 * rng.uniform(0, 1, i.shape)
    return np.ascontiguousarray(jac.ravel())


@needs_ref
@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4c_member(ref, corder):
    box, f = random_inputs(41, 10, 12, 9)
    jac = varying_jac(box, 42)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    up0 = to_layout(f["up"], box, corder)
    a, b = up0.copy(), up0.copy()
    args = (f["rho"], f["dcx"], f["dcy"], f["strx"], f["stry"], jac, f["cox"], f["coy"], 0.02)
    O.addsgd4c(corder, box, a, u, um, *args)
    ref.addsgd4c(corder, box, b, u, um, *args)
    assert relerr(a - up0, b - up0) < 1e-14


@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4c_is_addsgd4_on_rho_jac(corder):
    """addsgd4c divides by rho*jac and weights the x/y differences with
    rho*jac at the neighbours: with dcz = 0 it equals addsgd4 acting with
    the density rho*jac, for a jac that varies in space (a shifted index
    would change the weights and fail this)."""
    box, f = random_inputs(43, 11, 10, 12)
    jac = varying_jac(box, 44)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    up0 = to_layout(f["up"], box, corder)
    a, b = up0.copy(), up0.copy()
    O.addsgd4c(corder, box, a, u, um, f["rho"], f["dcx"], f["dcy"], f["strx"], f["stry"], jac, f["cox"], f["coy"],
               0.02)
    zeros = np.zeros_like(f["dcz"])
    O.addsgd4(corder, box, b, u, um, f["rho"] * jac, f["dcx"], f["dcy"], zeros, f["strx"], f["stry"], f["strz"],
              f["cox"], f["coy"], np.ones_like(f["coz"]), 0.02)
    assert relerr(a - up0, b - up0) < 1e-14


# ------------------------------------------------------------ curvilinear
@needs_ref
@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("forced", [False, True])
def test_freesurfcurvisg_member(ref, coefs, corder, forced):
    from common import hills_depth
    nx, ny, nz, h = 12, 11, 10, 0.1
    box, z = hills_depth(nx, ny, nz, h, seed=51, nhill=2, amp=0.2)
    n = npts(box)
    met, jac = np.zeros(4 * n), np.zeros(n)
    assert O.metric(corder, box, h, z, met, jac) == 0
    _, f = random_inputs(52, nx, ny, nz)
    nf = 3 * (nx + 4) * (ny + 4)
    # the reference reads its forcing array unconditionally: zero traction is an array of zeros
    forcing = np.random.default_rng(53).uniform(-1, 1, nf) if forced else np.zeros(nf)
    u0 = to_layout(f["u"], box, corder)
    a, b = u0.copy(), u0.copy()
    O.freesurfcurvisg(corder, box, nz, 5, a, f["mu"], f["la"], met, coefs[3], forcing, f["strx"], f["stry"])
    ref.freesurfcurvisg(corder, box, nz, 5, b, f["mu"], f["la"], met, coefs[3], forcing, f["strx"], f["stry"])
    assert not np.array_equal(a, u0)
    assert relerr(a, b) < 1e-13


# --------------------------------------------------------------- tables
@needs_ref
def test_stencil_coefficients_member(ref, coefs):
    mine = coefs
    theirs = ref.stencil_coefficients()
    # library order (acof, ghcof, bope, sbop) is the reference's argument order
    for a, b in zip(mine, theirs):
        assert np.array_equal(np.asarray(a).ravel(), b)


# --------------------------------------------------------- time functions
@needs_ref
@pytest.mark.parametrize("kind,freq,lo,hi", [(0, 5.0, -0.6, 0.6), (1, 12.5, -0.6, 0.6), (2, 2.0, -0.1, 0.6),
                                             (0, 16.7, -0.3, 0.3)])
def test_time_functions_member(kind, freq, lo, hi):
    for t in np.linspace(lo, hi, 257):
        for d2 in (0, 1):
            a = O.timefunc(kind, d2, freq, t)
            b = R.timefunc(kind, d2, freq, t)
            assert abs(a - b) <= 2e-15 * max(1.0, abs(b)), (kind, d2, t, a, b)


def test_ricker_closed_form():
    """Ricker (time_functions.C:175-182): g(0) = -1, g(+-1/(pi f sqrt 2)) = 0,
    g'' (0) = 6 (pi f)^2, zero mean."""
    f = 5.0
    assert O.timefunc(0, 0, f, 0.0) == -1.0
    tz = 1.0 / (np.pi * f * np.sqrt(2.0))
    assert abs(O.timefunc(0, 0, f, tz)) < 1e-15 and abs(O.timefunc(0, 0, f, -tz)) < 1e-15
    assert O.timefunc(0, 0, f, 0.5 * tz) < 0 < O.timefunc(0, 0, f, 1.5 * tz)
    assert O.timefunc(0, 1, f, 0.0) == pytest.approx(6 * (np.pi * f) ** 2, rel=1e-15)
    t = np.linspace(-1, 1, 40001)
    g = np.array([O.timefunc(0, 0, f, x) for x in t])
    assert abs(np.trapezoid(g, t)) < 1e-10
    # the peak of |g| is at 0, the side lobes at t = +-sqrt(3/2)/(pi f) with 2 e^(-3/2)
    ts = np.sqrt(1.5) / (np.pi * f)
    assert O.timefunc(0, 0, f, ts) == pytest.approx(2 * np.exp(-1.5), rel=1e-14)


def test_gaussian_closed_form():
    """Gaussian (time_functions.C:284-290): f/sqrt(2 pi) exp(-(f t)^2/2)."""
    f = 12.5
    assert O.timefunc(1, 0, f, 0.0) == pytest.approx(f / np.sqrt(2 * np.pi), rel=1e-15)
    assert O.timefunc(1, 0, f, 1 / f) == pytest.approx(f / np.sqrt(2 * np.pi) * np.exp(-0.5), rel=1e-15)
    assert O.timefunc(1, 1, f, 0.0) == pytest.approx(-f ** 3 / np.sqrt(2 * np.pi), rel=1e-15)


@needs_ref
@pytest.mark.parametrize("left,right", [(True, True), (True, False), (False, True)])
def test_supergrid_taper_member(left, right):
    """The library's taper (host set-up) against SuperGrid::stretching,
    dampingCoeff and cornerTaper at every grid point of an axis."""
    n, h, gp = 61, 0.25, 12
    x = (np.arange(-1, n + 3) - 1) * h
    s, d, c = R.supergrid_taper(left, 0.0, right, (n - 1) * h, gp * h, x)
    bc = [2 if left else 1, 2 if right else 1, 2, 2, 2, 2]
    sg = ew.host_supergrid(L.grid_desc(n, 30, 30, h, bc, sg_gp_thickness=gp))
    assert np.array_equal(sg["strx"], s)
    assert np.array_equal(sg["dcx"], d)
    assert np.array_equal(sg["cox"], c)


# ------------------------------------------------------- deck set-up runs
C1_DECK = """grid x=6.0 y=6.0 z=6.0 h=0.05
time t=0.6
testpointsource
supergrid gp=20
source x=3.0 y=3.0 z=3.0 Mxx=1 Myy=0.8 Mzz=0.6 Mxy=0.3 Mxz=-0.2 Myz=0.4 t0=0.32 freq=12.5 type=Gaussian
developer cfl=1.3 corder=1
"""


def small_deck(src, bc_top_free=True, steps=2):
    return """grid x=3.0 y=2.8 z=2.4 h=0.1
time steps=%d
testpointsource rho=1.3 cs=0.9 cp=1.7 halfspace=%d
supergrid gp=6
%s
developer cfl=1.3 corder=1
""" % (steps, 1 if bc_top_free else 0, src)


@needs_ref
@pytest.mark.parametrize("z0,top", [(1.23, True), (0.17, True), (0.05, True), (0.0, True), (1.11, False)])
def test_sources_dt_and_supergrid_of_a_deck(z0, top):
    """Grid-point sources (Source::set_grid_point_sources4), dt (computeDT)
    and supergrid arrays of a deck run equal the library's host set-up, for
    off-grid sources in the interior and next to a free surface (SBP norm
    weights, one-sided stencils)."""
    src = "source x=1.37 y=1.21 z=%g Mxx=0.7 Mxy=-0.3 Mxz=0.45 Myy=1.1 Myz=0.2 Mzz=-0.9 t0=0.2 freq=5 type=Ricker" % z0
    with R.RefEW(small_deck(src, top)) as e:
        box = e.boxes[0]
        nx, ny, nz = box[1] - 2, box[3] - 2, box[5] - 2
        bc = [2, 2, 2, 2, 0 if top else 2, 2]
        desc = L.grid_desc(nx, ny, nz, 0.1, bc, sg_gp_thickness=6)
        mine = ew.host_supergrid(desc)
        theirs = e.supergrid()
        for k in ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz"):
            assert np.array_equal(mine[k], theirs[k]), k
        ijkg, forces, _, _ = e.point_sources()
        ijk, frc = ew.host_moment_source(desc, 1.37, 1.21, z0, [0.7, -0.3, 0.45, 1.1, 0.2, -0.9])

        def merged(pts, vals):
            out = {}
            for p, v in zip(pts, vals):
                out[tuple(p)] = out.get(tuple(p), 0) + np.asarray(v)
            return {k: v for k, v in out.items() if np.any(v != 0)}

        a = merged(ijkg[:, :3].tolist(), forces)
        b = merged(ijk.tolist(), frc)
        assert set(a) == set(b)
        scale = max(np.max(np.abs(v)) for v in a.values())
        for k in a:
            assert np.max(np.abs(a[k] - b[k])) <= 1e-13 * scale, k
        # dt of the oracle's computeDT for the same material and CFL
        from oracle.loop import OracleRun
        mu, la = 1.3 * 0.9 ** 2, 1.3 * 1.7 ** 2 - 2 * 1.3 * 0.9 ** 2
        run = OracleRun(nx, ny, nz, 0.1, bc, 1.3, mu, la, [], sg_gp=6, nsteps=2)
        assert run.dt == pytest.approx(e.dt, rel=1e-15)


# ------------------------------------------------- exact solution, norms
@needs_ref
@pytest.mark.parametrize("kind,freq,t0,t", [("Gaussian", 12.5, 0.32, 0.6), ("Gaussian", 8.0, 0.1, 0.45),
                                            ("C6SmoothBump", 1.0, 0.0, 0.6), ("C6SmoothBump", 2.0, 0.1, 0.5)])
def test_exact_solution_member(kind, freq, t0, t):
    """The oracle's analytic moment-tensor solution (its own quadrature of
    the near-field integral) against EW::get_exact_point_source (closed
    forms), for an off-grid source with a full moment tensor, outside the
    ball of radius 4h around the source that the reference's error norm
    leaves out (EW.C:4541-4579).  Inside it the reference's closed form of
    the C6SmoothBump near-field integral (C6SBTP differences,
    EW.C:3846-3868) cancels: at i, j, k = 15, 14, 13 (R = 0.44 h) of the
    second C6 case a 40-digit evaluation gives 1.3100033324551346, the
    oracle 1.3100033324551095 and the reference 1.3100034549500308.  Both
    follow the reference in scaling time, not R, by the frequency of a C6
    source (EW.C:3935, 4075-4081)."""
    m6 = [1.0, 0.3, -0.2, 0.8, 0.4, 0.6]
    x0, y0, z0, h = 1.43, 1.31, 1.17, 0.1
    src = ("source x=%g y=%g z=%g Mxx=%g Mxy=%g Mxz=%g Myy=%g Myz=%g Mzz=%g t0=%g freq=%g type=%s"
           % (x0, y0, z0, *m6, t0, freq, kind))
    with R.RefEW(small_deck(src, bc_top_free=False, steps=1)) as e:
        box = e.boxes[0]
        theirs = e.exact_point_source(t)
    rho, cs, cp = 1.3, 0.9, 1.7
    tk = L.TF_GAUSSIAN if kind == "Gaussian" else L.TF_C6SMOOTHBUMP
    time, fr = (t - t0) * freq, freq
    if tk == L.TF_GAUSSIAN:
        fr = 1 / freq
        time = time * fr
    mine = np.zeros_like(theirs)
    # get_exact_point_source writes interleaved components (EW.C:3985-4010)
    O.exact_moment(box, h, 0.0, x0, y0, z0, rho, cp, cs, time, fr, m6, tk, mine, 0)
    ib, ie, jb, je, kb, ke = box
    k, j, i = np.meshgrid(np.arange(kb, ke + 1), np.arange(jb, je + 1), np.arange(ib, ie + 1), indexing="ij")
    r = np.sqrt(((i - 1) * h - x0) ** 2 + ((j - 1) * h - y0) ** 2 + ((k - 1) * h - z0) ** 2).ravel()
    far = np.repeat(r > 4 * h, 3)
    assert np.max(np.abs(theirs[far])) > 0
    assert relerr(mine[far], theirs[far]) < 1e-12
    assert relerr(mine, theirs) < (1e-12 if tk == L.TF_GAUSSIAN else 1e-6)


@needs_ref
def test_solerr3_member(ref):
    box, f = random_inputs(61, 12, 13, 14)
    _, g = random_inputs(62, 12, 13, 14)
    args = (box, 0.1, g["u"], f["u"], 0.0, 0.61, 0.52, 0.73, 0.4, (2, 10, 3, 11, 2, 12))
    a = O.solerr3(*args, 1)
    b = ref.solerr3(*args)
    assert np.allclose(a, b, rtol=1e-14, atol=0)


CURV_DECK = """grid x=1.28 y=1.28 z=5.0 h=0.02
time steps=0
block vp=2 vs=1 r=1.5
block vp=2.6 vs=1.35 r=1.7 z1=1.0
supergrid gp=30
topography order=6 input=gaussian zmax=2.0 gaussianAmp=0.1 gaussianXc=0.6 gaussianYc=0.4 gaussianLx=0.15 gaussianLy=0.10
developer corder=0
"""


@needs_ref
def test_curvilinear_dt_matrix_against_reference():
    """The curvilinear time-step bound: the oracle's principal-part matrix
    (all six entries, off-diagonal ones included) on the reference's own
    metric, Jacobian and material of its topography grid, with the Cartesian
    bound of the grid below, gives the dt the reference computed with
    LAPACK dspev (EW::computeDT, EW.C:5041-5146) for the deck of its
    curvilinear fixture (pytest/reference/topo/curvilinear.in, a second
    material block so that the minimum is not at a constant state)."""
    from oracle.loop import curvilinear_eigmax
    with R.RefEW(CURV_DECK) as e:
        assert e.ngrids == 2
        cfl = 1.3
        h0, _, _ = e.grid(0)
        rho0, mu0, la0 = (e.array(w, 0) for w in (2, 3, 4))
        dt_cart = cfl * h0 / np.sqrt(np.max((4 * mu0 + la0) / rho0))
        rho1, mu1, la1 = (e.array(w, 1) for w in (2, 3, 4))
        met, jac = e.array(8, 1), e.array(9, 1)
        assert met.size == 4 * jac.size == 4 * rho1.size
        dt_curv = cfl / np.sqrt(curvilinear_eigmax(met, jac, rho1, mu1, la1, 0))
        got = min(dt_cart, dt_curv)
        print("curvilinear dt: reference", e.dt, "oracle", got, "(Cartesian", dt_cart, "curvilinear", dt_curv, ")")
        assert dt_curv < dt_cart  # the bound of the topography grid decides
        assert abs(got - e.dt) <= 1e-13 * e.dt
        # the off-diagonal terms decide the bound here: without them, or
        # with another weight, it moves far beyond the agreement above (their
        # sign does not: where the bound is attained one of zp, zq vanishes
        # and a sign flip of the remaining term is a similarity)
        for scale in (0.0, 0.5, 2.0):
            bad = cfl / np.sqrt(curvilinear_eigmax(met, jac, rho1, mu1, la1, 0, offdiag=scale))
            print("off-diagonal terms scaled by", scale, "dt", bad)
            assert abs(bad - e.dt) > 1e-6 * e.dt
