"""GPU parity: every kernel-level entry point of libsw4hip.so and the EW time
loop (fused sweeps and the reference stage order) against the CPU oracle on
the same seeded inputs.  Bar: max|gpu - oracle| / max|oracle| <= 1e-12 (FP64;
the GPU contracts multiply-adds into FMAs, the oracle does not)."""
import numpy as np
import pytest

from common import grid_box, npts, random_inputs, relerr, to_layout
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


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("os45", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_rhs4sg(torch, coefs, corder, os45):
    from sw4lite_amd import ew
    acof, ghcof, bope, _ = coefs
    nx, ny, nz = 37, 21, 19
    box, f = random_inputs(100 + corder, nx, ny, nz)
    u = to_layout(f["u"], box, corder)
    want = np.zeros_like(u)
    os_ = [0, 0, 0, 0, *os45]
    O.rhs4sg(corder, box, nz, os_, acof, bope, ghcof, want, u, f["mu"], f["la"], 0.3, f["strx"], f["stry"], f["strz"])
    lu = dev(torch, np.zeros_like(u))
    ew.rhs4sg(corder, box, nz, os_, acof, bope, ghcof, lu, dev(torch, u), dev(torch, f["mu"]), dev(torch, f["la"]),
              0.3, dev(torch, f["strx"]), dev(torch, f["stry"]), dev(torch, f["strz"]))
    got = host(lu)
    print("rhs4sg corder", corder, "onesided", os45, "relerr", relerr(got, want), flush=True)
    assert relerr(got, want) < TOL


@pytest.mark.parametrize("corder", [0, 1])
def test_pointwise_updates(torch, corder):
    from sw4lite_amd import ew
    box, f = random_inputs(7, 20, 18, 16)
    g = {k: to_layout(f[k], box, corder) for k in ("u", "um", "up")}
    _, f2 = random_inputs(8, 20, 18, 16)
    lu, fo = to_layout(f2["u"], box, corder), to_layout(f2["um"], box, corder)
    dt = 0.013
    # predictor
    want = np.zeros_like(g["u"])
    O.predfort(corder, box, want, g["u"], g["um"], lu, fo, f["rho"], dt * dt)
    up = dev(torch, np.zeros_like(want))
    ew.predfort(corder, box, up, dev(torch, g["u"]), dev(torch, g["um"]), dev(torch, lu), dev(torch, fo),
                dev(torch, f["rho"]), dt * dt)
    assert relerr(host(up), want) < TOL
    # corrector
    want = g["up"].copy()
    O.corrfort(corder, box, want, lu, fo, f["rho"], dt ** 4)
    up = dev(torch, g["up"])
    ew.corrfort(corder, box, up, dev(torch, lu), dev(torch, fo), dev(torch, f["rho"]), dt ** 4)
    assert relerr(host(up) - g["up"], want - g["up"]) < TOL
    # second time difference
    want = np.zeros_like(g["u"])
    O.dpdmtfort(box, g["up"], g["u"], g["um"], want, 1 / dt ** 2)
    ua = dev(torch, np.zeros_like(want))
    ew.dpdmtfort(box, dev(torch, g["up"]), dev(torch, g["u"]), dev(torch, g["um"]), ua, 1 / dt ** 2)
    assert relerr(host(ua), want) < TOL


@pytest.mark.parametrize("corder,tt", [(0, 0), (1, 0), (1, 1)])
def test_force(torch, corder, tt):
    """sw4hip_force (EW::Force) against the oracle's: 40 grid-point sources
    of the three time functions, several at the same grid points (summed in
    the table's order), a second call with another table reusing none of
    the first one's device copy, and the same table again."""
    from sw4lite_amd import ew
    box = (-1, 22, -1, 20, -1, 18)
    npts = 24 * 22 * 20
    for call, seed in enumerate((31, 32, 31)):
        rng = np.random.default_rng(seed + 10 * corder + 100 * tt)
        n = 40
        ijk = np.stack([rng.integers(1, 21, n), rng.integers(1, 19, n), rng.integers(1, 17, n)], axis=1)
        ijk[5:12] = ijk[4]  # a point with eight sources
        frc = rng.uniform(-1, 1, (n, 3))
        freq = rng.uniform(2.0, 9.0, n)
        t0 = rng.uniform(0.0, 0.3, n)
        tf = rng.integers(0, 3, n)
        t = 0.21
        want = np.zeros(3 * npts)
        O.force(corder, box, want, ijk, frc, freq, t0, tf, t, tt)
        got = dev(torch, np.full(3 * npts, 5.0))  # every value is overwritten
        ew.force(corder, box, t, tt, ijk, frc, freq, t0, tf, got)
        assert np.count_nonzero(want) > 0
        err = relerr(host(got), want)
        print("force corder", corder, "tt", tt, "call", call, "relerr", err, flush=True)
        assert err < 1e-14


@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4(torch, corder):
    from sw4lite_amd import ew
    box, f = random_inputs(9, 33, 17, 21)
    u, um, up0 = (to_layout(f[k], box, corder) for k in ("u", "um", "up"))
    one = [f[k] for k in ("dcx", "dcy", "dcz", "strx", "stry", "strz", "cox", "coy", "coz")]
    want = up0.copy()
    O.addsgd4(corder, box, want, u, um, f["rho"], *one, 0.02)
    up = dev(torch, up0)
    ew.addsgd4(corder, box, up, dev(torch, u), dev(torch, um), dev(torch, f["rho"]), *[dev(torch, a) for a in one],
               0.02)
    got = host(up)
    print("addsgd4 corder", corder, "relerr", relerr(got - up0, want - up0), flush=True)
    assert relerr(got - up0, want - up0) < TOL


@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("bc", [[2, 2, 2, 2, 0, 2], [1, 1, 2, 2, 2, 0], [3, 3, 3, 3, 0, 0], [2, 2, 1, 1, 1, 1]])
def test_bcfortsg(torch, coefs, corder, bc):
    from sw4lite_amd import ew
    nx, ny, nz = 16, 15, 14
    box, f = random_inputs(10, nx, ny, nz)
    sbop = coefs[3]
    desc = L.grid_desc(nx, ny, nz, 0.1, bc, corder=corder)
    wind, nbc = ew.host_boundary_windows(desc)
    rng = np.random.default_rng(11)
    ni, nj = nx + 4, ny + 4
    bforce = []
    for s in range(6):
        m = ni * nj if bc[s] == 0 else max(int(nbc[s]), 1)
        bforce.append(rng.uniform(-1, 1, 3 * m))
    u0 = to_layout(f["u"], box, corder)
    want = u0.copy()
    O.bcfortsg(corder, box, wind, nx, ny, nz, want, 0.1, bc, sbop, f["mu"], f["la"], bforce, f["strx"], f["stry"])
    u = dev(torch, u0)
    ew.bcfortsg(corder, box, wind, nx, ny, nz, u, 0.1, bc, sbop, dev(torch, f["mu"]), dev(torch, f["la"]),
                [dev(torch, b) for b in bforce], dev(torch, f["strx"]), dev(torch, f["stry"]))
    got = host(u)
    assert np.count_nonzero(got != u0) > 0
    assert relerr(got, want) < TOL


# ---------------------------------------------------------------------------
def layered_material(nx, ny, nz, h, seed):
    rng = np.random.default_rng(seed)
    box = grid_box(nx, ny, nz)
    k = np.arange(box[4], box[5] + 1)
    z = (k - 1) * h
    cs = 1.0 + 1.5 * z / z.max()
    cs3 = np.broadcast_to(cs[:, None, None], (nz + 4, ny + 4, nx + 4))
    p = np.exp(0.05 * rng.standard_normal((nz + 4, ny + 4, nx + 4)))
    cs3 = cs3 * p
    cp3 = 1.8 * cs3
    rho = 1.7 * cp3 ** 0.25
    mu = rho * cs3 ** 2
    la = rho * cp3 ** 2 - 2 * mu
    return rho.ravel().copy(), mu.ravel().copy(), la.ravel().copy()


SOURCES = [dict(x=1.2, y=1.1, z=0.7, m=[0.3, 1.0, -0.2, -0.5, 0.4, 0.2], t0=0.12, freq=6.0, tfkind=L.TF_RICKER),
           dict(x=0.9, y=1.5, z=0.35, m=[1.0, 0.0, 0.5, -1.0, 0.0, 0.0], t0=0.1, freq=20.0, tfkind=L.TF_GAUSSIAN),
           dict(x=0.9, y=1.5, z=0.35, m=[0.2, 0.1, 0.0, 0.3, 0.0, -0.5], t0=0.0, freq=8.0, tfkind=L.TF_C6SMOOTHBUMP)]


@pytest.mark.parametrize("mode,corder,bc", [
    (0, 1, [2, 2, 2, 2, 0, 2]),
    (1, 1, [2, 2, 2, 2, 0, 2]),
    (1, 0, [2, 2, 2, 2, 0, 2]),
    (0, 1, [2, 2, 2, 2, 0, 0]),
    (0, 1, [2, 2, 2, 2, 2, 2]),
    (0, 1, [3, 3, 3, 3, 0, 1]),
])
def test_ew_steps_match_oracle(torch, mode, corder, bc):
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    nx, ny, nz, h = 30, 28, 26, 0.08
    rho, mu, la = layered_material(nx, ny, nz, h, 5)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, SOURCES, corder=corder, sg_gp=6, nsteps=12)
    e = EW(nx, ny, nz, h, bc, corder=corder, sg_gp_thickness=6)
    e.set_material(rho, mu, la)
    for s in SOURCES:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    e.setup(nsteps=12)
    inf = e.info()
    assert inf.dt == pytest.approx(ref.dt, rel=1e-15)
    src = e.sources()
    assert np.array_equal(src["ijk"], ref.src[0])
    assert relerr(src["force"], ref.src[1]) < 1e-15
    for _ in range(3):
        e.timesteploop(4, mode=mode)
        ref.step(4)
        u, um = e.get_solution()
        err = relerr(u, ref.U)
        print("EW mode", mode, "corder", corder, "bc", bc, "cycle", ref.cycle, "relerr", err, flush=True)
        assert err < TOL
        assert relerr(um, ref.Um) < TOL
    e.close()


def test_full_size_fused_matches_reference_order(torch):
    """512^3 (the bench workload's size, 137 M points per component): the
    fused sweeps and the reference stage order agree after 3 steps, and the
    solution is finite and non-trivial.  Size-independent check of the
    64-bit indexing and the launch geometry at full scale."""
    import bench
    from sw4lite_amd import EW
    n, h = 512, 20.0
    srcs = bench.sources(n, h, seed=2)
    rho, mu, la = bench.make_material(torch, n, h, seed=2)
    res = []
    for mode in (0, 1):
        e = EW(n, n, n, h, [2, 2, 2, 2, 0, 2], corder=1, sg_gp_thickness=30)
        e.set_material(rho, mu, la)
        for s in srcs:
            e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=500)
        e.timesteploop(3, mode=mode)
        u, _ = e.get_solution()
        res.append(u)
        e.close()
        torch.cuda.empty_cache()
    a, b = res
    assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
    err = relerr(a, b)
    print("512^3 fused vs reference order after 3 steps: relerr", err, "max|u|", np.max(np.abs(b)), flush=True)
    assert err < TOL


@pytest.mark.parametrize("variant,kchunk,bc,nx", [(*c, 45)[:4] for c in [
    (1, 32, [2, 2, 2, 2, 0, 2]),
    (1, 7, [2, 2, 2, 2, 0, 0]),
    (1, 5, [2, 2, 2, 2, 2, 2]),
    (2, 7, [2, 2, 2, 2, 0, 2]),
    (2, 32, [2, 2, 2, 2, 0, 0]),
    (3, 5, [2, 2, 2, 2, 0, 2]),
    (4, 32, [2, 2, 2, 2, 0, 2]),
    (4, 7, [2, 2, 2, 2, 0, 0]),
    (4, 5, [2, 2, 2, 2, 2, 2]),
    (4, 1, [2, 2, 2, 2, 0, 2]),
    (4, 64, [2, 2, 2, 2, 0, 2]),
    (5, 64, [2, 2, 2, 2, 0, 2]),
    (5, 5, [2, 2, 2, 2, 2, 2]),
    (6, 64, [2, 2, 2, 2, 0, 2]),
    (6, 7, [2, 2, 2, 2, 0, 0]),
    (6, 5, [2, 2, 2, 2, 2, 2]),
    (0, 32, [2, 2, 2, 2, 0, 2]),
    # variant 7 takes its asynchronous-load predictor only for even row lengths
    (7, 64, [2, 2, 2, 2, 0, 2], 44),
    (7, 7, [2, 2, 2, 2, 0, 0], 46),
    (7, 5, [2, 2, 2, 2, 2, 2], 44),
    (7, 1, [2, 2, 2, 2, 0, 2], 44),
    (7, 64, [2, 2, 2, 2, 0, 2]),
    # variant 8 stores the acceleration (even row lengths; otherwise it is 7)
    (8, 64, [2, 2, 2, 2, 0, 2], 44),
    (8, 7, [2, 2, 2, 2, 0, 0], 46),
    (8, 5, [2, 2, 2, 2, 2, 2], 44),
    (8, 7, [2, 2, 2, 2, 0, 2]),
    (8, 7, [3, 3, 3, 3, 0, 1], 44),
    # variant 9 runs the damping sweep only over the supergrid layers
    (9, 64, [2, 2, 2, 2, 0, 2], 96),
    (9, 7, [2, 2, 2, 2, 2, 2], 96),
    (9, 5, [2, 2, 2, 2, 0, 0], 45),
    # variant 10 takes four-byte asynchronous loads (and the stored
    # acceleration) for odd row lengths
    (10, 64, [2, 2, 2, 2, 0, 2], 45),
    (10, 7, [2, 2, 2, 2, 2, 2], 45),
    (10, 5, [2, 2, 2, 2, 0, 0], 44),
]])
def test_sweep_variants_match_oracle(torch, variant, kchunk, bc, nx):
    """The LDS-staged k-marching sweeps on a grid that is not a multiple of
    the 32x8 tile, with several k-chunks and a partial last chunk."""
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    ny, nz, h = 19, 41, 0.1
    rho, mu, la = layered_material(nx, ny, nz, h, 6)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, SOURCES, corder=1, sg_gp=6, nsteps=8)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
    e.set_kernel_variant(variant, kchunk)
    e.set_material(rho, mu, la)
    for s in SOURCES:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    e.setup(nsteps=8)
    e.timesteploop(8, mode=0)
    ref.step(8)
    u, um = e.get_solution()
    e.close()
    err = relerr(u, ref.U)
    print("sweep variant", variant, "kchunk", kchunk, "bc", bc, "nx", nx, "relerr", err, flush=True)
    assert err < TOL and relerr(um, ref.Um) < TOL


@pytest.mark.parametrize("variant", [7, 8])
def test_thin_grid_free_top_and_bottom(torch, variant):
    """nz = 12 (the smallest grid accepted) with free surfaces at top and
    bottom: the SBP rows of both ends cover every plane and the interior
    sweep has none, so the stored acceleration of variant 8 has nowhere to
    be stored and the step must run as variant 7 does."""
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    nx, ny, nz, h, bc = 44, 19, 12, 0.1, [2, 2, 2, 2, 0, 0]
    rho, mu, la = layered_material(nx, ny, nz, h, 8)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, SOURCES, corder=1, sg_gp=6, nsteps=6)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
    e.set_kernel_variant(variant, 64)
    e.set_material(rho, mu, la)
    for s in SOURCES:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    e.setup(nsteps=6)
    e.timesteploop(6, mode=0)
    ref.step(6)
    u, um = e.get_solution()
    e.close()
    err = relerr(u, ref.U)
    print("thin grid nz=12 variant", variant, "relerr", err, flush=True)
    assert np.max(np.abs(ref.U)) > 0 and err < TOL and relerr(um, ref.Um) < TOL


# supergrid and Dirichlet sides: the ghost shell's acceleration is final;
# periodic sides: the boundary conditions change the ghost frame afterwards
@pytest.mark.parametrize("bc,nx,kchunk", [([2, 2, 2, 2, 0, 2], 44, 7), ([2, 2, 2, 2, 0, 0], 46, 64),
                                          ([2, 2, 2, 2, 2, 2], 44, 5), ([3, 3, 3, 3, 0, 1], 44, 7),
                                          ([1, 1, 2, 2, 0, 2], 46, 7)])
def test_stored_acceleration_bitwise(torch, bc, nx, kchunk):
    """Variant 8 stores the acceleration in the predictor sweep, completes it
    where later stages change Up (sources, ghost frame, the planes around the
    sweep) and lets the corrector stage it: every value the corrector reads
    must be the one variant 7 forms on the fly, so the solutions agree to
    the last bit."""
    from sw4lite_amd import EW
    ny, nz, h = 19, 41, 0.1
    rho, mu, la = layered_material(nx, ny, nz, h, 6)
    res = []
    for variant in (7, 8):
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
        e.set_kernel_variant(variant, kchunk)
        e.set_material(rho, mu, la)
        for s in SOURCES:
            e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=8)
        e.timesteploop(8, mode=0)
        res.append(e.get_solution())
        e.close()
    (u7, um7), (u8, um8) = res
    assert np.max(np.abs(u7)) > 0
    ndiff = int(np.count_nonzero(u7 != u8)) + int(np.count_nonzero(um7 != um8))
    print("stored acceleration: bc", bc, "nx", nx, "kchunk", kchunk, "points differing from variant 7:", ndiff, flush=True)
    assert ndiff == 0


@pytest.mark.parametrize("bc,kchunk", [([2, 2, 2, 2, 0, 2], 7), ([2, 2, 2, 2, 2, 2], 5), ([1, 1, 2, 2, 0, 2], 64),
                                       ([2, 2, 1, 1, 2, 0], 9), ([3, 3, 3, 3, 0, 1], 7)])
def test_damping_layers_bitwise(torch, bc, kchunk):
    """Variant 9 runs the damping sweep only on the tiles and planes where
    a supergrid profile is nonzero (elsewhere the damping term is an exact
    zero): on a grid of three tiles in x, five in y and k-chunks that cut
    the z layers, the solution must equal variant 8's to the last bit
    (periodic x, y and a Dirichlet bottom: no layer, nothing runs)."""
    from sw4lite_amd import EW
    nx, ny, nz, h = 96, 40, 41, 0.05
    rho, mu, la = layered_material(nx, ny, nz, h, 9)
    res = []
    for variant in (8, 9):
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
        e.set_kernel_variant(variant, kchunk)
        e.set_material(rho, mu, la)
        for s in SOURCES:
            e.add_moment_source(s["x"] * 0.5, s["y"] * 0.5, s["z"] * 0.5, s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=8)
        e.timesteploop(8, mode=0)
        res.append(e.get_solution())
        e.close()
    (u8, um8), (u9, um9) = res
    assert np.max(np.abs(u8)) > 0
    ndiff = int(np.count_nonzero(u8 != u9)) + int(np.count_nonzero(um8 != um9))
    print("damping layers: bc", bc, "kchunk", kchunk, "points differing from variant 8:", ndiff, flush=True)
    assert ndiff == 0


@pytest.mark.parametrize("bc,kchunk", [([2, 2, 2, 2, 0, 2], 64), ([2, 2, 2, 2, 2, 2], 7), ([3, 3, 3, 3, 0, 1], 5)])
def test_odd_rows_bitwise(torch, bc, kchunk):
    """A grid of odd row length (nx = 95): variant 8 runs the register-staged
    sweeps and damps the whole box, variant 9 damps only the listed tiles of
    the supergrid layers, variant 10 stages by four-byte asynchronous loads
    and stores the acceleration; the solutions must agree to the last bit."""
    from sw4lite_amd import EW
    nx, ny, nz, h = 95, 40, 41, 0.05
    rho, mu, la = layered_material(nx, ny, nz, h, 9)
    res = []
    for variant in (8, 9, 10):
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
        e.set_kernel_variant(variant, kchunk)
        e.set_material(rho, mu, la)
        for s in SOURCES:
            e.add_moment_source(s["x"] * 0.5, s["y"] * 0.5, s["z"] * 0.5, s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=8)
        e.timesteploop(8, mode=0)
        res.append(e.get_solution())
        e.close()
    (u8, um8) = res[0]
    assert np.max(np.abs(u8)) > 0
    for v, (u, um) in zip((9, 10), res[1:]):
        ndiff = int(np.count_nonzero(u8 != u)) + int(np.count_nonzero(um8 != um))
        print("odd rows: bc", bc, "kchunk", kchunk, "variant", v, "points differing from variant 8:", ndiff, flush=True)
        assert ndiff == 0


def test_receivers_match_solution(torch):
    """Displacement and velocity receivers (EW::extractRecordData) record the
    solution at the nearest grid point at t=0 and after every step."""
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    nx, ny, nz, h = 30, 28, 26, 0.08
    bc = [2, 2, 2, 2, 0, 2]
    rho, mu, la = layered_material(nx, ny, nz, h, 7)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, SOURCES, corder=1, sg_gp=6, nsteps=6)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
    e.set_material(rho, mu, la)
    for s in SOURCES:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    recs = [(1.0, 1.05, 0.0, 0), (1.53, 0.77, 0.41, 1), (0.3, 2.0, 1.1, 0)]
    for x, y, z, m in recs:
        e.add_receiver(x, y, z, m)
    e.setup(nsteps=6)
    e.timesteploop(6, mode=0)
    got = e.receivers()
    e.close()
    assert got.shape == (7, 3, 3)
    npl = (nx + 4) * (ny + 4) * (nz + 4)
    want = np.zeros_like(got)
    for n in range(7):
        if n:
            ref.step(1)
        for r, (x, y, z, m) in enumerate(recs):
            i, j, k = (int(round(v / h)) + 1 for v in (x, y, z))
            p = (i + 1) + (nx + 4) * ((j + 1) + (ny + 4) * (k + 1))
            if m == 0 or n == 0:
                want[n, r] = [ref.U[p + c * npl] for c in range(3)]
            else:
                want[n, r] = [(ref.U[p + c * npl] - ref.Up[p + c * npl]) / (2 * ref.dt) for c in range(3)]
    # t = 0 velocity receivers record (U - Um)/(2 dt) of the zero initial data
    assert relerr(got[1:], want[1:]) < TOL
    assert np.all(got[0] == 0)
