"""GPU: the FP32 form of the fused step with the compensated time update
(EW step mode 2) against the FP64 oracle on the same seeded inputs.

The stencils, the supergrid damping and the material run in FP32 (unit
roundoff 2^-24 ~ 6e-8); the time update keeps u as the float pair U + C, and
the boundary closures, source time functions and receivers run in FP64.  The
bar for every output field (u, um, the receiver traces) is
max|fp32 - oracle| / max|oracle| <= 1e-5, the seismogram tolerance the FP32
form is specified to; each test prints the error it measured per field."""
import numpy as np
import pytest

from common import relerr
from sw4lite_amd import lib as L
from test_gpu_parity import SOURCES, layered_material

pytestmark = pytest.mark.gpu
FP32_TOL = 1e-5


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a visible GPU"
    L.load()
    return _t


def make_pair(nx, ny, nz, h, bc, seed, nsteps, sources=SOURCES, kchunk=None):
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    rho, mu, la = layered_material(nx, ny, nz, h, seed)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, sources, corder=1, sg_gp=6, nsteps=nsteps)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
    if kchunk is not None:
        e.set_kernel_variant(6, kchunk)
    e.set_material(rho, mu, la)
    for s in sources:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    return e, ref


def field_errors(u, um, ref):
    """Per-field, per-component errors of the solution against the oracle."""
    n = u.size // 3
    out = {}
    for name, a, b in (("u", u, ref.U), ("um", um, ref.Um)):
        out[name] = relerr(a, b)
        out[name + "_comp"] = [relerr(a[c * n:(c + 1) * n], b[c * n:(c + 1) * n]) for c in range(3)]
    return out


@pytest.mark.parametrize("bc", [
    [2, 2, 2, 2, 0, 2],
    [2, 2, 2, 2, 0, 0],
    [2, 2, 2, 2, 2, 2],
    [3, 3, 3, 3, 0, 1],
])
def test_comp_steps_match_oracle(torch, bc):
    nx, ny, nz, h = 30, 28, 26, 0.08
    e, ref = make_pair(nx, ny, nz, h, bc, 5, 12)
    e.setup(nsteps=12)
    for _ in range(3):
        e.timesteploop(4, mode=2)
        ref.step(4)
        u, um = e.get_solution()  # writes the float state back to U, Um
        err = field_errors(u, um, ref)
        print("FP32 comp bc", bc, "cycle", ref.cycle, err, flush=True)
        assert np.all(np.isfinite(u)) and np.max(np.abs(u)) > 0
        assert err["u"] < FP32_TOL and err["um"] < FP32_TOL
    e.close()


@pytest.mark.parametrize("kchunk,bc", [
    (5, [2, 2, 2, 2, 0, 2]),
    (7, [2, 2, 2, 2, 0, 0]),
    (64, [2, 2, 2, 2, 2, 2]),
    (1, [2, 2, 2, 2, 0, 2]),
])
def test_comp_sweep_chunks_match_oracle(torch, kchunk, bc):
    """The FP32 sweeps on a grid that is not a multiple of the 32x8 tile,
    with several k-chunks and a partial last chunk."""
    nx, ny, nz, h = 45, 19, 41, 0.1
    e, ref = make_pair(nx, ny, nz, h, bc, 6, 8, kchunk=kchunk)
    e.setup(nsteps=8)
    e.timesteploop(8, mode=2)
    ref.step(8)
    u, um = e.get_solution()
    e.close()
    err = field_errors(u, um, ref)
    print("FP32 comp kchunk", kchunk, "bc", bc, err, flush=True)
    assert err["u"] < FP32_TOL and err["um"] < FP32_TOL


def test_comp_state_round_trip(torch):
    """U, Um -> float state (U + C, D) -> U, Um: u to the pair's 48 bits,
    um to float rounding of u - um."""
    from sw4lite_amd import EW
    nx, ny, nz, h = 20, 18, 16, 0.1
    rho, mu, la = layered_material(nx, ny, nz, h, 3)
    e = EW(nx, ny, nz, h, [2, 2, 2, 2, 0, 2], corder=1, sg_gp_thickness=4)
    e.set_material(rho, mu, la)
    e.setup(nsteps=1)
    rng = np.random.default_rng(11)
    n = 3 * (nx + 4) * (ny + 4) * (nz + 4)
    u = rng.uniform(-1, 1, n) * 1e3
    um = u - rng.uniform(-1, 1, n)
    e.set_solution(u, um)
    e.timesteploop(0, mode=2)
    u2, um2 = e.get_solution()
    e.close()
    assert relerr(u2, u) < 2.0 ** -46
    assert np.max(np.abs((u2 - um2) - (u - um))) <= 2.0 ** -23 * np.max(np.abs(u - um))


def test_comp_receivers_match_oracle(torch):
    """Displacement (U + C) and velocity ((Dn + D)/(2 dt)) receivers of the
    FP32 form against the oracle's solution at the nearest grid point."""
    nx, ny, nz, h = 30, 28, 26, 0.08
    bc = [2, 2, 2, 2, 0, 2]
    e, ref = make_pair(nx, ny, nz, h, bc, 7, 6)
    recs = [(1.0, 1.05, 0.0, 0), (1.53, 0.77, 0.41, 1), (0.3, 2.0, 1.1, 0), (1.2, 1.1, 0.7, 1)]
    for x, y, z, m in recs:
        e.add_receiver(x, y, z, m)
    e.setup(nsteps=6)
    e.timesteploop(6, mode=2)
    got = e.receivers()
    e.close()
    assert got.shape == (7, 4, 3)
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
    for r, (_, _, _, m) in enumerate(recs):
        err = relerr(got[1:, r], want[1:, r])
        print("FP32 comp receiver", r, "mode", m, "relerr", err, flush=True)
        assert err < FP32_TOL


def test_comp_pointsource_error(torch):
    """The analytic point-source errors (Linf, L2) of the reference deck in the
    FP32 form equal the reference's PointSourceErr.txt values to 1e-4
    relative: FP32 rounding stays far below the discretisation error."""
    from sw4lite_amd import EW
    from test_pointsource import DECK, SRC, WANT
    d = DECK
    e = EW(d["nx"], d["ny"], d["nz"], d["h"], d["bc"], corder=1, sg_gp_thickness=d["sg_gp"])
    e.set_material_const(d["rho"], d["cs"], d["cp"])
    e.add_moment_source(SRC["x"], SRC["y"], SRC["z"], SRC["m"], SRC["t0"], SRC["freq"], SRC["tfkind"])
    e.setup(tmax=d["tmax"])
    e.timesteploop(mode=2)
    got = e.point_source_error()
    e.close()
    print("FP32 comp pointsource", got, WANT, flush=True)
    for g, w in zip(got, WANT):
        assert abs(g - w) <= 1e-4 * abs(w), (got, WANT)


def test_comp_full_size_matches_fp64(torch):
    """512^3 (the bench workload): the FP32 form against the FP64 fused step
    after 3 steps.  Size-independent check of the 64-bit indexing and the
    launch geometry of the FP32 kernels at full scale."""
    import bench
    from sw4lite_amd import EW
    n, h = 512, 20.0
    srcs = bench.sources(n, h, seed=2)
    rho, mu, la = bench.make_material(torch, n, h, seed=2)
    res = []
    for mode in (0, 2):
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
    assert np.all(np.isfinite(b)) and np.max(np.abs(b)) > 0
    err = relerr(b, a)
    print("512^3 FP32 comp vs FP64 after 3 steps: relerr", err, "max|u|", np.max(np.abs(a)), flush=True)
    assert err < FP32_TOL


@pytest.mark.parametrize("bc,kchunk", [([2, 2, 2, 2, 0, 2], 7), ([2, 2, 2, 2, 0, 0], 64), ([3, 3, 3, 3, 0, 1], 5)])
def test_comp_async_sweeps_bitwise(torch, bc, kchunk):
    """From kernel variant 7 on, the FP32 sweeps bring their planes to LDS by
    asynchronous loads where the row length allows it (nx + 4 divisible by 4,
    as here).  The arithmetic is the same in the same order, so the state
    after mode-2 steps equals variant 6's to the last bit; and it meets the
    FP32 bar against the FP64 oracle."""
    from sw4lite_amd import EW
    nx, ny, nz, h = 44, 19, 41, 0.1
    rho, mu, la = layered_material(nx, ny, nz, h, 6)
    res = []
    for variant in (6, 8):
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=6)
        e.set_kernel_variant(variant, kchunk)
        e.set_material(rho, mu, la)
        for s in SOURCES:
            e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=10)
        e.timesteploop(10, mode=2)
        res.append(e.get_solution())
        e.close()
    (u6, um6), (u8, um8) = res
    assert np.max(np.abs(u6)) > 0
    ndiff = int(np.count_nonzero(u6 != u8)) + int(np.count_nonzero(um6 != um8))
    print("FP32 asynchronous-load sweeps: bc", bc, "kchunk", kchunk, "points differing from variant 6:", ndiff,
          flush=True)
    assert ndiff == 0


def test_comp_async_sweeps_match_oracle(torch):
    e, ref = make_pair(44, 19, 41, 0.1, [2, 2, 2, 2, 0, 2], 6, 10)
    e.set_kernel_variant(8, 7)
    e.setup(nsteps=10)
    e.timesteploop(10, mode=2)
    ref.step(10)
    u, um = e.get_solution()
    e.close()
    err = field_errors(u, um, ref)
    print("FP32 asynchronous-load sweeps vs oracle", err, flush=True)
    assert err["u"] < FP32_TOL and err["um"] < FP32_TOL
