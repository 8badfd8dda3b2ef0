"""The benchmark configurations as parity cases.

C1, the Gaussian point source (121^3, h = 0.05, rho = 1, cs = 1,
cp = sqrt 3, a full moment tensor at the centre with a Gaussian of
sigma = 0.08 (freq 12.5) centred at t0 = 0.32, supergrid of 20 points on
all six faces, CFL 1.3, t = 0.6): the reference program run on that deck
(tests/golden/make_ref_runs.py) gives dt = 0.6/21 and the analytic-solution
errors Linf, L2 and max|u_exact| of tests/golden/c1_pointsource.json.  The
CPU oracle loop and the HIP library (fused sweeps and the reference stage
order) must give the same numbers to 1e-11 relative, and the HIP field must
equal the oracle's to 1e-12.

C3, the curvilinear workload (a grid under six seeded Gaussian hills of up to
15% of the depth, linear vertical stretching, the layered + lognormal
material, free surface on the curved top, supergrid elsewhere, one Ricker
source at 4 Hz): at 64 x 64 x 48 points the EW loop (reference stage order
with rhs4sgcurv, addsgd4cfort and the curved closure) must equal the
oracle's loop to 1e-12 after 10 steps; at the full 384 x 384 x 256 points,
a few steps must be finite, non-trivial and linear in the source (the moment
tensor doubled doubles every value: scaling by two commutes with every
rounding of a linear scheme).

C4, the precision workload at 1024^3 (h = 10 m): the FP32 form with the
compensated update against the FP64 step on 64 surface receivers over 100
steps, each component within 1e-5 relative L2 (the bench's precision
workload with shallow sources, 40-80 m deep, receivers within 150 m of their
epicentres and the wavelets centred at 30 ms, so that the traces carry the
arrivals within the 0.16 s the run spans).
"""
import json
import os

import numpy as np
import pytest

from common import ROOT, relerr
from sw4lite_amd import lib as L

C1 = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_pointsource.json")))
C1_WANT = (C1["linf"], C1["l2"], C1["solinf"])
C1_GRID = dict(nx=121, ny=121, nz=121, h=0.05, bc=[L.BC_SUPERGRID] * 6, sg_gp=20, rho=1.0, cs=1.0,
               cp=np.sqrt(3.0), tmax=0.6)
C1_SRC = dict(x=3.0, y=3.0, z=3.0, m=[1.0, 0.3, -0.2, 0.8, 0.4, 0.6], t0=0.32, freq=12.5, tfkind=L.TF_GAUSSIAN)
C1_TOL = 1e-11


def test_c1_deck_is_the_workload():
    deck = C1["deck"]
    assert "grid x=6.0 y=6.0 z=6.0 h=0.05" in deck and "supergrid gp=20" in deck and "time t=0.6" in deck
    assert "testpointsource\n" in deck and "cfl=1.3" in deck
    assert "Mxx=1 Myy=0.8 Mzz=0.6 Mxy=0.3 Mxz=-0.2 Myz=0.4 t0=0.32 freq=12.5 type=Gaussian" in deck
    assert C1["box"] == [-1, 123, -1, 123, -1, 123] and C1["nsteps"] == 21
    assert [float(v) for v in C1["printed"][1:]] == pytest.approx(list(C1_WANT), rel=1e-11)


def c1_oracle():
    from oracle.loop import OracleRun
    d = C1_GRID
    mu = d["rho"] * d["cs"] ** 2
    la = d["rho"] * d["cp"] ** 2 - 2 * mu
    run = OracleRun(d["nx"], d["ny"], d["nz"], d["h"], d["bc"], d["rho"], mu, la, [C1_SRC], sg_gp=d["sg_gp"],
                    tmax=d["tmax"])
    return run


def test_c1_cpu_oracle():
    run = c1_oracle()
    assert run.nsteps == C1["nsteps"] and run.dt == pytest.approx(C1["dt"], rel=1e-15)
    run.step(run.nsteps)
    li, l2, xli, _ = run.point_source_error(C1_SRC)
    for g, w in zip((li, l2, xli), C1_WANT):
        assert abs(g - w) <= C1_TOL * abs(w), ((li, l2, xli), C1_WANT)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_c1_gpu(mode):
    from sw4lite_amd import EW
    d = C1_GRID
    print("C1 121^3 Gaussian point source, mode", mode, flush=True)
    e = EW(d["nx"], d["ny"], d["nz"], d["h"], d["bc"], corder=1, sg_gp_thickness=d["sg_gp"])
    try:
        e.set_material_const(d["rho"], d["cs"], d["cp"])
        s = C1_SRC
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(tmax=d["tmax"])
        inf = e.info()
        assert inf.nsteps == C1["nsteps"] and inf.dt == pytest.approx(C1["dt"], rel=1e-15)
        e.timesteploop(mode=mode)
        got = e.point_source_error()
        u, _ = e.get_solution()
    finally:
        e.close()
    print("C1 mode", mode, got, C1_WANT, flush=True)
    for g, w in zip(got, C1_WANT):
        assert abs(g - w) <= C1_TOL * abs(w), (got, C1_WANT)
    if mode == 0:
        run = c1_oracle()
        run.step(run.nsteps)
        assert relerr(u, run.U) < 1e-12


@pytest.mark.gpu
def test_c3_reduced_matches_oracle():
    import torch
    import bench
    from oracle.loop import OracleRun
    from sw4lite_amd import EW
    nx, ny, nz, h, nsteps = 64, 64, 48, 20.0, 10
    _, _, z = bench.hills_depth(torch, nx, ny, nz, h, seed=3)
    rho, mu, la = (a.cpu().numpy() for a in bench.make_material(torch, (nx, ny, nz), h, seed=3, z=z))
    z = z.cpu().numpy()
    bc = [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID]
    src = [dict(x=0.45 * (nx - 1) * h, y=0.55 * (ny - 1) * h, z=0.5 * (nz - 1) * h,
                m=list(bench.double_couple(np.random.default_rng(3), 1e15)), t0=0.06, freq=4.0, tfkind=L.TF_RICKER)]
    print("C3 reduced 64x64x48 curvilinear, %d steps" % nsteps, flush=True)
    ref = OracleRun(nx, ny, nz, h, bc, rho, mu, la, src, corder=1, sg_gp=8, nsteps=nsteps, zgrid=z)
    e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=8)
    try:
        e.set_material(rho, mu, la)
        e.set_topography(z)
        for s in src:
            e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        e.setup(nsteps=nsteps)
        assert abs(e.info().dt - ref.dt) < 1e-13 * ref.dt
        e.timesteploop(nsteps, mode=1)
        u, um = e.get_solution()
    finally:
        e.close()
    ref.step(nsteps)
    err = relerr(u, ref.U)
    print("C3 reduced: relerr", err, "max|u|", np.max(np.abs(ref.U)), flush=True)
    assert np.max(np.abs(ref.U)) > 0 and err < 1e-12 and relerr(um, ref.Um) < 1e-12


@pytest.mark.gpu
def test_c3_full_size_properties():
    import torch
    import bench
    from sw4lite_amd import EW
    nx, ny, nz, h, nsteps = 384, 384, 256, 20.0, 3
    bc = [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID]
    m = np.array(bench.double_couple(np.random.default_rng(3), 1e15))
    res = []
    for scale in (1.0, 2.0):
        print("C3 384x384x256 curvilinear, moment x%g, %d steps" % (scale, nsteps), flush=True)
        _, _, z = bench.hills_depth(torch, nx, ny, nz, h, seed=3)
        rho, mu, la = bench.make_material(torch, (nx, ny, nz), h, seed=3, z=z)
        e = EW(nx, ny, nz, h, bc, corder=1, sg_gp_thickness=30)
        try:
            e.set_material(rho, mu, la)
            e.set_topography(z)
            del rho, mu, la, z
            e.add_moment_source(0.5 * (nx - 1) * h, 0.5 * (ny - 1) * h, 0.4 * (nz - 1) * h, list(scale * m), 0.0, 4.0,
                                L.TF_RICKER)
            e.setup(nsteps=400)
            e.timesteploop(nsteps, mode=1)
            res.append(e.get_solution()[0])
        finally:
            e.close()
            torch.cuda.empty_cache()
    u1, u2 = res
    assert np.all(np.isfinite(u1)) and np.max(np.abs(u1)) > 0
    dev = np.max(np.abs(u2 - 2 * u1)) / np.max(np.abs(u2))
    print("C3 full size: max|u| %.6g, nonzero points %d, |u(2M) - 2 u(M)| / max %.3g" %
          (np.max(np.abs(u1)), int(np.count_nonzero(u1)), dev), flush=True)
    assert dev <= 1e-15


@pytest.mark.gpu
def test_c4_fp32_against_fp64_1024():
    import torch
    import bench
    steps = 100
    print("C4 1024^3 FP32 compensated vs FP64, %d steps, 64 receivers" % steps, flush=True)
    # shallow sources and near receivers, so that the waves reach the
    # receivers within the 0.16 s the run spans
    traces, secs, dt, recs = bench.precision_traces(torch, 1024, steps, t0=0.03, log=False, depth=(40.0, 80.0),
                                                    radius=150.0)
    ref, got = traces[0], traces[2]
    assert ref.shape == (steps + 1, 64, 3)
    # every component carries signal well above the numerical precursors
    assert min(float(np.max(np.abs(ref[..., c]))) for c in range(3)) > 1e-6
    rel_l2 = [float(np.linalg.norm(got[..., c] - ref[..., c]) / np.linalg.norm(ref[..., c])) for c in range(3)]
    max_err = [float(np.max(np.abs(got[..., c] - ref[..., c])) / np.max(np.abs(ref[..., c]))) for c in range(3)]
    print("C4: relative L2 per component", rel_l2, "max error", max_err, "FP64 %.1f s FP32 %.1f s" % (secs[0], secs[2]),
          flush=True)
    assert max(rel_l2) <= 1e-5
