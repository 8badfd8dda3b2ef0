#!/usr/bin/env python3
"""Throughput of the sw4lite elastic-wave time step on one MI355X.

Workload (the reference path at scale): FP64 Cartesian 4th-order SBP solver
on a 512^3 grid, h = 20 m; layered material cs = 1000 -> 3500 m/s linear in
depth, cp = 1.8 cs, rho = 1700 (cp / 1 km/s)^0.25 kg/m^3, times a seeded
lognormal perturbation exp(0.05 g) of cs and cp, g a unit-variance field of
white noise smoothed by three passes of a 17-point box filter per axis (a
near-Gaussian of 8.5 cells); free surface on top, supergrid
of 30 points on the other five faces; four seeded random double-couple
sources with 5 Hz Ricker time functions.  All inputs are synthetic and built
on the GPU before timing.

A step is one full time step (predictor sweep, sources, ghost-point BC,
corrector + supergrid sweep, sources, BC).  The metric is interior grid-point
updates per second: 512^3 * steps / seconds, inputs resident in HBM.

The JSON line adds
  roofline: the dominant kernel (the slower interior sweep), its
    algorithmic bytes per launch (predictor: 12 FP64 words per point, read
    U, Um (3 components each), mu, lambda, rho, write Up) over its average
    launch time from hipEvents on the stream it runs on, against 8 TB/s;
    traffic is the HBM bytes per launch measured with rocprofv3 counters;
    moved_* the bytes the kernel moves by design (with the stored
    acceleration the predictor also writes A: 15 words);
  step_roofline: the whole step on its 27 algorithmic words (216 B) per
    point against 8 TB/s;
  cpu_baseline: the same step on the host cores for a bounded slab of the
    workload, with the stencil stages (rhs4sg_rev, addsgd4) from the
    reference's own sources (oracle/_ref) when they are present.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt updates/s per RK step (FP64 vs FP32-comp, %HBM roofline); adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
CORR_BYTES_PER_PT = 15 * 8     # corrector sweep: Up, U, Um, mu, lambda, rho in, Up' out (FP64)
PRED_BYTES_PER_PT = 12 * 8     # predictor sweep, FP64
# kernel variant 8 (even row length): the predictor also writes the
# acceleration A, the corrector reads A, Up, mu, lambda, rho and writes Up'
PRED_ACC_BYTES_PER_PT = 15 * 8
CORR_ACC_BYTES_PER_PT = 12 * 8
STEP_BYTES_PER_PT = 27 * 8     # whole step, compulsory traffic
# FP32 form with the compensated update (float words): the predictor reads U,
# mu, lambda, rho and writes A; the corrector (damping included) reads A, D,
# U, C, mu, lambda, rho and writes Dn, U, C
COMP_CORR_BYTES_PER_PT = 24 * 4
COMP_PRED_BYTES_PER_PT = 9 * 4
COMP_STEP_BYTES_PER_PT = 33 * 4
STAGES = ["pred_interior", "pred_boundary_rows_shell", "pred_sources_bc", "corr_interior",
          "corr_interior_damping", "corr_boundary_rows_shell", "corr_sources_bc"]


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None, help="timed steps (default 100; 800 for the precision workload)")
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=None, help="grid points per axis (default 512; 1024 for the precision workload)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp32", action="store_true", help="skip the FP32 compensated form")
    p.add_argument("--cpu-nz", type=int, default=192, help="depth of the CPU-baseline slab")
    p.add_argument("--cpu-steps", type=int, default=12)
    p.add_argument("--variant", type=int, default=10, help="interior sweep kernels: 0 thread/point, 1 LDS k-march, 2 with prefetch, "
                        "3 prefetch without register cap, 4 as 3 with the damping in the corrector sweep, "
                        "5 corrector as 4 and predictor as 1, 6 as 1 with an LDS damping sweep, "
                        "7 as 6 with the predictor planes loaded asynchronously into LDS (even row lengths), "
                        "8 as 7 with the acceleration stored by the predictor and staged by the corrector, "
                        "9 as 8 with the damping sweep only over the supergrid layers, "
                        "10 as 9 with rows of odd length staged by four-byte asynchronous loads")
    p.add_argument("--kchunk", type=int, default=0,
                   help="planes per workgroup of the LDS sweeps (0: the library's defaults, 64 for the predictor "
                        "and corrector sweeps, 32 for the damping sweep)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also launch a plain device copy of 1 GiB (PMC scaling)")
    p.add_argument("--workload", choices=["cartesian", "curvilinear", "precision", "adjoint", "batch",
                                          "batch-adjoint"],
                   default="cartesian",
                   help="cartesian: the full time step at N^3 (default); curvilinear: the curvilinear-grid "
                        "kernels (rhs4sgcurv, addsgd4c, free-surface closure) on a 384x384x256 grid under "
                        "Gaussian hills; precision: FP32 compensated against FP64 on 64 surface receivers "
                        "(1024^3, h=10 m, 800 steps); adjoint: gradient of a receiver misfit against the forward run "
                        "(256^3, h=25 m, 600 steps, 128 surface receivers); batch: 48 shots of N^3, N uniform in "
                        "128..288, h=25 m, stepped alone, batched, batched with graph replay; batch-adjoint: the "
                        "gradients of every shot of the batch workload against the batched forward run")
    p.add_argument("--ckpt", type=int, default=0,
                   help="adjoint workload: steps between checkpoints (0: from free HBM, the whole history when it fits)")
    p.add_argument("--shots", type=int, default=48, help="batch workloads: number of shots")
    p.add_argument("--concurrent", type=int, default=4, help="batch-adjoint workload: shots whose gradients run at once")
    a = p.parse_args()
    if a.n is None:
        a.n = {"precision": 1024, "adjoint": 256}.get(a.workload, 512)
    if a.steps is None:
        a.steps = {"precision": 800, "adjoint": 600, "batch": 30, "batch-adjoint": 100}.get(a.workload, 100)
    return a


def make_material(torch, n, h, seed, device="cuda", z=None, width=17):
    """rho, mu, lambda over the box of n = N or (nx, ny, nz) interior points
    (two ghost points per side); depth from the grid depth field z (same box,
    flattened) when given, else z = (k-1) h.  The perturbation is smoothed by
    three passes of a `width`-point box filter per axis (17: a near-Gaussian
    of 8.5 cells)."""
    dev = torch.device(device)
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    shp = (nz + 4, ny + 4, nx + 4)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    g = torch.randn(shp, generator=gen, device=dev, dtype=torch.float64)
    F = torch.nn.functional
    W = width

    def box(x):  # along the last axis, zero padded
        c = F.pad(x, (W // 2 + 1, W // 2)).cumsum(-1)
        return (c[..., W:] - c[..., :-W]) / W

    for axis in range(3):
        g = g.movedim(axis, 2)
        for _ in range(3):
            g = box(g)
        g = g.movedim(2, axis)
    g = (g - g.mean()) / g.std()
    p = torch.exp(0.05 * g)
    del g
    if z is None:
        k = torch.arange(-1, nz + 3, device=dev, dtype=torch.float64)
        depth = ((k - 1) * h).view(-1, 1, 1)
    else:
        depth = z.view(shp).clamp(min=0.0)
    cs = (1000.0 + 2500.0 * depth / ((nz - 1) * h)) * p
    cp = 1.8 * cs
    rho = 1700.0 * (cp / 1000.0) ** 0.25
    mu = rho * cs * cs
    la = rho * cp * cp - 2 * mu
    return rho.contiguous().view(-1), mu.contiguous().view(-1), la.contiguous().view(-1)


def double_couple(rng, m0):
    """Moment tensor of a random strike/dip/rake double couple (Aki & Richards 4.88)."""
    st, dp, rk = rng.uniform(0, 2 * np.pi), rng.uniform(0.2, 0.5 * np.pi), rng.uniform(-np.pi, np.pi)
    nrm = np.array([-np.sin(dp) * np.sin(st), np.sin(dp) * np.cos(st), -np.cos(dp)])
    slip = np.array([np.cos(rk) * np.cos(st) + np.cos(dp) * np.sin(rk) * np.sin(st),
                     np.cos(rk) * np.sin(st) - np.cos(dp) * np.sin(rk) * np.cos(st),
                     -np.sin(rk) * np.sin(dp)])
    M = m0 * (np.outer(nrm, slip) + np.outer(slip, nrm))
    return [M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]]


def sources(n, h, seed, nsrc=4):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsrc):
        x = rng.uniform(40 * h, (n - 41) * h)
        y = rng.uniform(40 * h, (n - 41) * h)
        z = rng.uniform(6 * h, 0.5 * (n - 1) * h)
        out.append(dict(x=x, y=y, z=z, m=double_couple(rng, 1e15), t0=0.24, freq=5.0, tfkind=0))
    return out


def cpu_baseline(n, h, nz, steps, srcs):
    """Bounded CPU sample: an n x n x nz slab of the workload (same boundary
    types, layered material without the random perturbation, the sources that
    fall inside).  Preferred: the reference program itself (oracle/_ref,
    built from the reference's sources) running the slab's deck on the host
    cores, timed as the difference of a 1-step and a (1 + steps)-step run;
    otherwise the loop of oracle/loop.py with the reference's stencils."""
    import subprocess
    import tempfile
    from oracle import refbench
    model, ncpu, avail = refbench.host_description()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    inside = [s for s in srcs if s["z"] < (nz - 6) * h] or [dict(srcs[0], z=0.5 * (nz - 1) * h)]
    host = "%s, nproc %d, %d usable, OMP threads %d" % (model, ncpu, avail, threads)
    from oracle import refew
    if refew.available():
        with tempfile.TemporaryDirectory(prefix="cpubase_") as tmp:
            fn = os.path.join(tmp, "srcs.json")
            json.dump(inside, open(fn, "w"))
            try:
                r = subprocess.run([sys.executable, "-m", "oracle.refbench", str(n), str(n), str(nz), repr(h),
                                    repr((n - 1) * h), str(steps), fn], capture_output=True, text=True,
                                   timeout=300, cwd=ROOT)
                res = json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
            except (subprocess.TimeoutExpired, ValueError, IndexError):
                res = None
        if res is not None:
            return {
                "value": n * n * nz * steps / res["seconds"],
                "unit": "grid-pt updates/s",
                "cores": threads,
                "kind": "reference",
                "sample": "the reference program (its own time-step loop and OpenMP C kernels, built from its "
                          "sources) on a %dx%dx%d slab of the workload (layered material as linear blocks, no "
                          "perturbation, free top, supergrid 30, %d Ricker sources): %d steps timed as the "
                          "difference of a %d-step and a %d-step run (%.1f s and %.1f s); host %s"
                          % (n, n, nz, len(inside), steps, 1, 1 + steps, res["runs"][0], res["runs"][1], host),
                "cpu_model": model, "nproc": ncpu,
            }
    from oracle.loop import OracleRun
    from oracle import oracle as O
    from sw4lite_amd import lib as L
    kernels = "reference" if O.reference() is not None else "oracle"
    box = (-1, n + 2, -1, n + 2, -1, nz + 2)
    k = np.arange(box[4], box[5] + 1)
    z = (k - 1) * h
    cs = (1000.0 + 2500.0 * z / ((n - 1) * h))[:, None, None] * np.ones((1, n + 4, n + 4))
    cp = 1.8 * cs
    rho = 1700.0 * (cp / 1000.0) ** 0.25
    mu = rho * cs * cs
    la = rho * cp * cp - 2 * mu
    bc = [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID]
    run = OracleRun(n, n, nz, h, bc, rho.ravel(), mu.ravel(), la.ravel(), inside, sg_gp=30, nsteps=steps + 1,
                    kernels=kernels)
    run.step(1)
    t0 = time.perf_counter()
    run.step(steps)
    dt = time.perf_counter() - t0
    return {
        "value": n * n * nz * steps / dt,
        "unit": "grid-pt updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "%dx%dx%d slab of the workload (layered material, no perturbation), %d timed steps after 1 "
                  "warm-up, %.1f s; stencil stages %s, pointwise stages from the oracle restatement; host %s"
                  % (n, n, nz, steps, dt, "from the reference sources" if kernels == "reference" else "oracle",
                     host),
        "cpu_model": model, "nproc": ncpu,
    }


CURV_FLOP_INTERIOR = 2126   # rhs4sgcurv.C:47-50, arithmetic ops per interior point
CURV_FLOP_BOUNDARY = 6049   # per point of the six SBP rows
CURV_WORDS = 13             # u(3), mu, lambda, met(4), jac in; lu(3) out
FP64_PEAK_TFLOPS = 78.6


def hills_depth(torch, nx, ny, nz, h, seed, nhill=6):
    """Grid depth over the box: the top is a sum of nhill seeded Gaussian
    hills (width 20-60 cells, total height 15% of the depth), mapped linearly
    in k to the flat bottom z = (nz-1) h."""
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    depth = (nz - 1) * h
    x = (torch.arange(-1, nx + 3, device=dev, dtype=torch.float64) - 1) * h
    y = (torch.arange(-1, ny + 3, device=dev, dtype=torch.float64) - 1) * h
    Y, X = torch.meshgrid(y, x, indexing="ij")
    top = torch.zeros_like(X)
    for _ in range(nhill):
        xc, yc = rng.uniform(0.15, 0.85) * (nx - 1) * h, rng.uniform(0.15, 0.85) * (ny - 1) * h
        w = rng.uniform(20, 60) * h
        top -= rng.uniform(0.3, 1.0) * torch.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2 * w * w))
    top *= 0.15 * depth / top.abs().max()
    k = torch.arange(-1, nz + 3, device=dev, dtype=torch.float64)
    zeta = ((k - 1) / (nz - 1)).view(-1, 1, 1)
    z = top.unsqueeze(0) * (1 - zeta) + depth * zeta
    xs = X.unsqueeze(0).expand(nz + 4, -1, -1).contiguous().view(-1)
    ys = Y.unsqueeze(0).expand(nz + 4, -1, -1).contiguous().view(-1)
    return xs, ys, z.contiguous().view(-1)


def cpu_baseline_curv(nx, ny, nz_slab):
    """The oracle's rhs4sgcurv (OpenMP) on a bounded slab of the workload's
    shape: nx x ny x nz_slab interior points under one hill, 3 evaluations."""
    from oracle import oracle as O
    from sw4lite_amd import ew
    acof, ghcof, bope, _ = ew.stencil_coefficients()
    h = 20.0
    ib, ie, jb, je, kb, ke = -1, nx + 2, -1, ny + 2, -1, nz_slab + 2
    box = (ib, ie, jb, je, kb, ke)
    shp = (ke - kb + 1, je - jb + 1, ie - ib + 1)
    n = shp[0] * shp[1] * shp[2]
    x = (np.arange(ib, ie + 1) - 1) * h
    y = (np.arange(jb, je + 1) - 1) * h
    X, Y = np.meshgrid(x, y)
    top = -0.1 * 255 * h * np.exp(-((X - x.mean()) ** 2 + (Y - y.mean()) ** 2) / (2 * (40 * h) ** 2))
    zeta = (np.arange(kb, ke + 1) - 1) / 255.0
    z = np.ascontiguousarray((top[None] * (1 - zeta[:, None, None]) + 255 * h * zeta[:, None, None]).reshape(-1))
    met, jac = np.zeros(4 * n), np.zeros(n)
    O.metric(1, box, h, z, met, jac)
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, 3 * n)
    mu, la = np.full(n, 3e10), np.full(n, 4e10)
    strx, stry = np.ones(shp[2]), np.ones(shp[1])
    lu = O.filled(3 * n)
    O.rhs4sgcurv(1, box, u, mu, la, met, jac, lu, [0, 0, 0, 0, 1, 0], acof, bope, ghcof, strx, stry)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        O.rhs4sgcurv(1, box, u, mu, la, met, jac, lu, [0, 0, 0, 0, 1, 0], acof, bope, ghcof, strx, stry)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nx * ny * nz_slab * reps / dt, "unit": "grid-pt evaluations/s", "cores": threads,
            "kind": "port",
            "sample": "oracle rhs4sgcurv_rev restatement, %dx%dx%d slab under one hill, top SBP rows, %d "
                      "evaluations after 1 warm-up, %.1f s, OMP threads %d" % (nx, ny, nz_slab, reps, dt, threads)}


def main_curvilinear(a, torch):
    from sw4lite_amd import ew
    from sw4lite_amd import lib as L
    nx, ny, nz, h = 384, 384, 256, 20.0
    box = (-1, nx + 2, -1, ny + 2, -1, nz + 2)
    npts = (nx + 4) * (ny + 4) * (nz + 4)
    dev = torch.device("cuda")
    x, y, z = hills_depth(torch, nx, ny, nz, h, seed=3)
    met = torch.empty(4 * npts, dtype=torch.float64, device=dev)
    jac = torch.empty(npts, dtype=torch.float64, device=dev)
    ew.metric(1, box, x, y, z, met, jac)
    del x, y
    rho, mu, la = make_material(torch, (nx, ny, nz), h, seed=3, z=z)
    del z
    desc = L.grid_desc(nx, ny, nz, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1,
                       sg_gp_thickness=30)
    sg = {k: torch.from_numpy(v).to(dev) for k, v in ew.host_supergrid(desc).items()}
    acof, ghcof, bope, sbop = ew.stencil_coefficients()
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)
    u = 1e-3 * torch.randn(3 * npts, generator=gen, device=dev, dtype=torch.float64)
    um = u + 1e-5 * torch.randn(3 * npts, generator=gen, device=dev, dtype=torch.float64)
    lu = torch.zeros_like(u)
    up = torch.zeros_like(u)
    onesided = [0, 0, 0, 0, 1, 0]

    def rhs():
        ew.rhs4sgcurv(1, box, u, mu, la, met, jac, lu, onesided, acof, bope, ghcof, sg["strx"], sg["stry"])

    def damp():
        ew.addsgd4c(1, box, up, u, um, rho, sg["dcx"], sg["dcy"], sg["strx"], sg["stry"], jac, sg["cox"], sg["coy"],
                    0.02)

    def bc():
        ew.freesurfcurvisg(1, box, nz, 5, u, mu, la, met, sbop, None, sg["strx"], sg["stry"])

    def timed(fn, reps):
        for _ in range(a.warmup):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t0 = time.perf_counter()
    ms_rhs = timed(rhs, a.steps)
    ms_damp = timed(damp, a.steps)
    ms_bc = timed(bc, a.steps)
    del lu, up, um
    # the full time step on the same grid: EW in the reference stage order
    from sw4lite_amd import EW
    e = EW(nx, ny, nz, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
    x, y, z = hills_depth(torch, nx, ny, nz, h, seed=3)
    del x, y
    e.set_material(rho, mu, la)
    e.set_topography(z)
    del z
    m = double_couple(np.random.default_rng(3), 1e15)
    e.add_moment_source(0.5 * (nx - 1) * h, 0.45 * (ny - 1) * h, 0.4 * (nz - 1) * h, m, 0.3, 4.0, 0)
    e.setup(nsteps=400)
    e.timesteploop(a.warmup, mode=1)
    e.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nstep = max(1, a.steps // 2)
    e.timesteploop(nstep, mode=1)
    e.sync()
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t1) / nstep
    dt_curv = e.info().dt
    e.close()
    wall = time.perf_counter() - t0
    pts = float(nx) * ny * nz
    flop = float(nx) * ny * ((nz - 6) * CURV_FLOP_INTERIOR + 6 * CURV_FLOP_BOUNDARY)
    tflops = flop / (ms_rhs * 1e-3) / 1e12
    gbs = pts * CURV_WORDS * 8 / (ms_rhs * 1e-3) / 1e9
    line = {
        "metric": "rhs4sgcurv grid-pt evaluations/s (curvilinear divergence of stress, FP64)",
        "value": pts / (ms_rhs * 1e-3),
        "unit": "grid-pt evaluations/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_rhs,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "%dx%dx%d curvilinear grid under 6 seeded Gaussian hills (15%% of depth, widths 20-60 "
                        "cells), linear vertical stretching, h=20 m, layered+lognormal material, top SBP rows; "
                        "one rhs4sgcurv evaluation per step" % (nx, ny, nz),
            "addsgd4c_ms": ms_damp, "freesurfcurvisg_ms": ms_bc, "wall_s": wall,
            "hbm_GBps_at_13_words": gbs,
            "full_step": {"ms_per_step": step_s * 1e3, "grid_pt_updates_per_s": pts / step_s, "steps": nstep,
                          "dt": dt_curv, "order": "reference stage order (Force, rhs4sgcurv, predictor, BC, "
                                                  "Force_tt, dpdmt, rhs4sgcurv, corrector, addsgd4c, BC)",
                          "source": "one double couple, Ricker 4 Hz"},
        },
        "roofline": {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_rhs4curv",
                     "flop_per_launch": flop, "avg_launch_ms": ms_rhs},
        "cpu_baseline": None,
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_curv(nx, ny, 24)
    print(json.dumps(line), flush=True)


def bench_fp32(a, e, torch, n, fp64_value):
    """The same steps in the FP32 form with the compensated time update
    (EW step mode 2), continuing from the FP64 state."""
    e.timesteploop(a.warmup, mode=2)
    e.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stage_ms = e.timesteploop_timed(a.steps, fp32=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pts = float(n) ** 3
    value = pts * a.steps / wall
    corr_pts = float(n) * n * (n - 6)
    corr_ms = stage_ms[3] / a.steps
    pred_ms = stage_ms[0] / a.steps
    achieved = corr_pts * COMP_CORR_BYTES_PER_PT / (corr_ms * 1e-3) / 1e9
    return {
        "value": value, "unit": "grid-pt updates/s", "ms_per_step": 1e3 * wall / a.steps,
        "dtype": "f32 stencils, compensated f32 pair update, f64 boundary closure and sources",
        "speedup_vs_fp64": value / fp64_value,
        "step_hbm_fraction": value * COMP_STEP_BYTES_PER_PT / (HBM_PEAK_GBS * 1e9),
        "stage_ms_per_step": {k: v / a.steps for k, v in zip(STAGES, stage_ms.tolist())},
        "pred_interior_GBps": corr_pts * COMP_PRED_BYTES_PER_PT / (pred_ms * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "kernel": "k_comp_dma<true>" if a.variant >= 7 and (n + 4) % 4 == 0 else "k_comp_sweep<true>",
                     "bytes_per_launch": corr_pts * COMP_CORR_BYTES_PER_PT, "avg_launch_ms": corr_ms},
    }


def precision_traces(torch, n, steps, t0=0.24, log=True, depth=(200.0, 800.0), radius=1500.0):
    """The precision workload at n^3 (h = 10 m * 1024/n, the smoothing length
    of the perturbation kept at 170 m), run once with the FP64 fused step and
    once with the FP32 form with the compensated update from the same zero
    initial state; 64 seeded surface receivers record displacement after
    every step.  The four double couples (Ricker 5 Hz centred at t0) sit
    depth[0]..depth[1] m deep (200-800) and the receivers within radius m
    (1.5 km) of their epicentres.
    Returns ({0: FP64 traces, 2: FP32 traces} [steps+1, 64, 3], seconds per
    mode, dt, receivers)."""
    from sw4lite_amd import EW
    from sw4lite_amd import lib as L
    h = 10.0 * 1024 / n
    span = (n - 1) * h
    rng = np.random.default_rng(4)
    srcs = []
    for _ in range(4):
        x, y = rng.uniform(0.2 * span, 0.8 * span), rng.uniform(0.2 * span, 0.8 * span)
        srcs.append(dict(x=x, y=y, z=rng.uniform(*depth), m=double_couple(rng, 1e15), t0=t0, freq=5.0,
                         tfkind=0))
    recs = []
    for r in range(64):
        s0 = srcs[r % 4]
        rad, ang = rng.uniform(0.0, radius), rng.uniform(0.0, 2 * np.pi)
        recs.append((s0["x"] + rad * np.cos(ang), s0["y"] + rad * np.sin(ang), 0.0))
    traces, secs = {}, {}
    dt = None
    for mode in (0, 2):
        rho, mu, la = make_material(torch, n, h, seed=4, width=33)
        torch.cuda.empty_cache()
        e = EW(n, n, n, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
        e.set_material(rho, mu, la)
        del rho, mu, la
        torch.cuda.empty_cache()
        for s in srcs:
            e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
        for x, y, z in recs:
            e.add_receiver(x, y, z, 0)
        e.setup(nsteps=steps)
        dt = e.info().dt
        done, c0 = 0, time.perf_counter()
        while done < steps:
            chunk = min(100, steps - done)
            e.timesteploop(chunk, mode=mode)
            e.sync()
            done += chunk
            if log:
                sys.stderr.write("precision: mode %d, %d/%d steps, %.1f s\n" % (mode, done, steps,
                                                                              time.perf_counter() - c0))
                sys.stderr.flush()
        secs[mode] = time.perf_counter() - c0
        traces[mode] = e.receivers()
        e.close()
        torch.cuda.empty_cache()
    return traces, secs, dt, recs


def main_precision(a, torch):
    """The precision workload (precision_traces at n^3, a.steps steps).
    Reported: per component, the relative L2 error of the FP32 traces against
    the FP64 traces over all receivers and samples, and the maximum error
    relative to the largest FP64 amplitude."""
    n, steps = a.n, a.steps
    h = 10.0 * 1024 / n
    traces, secs, dt, recs = precision_traces(torch, n, steps)
    ref, got = traces[0], traces[2]
    d = got - ref
    rel_l2 = [float(np.linalg.norm(d[..., c]) / np.linalg.norm(ref[..., c])) for c in range(3)]
    max_err = [float(np.max(np.abs(d[..., c])) / np.max(np.abs(ref[..., c]))) for c in range(3)]
    pts = float(n) ** 3
    line = {
        "metric": "FP32-compensated seismogram error vs FP64 (relative L2, worst component)",
        "value": max(rel_l2), "unit": "relative L2", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * secs[2] / steps, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 stencils, compensated f32 pair update, f64 boundary closure and sources",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 Cartesian step, h=%g m, layered+lognormal material (smoothing 170 m), free top, "
                        "supergrid 30, 4 double couples 200-800 m deep (Ricker 5 Hz), 64 surface receivers "
                        "within 1.5 km of the epicentres" % (n, h),
            "grid": [n, n, n], "h": h, "dt": dt, "receivers": len(recs),
            "rel_l2_per_component": rel_l2, "max_err_per_component": max_err, "tolerance": 1e-5,
            "fp64_ms_per_step": 1e3 * secs[0] / steps, "fp32_ms_per_step": 1e3 * secs[2] / steps,
            "fp64_grid_pt_updates_per_s": pts * steps / secs[0], "fp32_grid_pt_updates_per_s": pts * steps / secs[2],
            "max_abs_fp64_trace": float(np.max(np.abs(ref))),
        },
    }
    print(json.dumps(line), flush=True)


def main_adjoint(a, torch):
    """One shot of the batched-adjoint workload: n^3 grid at h = 25 m with the
    workload material, a shallow double couple (Ricker 5 Hz), 128 seeded
    surface receivers, `steps` steps.  Times the forward run (fused FP64
    steps recording the receivers) and the gradient of J = 1/2 |d - d_obs|^2
    with respect to rho, mu, lambda (sw4hip_ew_gradient: forward with
    checkpoints every --ckpt steps, reverse sweep recomputing one segment at
    a time; --ckpt 0 lets the library choose the interval from free HBM,
    which at 256^3 x 600 steps holds the whole history, and the run is then
    repeated with checkpoints every 50 steps for comparison).  The library
    keeps its gradient workspace between calls: each setting is run twice
    and the second call timed (the first, which allocates, is reported).
    value = gradient seconds / forward seconds."""
    from sw4lite_amd import EW
    from sw4lite_amd import lib as L
    n, h, steps = a.n, 25.0, a.steps
    span = (n - 1) * h
    rng = np.random.default_rng(6)
    rho, mu, la = make_material(torch, n, h, seed=6)
    torch.cuda.empty_cache()
    e = EW(n, n, n, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
    e.set_material(rho, mu, la)
    e.sync()
    e.add_moment_source(0.5 * span, 0.5 * span, 500.0, double_couple(rng, 1e15), 0.24, 5.0, 0)
    for _ in range(128):
        e.add_receiver(rng.uniform(0.15 * span, 0.85 * span), rng.uniform(0.15 * span, 0.85 * span), 0.0, 0)
    e.setup(nsteps=steps)
    e.reset()
    e.timesteploop(min(5, steps), mode=0)
    e.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.reset()
    e.timesteploop(steps, mode=0)
    e.sync()
    t_fwd = time.perf_counter() - t0
    tr = e.receivers()
    dobs = 0.9 * tr
    out = [torch.empty_like(rho), torch.empty_like(mu), torch.empty_like(la)]
    used, nbytes = e.gradient_plan(steps, a.ckpt)

    def timed_gradient(ck):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.gradient(tr - dobs, steps, ckpt=ck, out=out)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # the first call allocates the workspace the library keeps for later
    # gradients (an inversion computes many); the second one is timed
    t_first = timed_gradient(a.ckpt)
    t_grad = timed_gradient(a.ckpt)
    gn = [float(o.abs().max()) for o in out]
    fixed = {}
    if a.ckpt <= 0 and used != min(50, steps):
        g_auto = [o.clone() for o in out]
        c50 = min(50, steps)
        e.release_workspace()
        timed_gradient(c50)
        t50 = timed_gradient(c50)
        fixed = {"checkpoint_every": c50, "device_bytes": e.gradient_plan(steps, c50)[1], "gradient_s": t50,
                 "ratio": t50 / t_fwd,
                 "max_rel_diff_vs_first": max(float((x - y).abs().max() / y.abs().max()) for x, y in zip(out, g_auto))}
    fp32 = {}
    if not a.no_fp32:
        # the FP32 form: its forward run, and its gradient (forward and
        # recomputation in FP32 from float checkpoints, reverse in FP64)
        e.release_workspace()
        c32 = a.ckpt if a.ckpt > 0 else min(50, steps)
        e.reset()
        e.timesteploop(min(5, steps), mode=2)
        e.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.reset()
        e.timesteploop(steps, mode=2)
        e.sync()
        t_fwd32 = time.perf_counter() - t0
        tr32 = e.receivers()

        def timed_gradient32():
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            e.gradient(tr32 - dobs, steps, ckpt=c32, out=out, fp32=True)
            torch.cuda.synchronize()
            return time.perf_counter() - t1

        timed_gradient32()
        t_g32 = timed_gradient32()
        fp32 = {"checkpoint_every": c32, "forward_s": t_fwd32, "gradient_s": t_g32, "ratio": t_g32 / t_fwd32,
                "gradient_over_fp64_forward": t_g32 / t_fwd,
                "trace_rel_l2_vs_fp64": float(np.linalg.norm(tr32 - tr) / np.linalg.norm(tr))}
    e.close()
    pts = float(n) ** 3
    line = {
        "metric": "adjoint/forward cost ratio (gradient of a receiver misfit w.r.t. rho, mu, lambda)",
        "value": t_grad / t_fwd, "unit": "ratio", "n_gpus": 1, "steps": steps, "warmup": min(5, steps),
        "ms_per_step": 1e3 * t_grad / steps, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "%d^3 Cartesian FP64 step, h=%g m, layered+lognormal material, free top, supergrid 30, "
                        "one double couple at 500 m (Ricker 5 Hz), 128 surface receivers, one shot" % (n, h),
            "grid": [n, n, n], "h": h, "checkpoint_every": used, "checkpoint_arg": a.ckpt,
            "gradient_device_bytes": nbytes, "forward_s": t_fwd, "gradient_s": t_grad,
            "gradient_first_call_s": t_first,
            "fixed_checkpoints": fixed,
            "fp32_form": fp32,
            "forward_grid_pt_updates_per_s": pts * steps / t_fwd,
            "max_abs_gradient": {"rho": gn[0], "mu": gn[1], "lambda": gn[2]},
        },
    }
    print(json.dumps(line), flush=True)


def main_batch(a, torch):
    """Batched shots: --shots grids of N^3 points, N drawn uniformly from
    128..288 (seeded), h = 25 m, workload material, one shallow double couple
    and 128 surface receivers each.  The same number of steps three ways:
    every shot stepped alone (fused step per shot), batched (one launch per
    sweep for all shots), batched with the three-step period replayed as a
    hipGraph; the k-block is autotuned first.  value = batched + graph
    grid-point updates per second over all shots."""
    from sw4lite_amd import EW
    from sw4lite_amd import lib as L
    from sw4lite_amd.ew import Batch
    rng = np.random.default_rng(8)
    h, steps = 25.0, a.steps
    sizes = [int(v) for v in rng.integers(128, 289, size=a.shots)]
    shots = []
    for q, n in enumerate(sizes):
        rho, mu, la = make_material(torch, n, h, seed=100 + q)
        e = EW(n, n, n, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
        e.set_material(rho, mu, la)
        e.sync()
        del rho, mu, la
        span = (n - 1) * h
        e.add_moment_source(rng.uniform(0.3, 0.7) * span, rng.uniform(0.3, 0.7) * span, rng.uniform(200, 1000),
                            double_couple(rng, 1e15), 0.24, 5.0, 0)
        for _ in range(128):
            e.add_receiver(rng.uniform(0.15, 0.85) * span, rng.uniform(0.15, 0.85) * span, 0.0, 0)
        e.setup(nsteps=3 * steps + 10)
        shots.append(e)
    torch.cuda.empty_cache()
    pts = float(sum(n ** 3 for n in sizes))
    b = Batch(shots)
    best, tune_ms = b.tune()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def alone():
        for e in shots:
            e.timesteploop(steps, mode=0)
        for e in shots:
            e.sync()

    b.step(3, graph=True)  # warm-up and graph capture
    t_alone = timed(alone)
    t_eager = timed(lambda: b.step(steps, graph=False))
    t_graph = timed(lambda: b.step(steps, graph=True))
    b.close()
    for e in shots:
        e.close()
    line = {
        "metric": "batched shots: grid-pt updates/s over all shots (FP64 fused step)",
        "value": pts * steps / t_graph, "unit": "grid-pt updates/s", "n_gpus": 1, "steps": steps, "warmup": 3,
        "ms_per_step": 1e3 * t_graph / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "%d shots of N^3, N uniform in 128..288 (seeded), h=25 m, layered+lognormal material, free "
                        "top, supergrid 30, one double couple, 128 surface receivers each" % a.shots,
            "sizes": sizes, "total_points": pts, "autotuned_kchunk": best, "tune_ms_per_step": tune_ms,
            "alone_updates_per_s": pts * steps / t_alone, "batched_updates_per_s": pts * steps / t_eager,
            "graph_updates_per_s": pts * steps / t_graph,
            "ms_per_step": {"alone": 1e3 * t_alone / steps, "batched": 1e3 * t_eager / steps,
                            "batched_graph": 1e3 * t_graph / steps},
        },
    }
    print(json.dumps(line), flush=True)


def main_batch_adjoint(a, torch):
    """The batched-adjoint workload: --shots grids of N^3 points, N drawn
    uniformly from 128..288 (seeded), h = 25 m, workload material, one shallow
    double couple (Ricker 5 Hz) and 128 surface receivers each.  The observed
    data d_obs come from the same shots with the material raised by 3% in a
    Gaussian blob (width 15% of the grid) at the centre; the misfit is
    J = 1/2 |d - d_obs|^2 over all shots and samples.  Timed: the forward run
    of all shots (batched, graph replay) and the gradient of J with respect
    to rho, mu, lambda of every shot (sw4hip_batch_gradient: up to
    --concurrent shots at once, checkpoints every --ckpt steps, default 50).
    value = gradient seconds / forward seconds."""
    from sw4lite_amd import EW
    from sw4lite_amd import lib as L
    from sw4lite_amd.ew import Batch
    rng = np.random.default_rng(8)
    h, steps = 25.0, a.steps
    ckpt = a.ckpt if a.ckpt > 0 else 50
    sizes = [int(v) for v in rng.integers(128, 289, size=a.shots)]
    shots, mats = [], []
    for q, n in enumerate(sizes):
        rho, mu, la = make_material(torch, n, h, seed=100 + q)
        e = EW(n, n, n, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
        span = (n - 1) * h
        e.add_moment_source(rng.uniform(0.3, 0.7) * span, rng.uniform(0.3, 0.7) * span, rng.uniform(200, 1000),
                            double_couple(rng, 1e15), 0.24, 5.0, 0)
        for _ in range(128):
            e.add_receiver(rng.uniform(0.15, 0.85) * span, rng.uniform(0.15, 0.85) * span, 0.0, 0)
        # the blob-perturbed material for the observed data
        k = torch.arange(-1, n + 3, device="cuda", dtype=torch.float64)
        x = (k - 1) * h - 0.5 * span
        r2 = x.view(-1, 1, 1) ** 2 + x.view(1, -1, 1) ** 2 + x.view(1, 1, -1) ** 2
        blob = (1 + 0.03 * torch.exp(-r2 / (2 * (0.15 * span) ** 2))).view(-1)
        e.set_material(rho * blob, mu * blob, la * blob)
        e.setup(nsteps=steps)
        e.reset()
        e.timesteploop(steps, mode=0)
        dobs = e.receivers()[: steps + 1]
        e.set_material(rho, mu, la)
        e.sync()
        del rho, mu, la, blob, r2
        shots.append(e)
        mats.append(dobs)
        sys.stderr.write("batch-adjoint: shot %d/%d (%d^3) set up\n" % (q + 1, a.shots, n))
        sys.stderr.flush()
    torch.cuda.empty_cache()
    pts = float(sum(n ** 3 for n in sizes))
    b = Batch(shots)
    b.set_kchunk(32)

    def forward():
        for e in shots:
            e.reset()
        b.step(steps, graph=True)
        for e in shots:
            e.sync()

    forward()  # warm-up and graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    forward()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    gtr = [e.receivers()[: steps + 1] - d for e, d in zip(shots, mats)]
    J = 0.5 * sum(float(np.sum(g * g)) for g in gtr)
    torch.cuda.synchronize()
    # the first call allocates the batch's work memory, which later calls reuse
    print("batch-adjoint: first gradient call", flush=True)
    t0 = time.perf_counter()
    b.gradient(gtr, steps, ckpt=ckpt, max_concurrent=a.concurrent)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    print("batch-adjoint: first gradient call %.1f s; timed call" % t_first, flush=True)
    t0 = time.perf_counter()
    grads = b.gradient(gtr, steps, ckpt=ckpt, max_concurrent=a.concurrent)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    gmax = [max(float(np.max(np.abs(g[c]))) for g in grads) for c in range(3)]
    del grads
    fp32 = {}
    if not a.no_fp32:
        # the FP32 form's gradients of the same shots (forward and its
        # recomputation in FP32 from float checkpoints, reverse in FP64),
        # against the FP64 batched forward (the work memory of the FP64
        # gradients is reused)
        torch.cuda.synchronize()
        print("batch-adjoint: FP32-form gradient call", flush=True)
        t0 = time.perf_counter()
        b.gradient(gtr, steps, ckpt=ckpt, max_concurrent=a.concurrent, fp32=True)
        torch.cuda.synchronize()
        t_g32 = time.perf_counter() - t0
        fp32 = {"gradient_s": t_g32, "gradient_over_fp64_forward": t_g32 / t_fwd}
    b.close()
    for e in shots:
        e.close()
    line = {
        "metric": "batched adjoint/forward cost ratio (gradients of a receiver misfit w.r.t. rho, mu, lambda)",
        "value": t_grad / t_fwd, "unit": "ratio", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": 1e3 * t_grad / steps, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "%d shots of N^3, N uniform in 128..288 (seeded), h=25 m, layered+lognormal material, free "
                        "top, supergrid 30, one double couple, 128 surface receivers each; d_obs from the material "
                        "+3%% in a central Gaussian blob" % a.shots,
            "sizes": sizes, "total_points": pts, "checkpoint_every": ckpt, "concurrent_shots": a.concurrent,
            "forward_s": t_fwd, "gradient_s": t_grad, "gradient_first_call_s": t_first, "misfit": J,
            "forward_grid_pt_updates_per_s": pts * steps / t_fwd,
            "max_abs_gradient": {"rho": gmax[0], "mu": gmax[1], "lambda": gmax[2]},
            "fp32_form": fp32,
        },
    }
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        sys.stderr.write("bench.py: this build runs on one GPU only; --gpus %d is not supported\n" % a.gpus)
        sys.exit(2)
    import torch
    from sw4lite_amd import EW
    from sw4lite_amd import lib as L
    if not torch.cuda.is_available():
        sys.stderr.write("bench.py: no GPU visible\n")
        sys.exit(1)
    if a.workload == "curvilinear":
        return main_curvilinear(a, torch)
    if a.workload == "precision":
        return main_precision(a, torch)
    if a.workload == "adjoint":
        return main_adjoint(a, torch)
    if a.workload == "batch":
        return main_batch(a, torch)
    if a.workload == "batch-adjoint":
        return main_batch_adjoint(a, torch)
    n, h = a.n, 20.0
    srcs = sources(n, h, seed=2)
    rho, mu, la = make_material(torch, n, h, seed=2)
    e = EW(n, n, n, h, [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID], corder=1, sg_gp_thickness=30)
    e.set_kernel_variant(a.variant, a.kchunk)
    e.set_material(rho, mu, la)
    del rho, mu, la
    torch.cuda.empty_cache()
    for s in srcs:
        e.add_moment_source(s["x"], s["y"], s["z"], s["m"], s["t0"], s["freq"], s["tfkind"])
    e.setup(nsteps=500)
    if a.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    e.timesteploop(a.warmup, mode=0)
    e.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stage_ms = e.timesteploop_timed(a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    inf = e.info()
    pts = float(n) ** 3
    value = pts * a.steps / wall
    # dominant kernel: the slower of the predictor and corrector sweeps over
    # the interior planes k = 7..n.  Its roofline counts the algorithmic bytes
    # of the unit of work (a point of the sweep): the predictor reads U, Um,
    # mu, lambda, rho and writes Up (12 words); a corrector that stages a
    # stored acceleration moves 12 words (A, Up, mu, lambda, rho in; Up' out),
    # one that forms it reads Up, U, Um, mu, lambda, rho (15 words).  The
    # bytes the kernels actually move with the stored acceleration (the
    # predictor writes A: 15 words) are reported beside it.
    corr_pts = float(n) * n * (n - 6)
    corr_ms = stage_ms[3] / a.steps
    pred_ms = stage_ms[0] / a.steps
    even = (n + 4) % 2 == 0
    stored_acc = a.variant >= 8 and (even or a.variant >= 10)
    if stored_acc:
        dw = "" if even else "_dw"  # odd rows: four-byte asynchronous loads
        kern = {"pred": ("k_sweep_dma%s<0>" % dw, PRED_BYTES_PER_PT, PRED_ACC_BYTES_PER_PT),
                "corr": ("k_sweep_dma%s<2>" % dw, CORR_ACC_BYTES_PER_PT, CORR_ACC_BYTES_PER_PT)}
    elif a.variant == 7 and (n + 4) % 2 == 0:
        kern = {"pred": ("k_sweep_dma<0>", PRED_BYTES_PER_PT, PRED_BYTES_PER_PT),
                "corr": ("k_sweep_lds<true,...>", CORR_BYTES_PER_PT, CORR_BYTES_PER_PT)}
    elif a.variant >= 1:
        kern = {"pred": ("k_sweep_lds<false,...>", PRED_BYTES_PER_PT, PRED_BYTES_PER_PT),
                "corr": ("k_sweep_lds<true,...>", CORR_BYTES_PER_PT, CORR_BYTES_PER_PT)}
    else:
        kern = {"pred": ("k_pred_interior", PRED_BYTES_PER_PT, PRED_BYTES_PER_PT),
                "corr": ("k_corr_interior", CORR_BYTES_PER_PT, CORR_BYTES_PER_PT)}
    dom = "pred" if pred_ms > corr_ms else "corr"
    dom_kernel, dom_bpp, dom_moved = kern[dom]
    dom_ms = pred_ms if dom == "pred" else corr_ms
    achieved = corr_pts * dom_bpp / (dom_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_sweeps_512.json")
    if os.path.exists(tf) and n == 512:
        traffic = json.load(open(tf)).get("bytes_per_launch", {}).get(dom_kernel)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 Cartesian FP64 step, h=20 m, layered+lognormal material, free top, supergrid 30, "
                        "4 double couples (Ricker 5 Hz)" % n,
            "grid": [n, n, n], "h": h, "dt": inf.dt, "kernel_variant": a.variant, "kchunk": a.kchunk, "sources": len(srcs), "grid_point_sources": inf.nsources,
            "step_hbm_fraction": value * STEP_BYTES_PER_PT / (HBM_PEAK_GBS * 1e9),
            "stage_ms_per_step": {k: v / a.steps for k, v in zip(STAGES, stage_ms.tolist())},
            "pred_interior_GBps": corr_pts * kern["pred"][1] / (pred_ms * 1e-3) / 1e9,
            "corr_interior_GBps": corr_pts * kern["corr"][1] / (corr_ms * 1e-3) / 1e9,
            "stored_acceleration": stored_acc,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": dom_kernel, "bytes_per_launch": corr_pts * dom_bpp, "bytes_per_pt": dom_bpp,
            "avg_launch_ms": dom_ms,
            "moved_bytes_per_launch": corr_pts * dom_moved,
            "moved_GBps": corr_pts * dom_moved / (dom_ms * 1e-3) / 1e9,
        },
        "step_roofline": {
            "bound": "hbm", "bytes_per_pt": STEP_BYTES_PER_PT, "achieved": value * STEP_BYTES_PER_PT / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": value * STEP_BYTES_PER_PT / (HBM_PEAK_GBS * 1e9),
        },
        "cpu_baseline": None,
    }
    if not a.no_fp32:
        line["fp32_comp"] = bench_fp32(a, e, torch, n, value)
    e.close()
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, h, a.cpu_nz, a.cpu_steps, srcs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
