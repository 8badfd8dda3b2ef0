#!/usr/bin/env python3
"""Throughput of the sw4lite elastic-wave time step on one MI355X.

Workload (the reference path at scale): FP64 Cartesian 4th-order SBP solver
on a 512^3 grid, h = 20 m; layered material cs = 1000 -> 3500 m/s linear in
depth, cp = 1.8 cs, rho = 1700 (cp / 1 km/s)^0.25 kg/m^3, times a seeded
lognormal perturbation exp(0.05 g) of cs and cp, g a unit-variance field of
white noise smoothed by a Gaussian of 8 cells; free surface on top, supergrid
of 30 points on the other five faces; four seeded random double-couple
sources with 5 Hz Ricker time functions.  All inputs are synthetic and built
on the GPU before timing.

A step is one full time step (predictor sweep, sources, ghost-point BC,
corrector + supergrid sweep, sources, BC).  The metric is interior grid-point
updates per second: 512^3 * steps / seconds, inputs resident in HBM.

The JSON line adds
  roofline: the dominant kernel (corrector sweep, interior points), its
    algorithmic bytes per launch (15 FP64 words per point: read Up, U, Um (3
    components each), mu, lambda, rho, write Up') over its average launch
    time from hipEvents on the stream it runs on, against 8 TB/s;
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
CORR_BYTES_PER_PT = 15 * 8     # corrector + damping sweep, FP64
PRED_BYTES_PER_PT = 12 * 8     # predictor sweep, FP64
STEP_BYTES_PER_PT = 27 * 8     # whole step, compulsory traffic


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=512, help="grid points per axis")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-nz", type=int, default=128, help="depth of the CPU-baseline slab")
    p.add_argument("--cpu-steps", type=int, default=3)
    p.add_argument("--variant", type=int, default=1, help="interior sweep kernels: 0 thread/point, 1 LDS k-march")
    p.add_argument("--kchunk", type=int, default=32, help="planes per workgroup of the LDS sweeps")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also launch a plain device copy of 1 GiB (PMC scaling)")
    return p.parse_args()


def make_material(torch, n, h, seed):
    """rho, mu, lambda over the (n+4)^3 box (two ghost points per side)."""
    dev = torch.device("cuda")
    m = n + 4
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    g = torch.randn((m, m, m), generator=gen, device=dev, dtype=torch.float32)
    sig = 8.0
    r = int(4 * sig)
    x = torch.arange(-r, r + 1, device=dev, dtype=torch.float32)
    w = torch.exp(-0.5 * (x / sig) ** 2)
    w = (w / w.sum()).view(1, 1, -1)
    F = torch.nn.functional
    for axis in range(3):  # separable Gaussian smoothing along k, j, i
        g = g.movedim(axis, 2).contiguous()
        sh = g.shape
        g = F.conv1d(g.reshape(-1, 1, sh[2]), w, padding=r).reshape(sh)
        g = g.movedim(2, axis).contiguous()
    g = (g - g.mean()) / g.std()
    p = torch.exp(0.05 * g.double())
    del g
    k = torch.arange(-1, n + 3, device=dev, dtype=torch.float64)
    z = (k - 1) * h
    cs = (1000.0 + 2500.0 * z / ((n - 1) * h)).view(m, 1, 1) * p
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
    fall inside), one warm-up step and `steps` timed steps."""
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
    inside = [s for s in srcs if s["z"] < (nz - 6) * h] or [dict(srcs[0], z=0.5 * (nz - 1) * h)]
    bc = [L.BC_SUPERGRID] * 4 + [L.BC_FREE, L.BC_SUPERGRID]
    run = OracleRun(n, n, nz, h, bc, rho.ravel(), mu.ravel(), la.ravel(), inside, sg_gp=30, nsteps=steps + 1,
                    kernels=kernels)
    run.step(1)
    t0 = time.perf_counter()
    run.step(steps)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": n * n * nz * steps / dt,
        "unit": "grid-pt updates/s",
        "cores": threads,
        "kind": "reference" if kernels == "reference" else "port",
        "sample": "%dx%dx%d slab of the workload (layered material, no perturbation), %d timed steps after 1 "
                  "warm-up, %.1f s; stencil stages %s, pointwise stages from the oracle restatement; "
                  "OMP threads %d" % (n, n, nz, steps, dt,
                                      "from the reference sources" if kernels == "reference" else "oracle",
                                      threads),
    }


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
    # dominant kernel: corrector sweep over the interior planes k = 7..n
    corr_pts = float(n) * n * (n - 6)
    corr_ms = stage_ms[3] / a.steps
    achieved = corr_pts * CORR_BYTES_PER_PT / (corr_ms * 1e-3) / 1e9
    pred_ms = stage_ms[0] / a.steps
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_corrector_512.json")
    if os.path.exists(tf) and n == 512:
        traffic = json.load(open(tf)).get("bytes_per_launch")
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
            "stage_ms_per_step": {k: v / a.steps for k, v in zip(
                ["pred_interior", "pred_boundary_rows_shell", "pred_sources_bc", "corr_interior",
                 "corr_boundary_rows_shell", "corr_sources_bc"], stage_ms.tolist())},
            "pred_interior_GBps": corr_pts * PRED_BYTES_PER_PT / (pred_ms * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_sweep_lds<true>" if a.variant == 1 else "k_corr_interior", "bytes_per_launch": corr_pts * CORR_BYTES_PER_PT,
            "avg_launch_ms": corr_ms,
        },
        "cpu_baseline": None,
    }
    e.close()
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, h, a.cpu_nz, a.cpu_steps, srcs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
