"""Shared helpers of the test-suite: seeded inputs on a small grid, layout
conversion and the parity metric."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def grid_box(nx, ny, nz):
    """Index box incl. two ghost points: (ib, ie, jb, je, kb, ke)."""
    return (-1, nx + 2, -1, ny + 2, -1, nz + 2)


def npts(box):
    ib, ie, jb, je, kb, ke = box
    return (ie - ib + 1) * (je - jb + 1) * (ke - kb + 1)


def random_inputs(seed, nx, ny, nz):
    """Seeded random fields: displacements in [-1,1], positive mu, lambda,
    rho, 1-D stretching/damping/corner profiles in their natural ranges."""
    rng = np.random.default_rng(seed)
    box = grid_box(nx, ny, nz)
    n = npts(box)
    f = {}
    f["u"] = rng.uniform(-1, 1, 3 * n)
    f["um"] = rng.uniform(-1, 1, 3 * n)
    f["up"] = rng.uniform(-1, 1, 3 * n)
    f["mu"] = rng.uniform(0.5, 2.0, n)
    f["la"] = rng.uniform(0.5, 2.0, n)
    f["rho"] = rng.uniform(0.8, 1.5, n)
    for ax, m in (("x", nx + 4), ("y", ny + 4), ("z", nz + 4)):
        f["str" + ax] = rng.uniform(0.2, 1.0, m)
        f["dc" + ax] = rng.uniform(0.0, 1.0, m)
        f["co" + ax] = rng.uniform(0.33, 1.0, m)
    return box, f


def to_layout(v, box, corder):
    """Convert a component-blocked vector field (3*npts) to layout corder."""
    if corder == 1:
        return v.copy()
    n = npts(box)
    return np.ascontiguousarray(v.reshape(3, n).T.reshape(-1))


def relerr(a, b):
    """max |a - b| / max |b| (the parity metric of every test)."""
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return d / s if s > 0 else d


def hills_depth(nx, ny, nz, h, seed, nhill=3, amp=0.15):
    """Grid depth z(i,j,k) over the box: the top is a sum of Gaussian hills
    (height <= amp * depth), mapped linearly in k to a flat bottom."""
    rng = np.random.default_rng(seed)
    box = grid_box(nx, ny, nz)
    ib, ie, jb, je, kb, ke = box
    x = (np.arange(ib, ie + 1) - 1) * h
    y = (np.arange(jb, je + 1) - 1) * h
    depth = (nz - 1) * h
    X, Y = np.meshgrid(x, y)  # (nj, ni)
    top = np.zeros_like(X)
    for _ in range(nhill):
        xc, yc = rng.uniform(0.2, 0.8) * x[-1], rng.uniform(0.2, 0.8) * y[-1]
        L = rng.uniform(0.15, 0.35) * min(x[-1], y[-1])
        top -= rng.uniform(0.3, 1.0) * amp * depth / nhill * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / L ** 2)
    zeta = (np.arange(kb, ke + 1) - 1) / (nz - 1)
    z = top[None, :, :] * (1 - zeta[:, None, None]) + depth * zeta[:, None, None]
    return box, np.ascontiguousarray(z.reshape(-1))
