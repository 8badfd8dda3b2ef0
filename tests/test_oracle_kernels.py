"""Pins the CPU oracle's stencil kernels:
  * against the reference kernels compiled from their own sources
    (oracle/_ref, when it is built in this container);
  * against the committed golden vectors made from them
    (tests/golden/kernels.npz, make_golden.py);
  * against properties the mathematics fixes: L(constant) = 0, linearity,
    exactness of the 4th-order operators on low-degree polynomials in the
    interior and on the SBP boundary rows, and the vanishing of supergrid
    damping when u = um.
"""
import json
import os

import numpy as np
import pytest

from common import ROOT, grid_box, npts, random_inputs, relerr, to_layout
from oracle import oracle as O

GOLD = os.path.join(ROOT, "tests", "golden")


def run_rhs(corder, box, nz, os_, coefs, u, f, h=0.1, lib=None):
    acof, ghcof, bope, _ = coefs
    lu = np.zeros_like(u)
    O.rhs4sg(corder, box, nz, os_, acof, bope, ghcof, lu, u, f["mu"], f["la"], h, f["strx"], f["stry"],
             f["strz"], lib=lib)
    return lu


def test_golden_inputs_reproduce():
    meta = json.load(open(os.path.join(GOLD, "kernels.json")))
    nx, ny, nz = meta["grid"]
    for name, case in meta["cases"].items():
        _, f = random_inputs(case["seed"], nx, ny, nz)
        for k, v in case["checksums"].items():
            assert abs(float(np.sum(f[k])) - v) <= 1e-12 * max(1.0, abs(v)), (name, k)


@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("os45", [(0, 0), (1, 0), (1, 1)])
def test_rhs4sg_matches_golden(coefs, corder, os45):
    meta = json.load(open(os.path.join(GOLD, "kernels.json")))
    gold = np.load(os.path.join(GOLD, "kernels.npz"))
    nx, ny, nz = meta["grid"]
    name = "rhs4sg_rev_os%d%d" % os45
    box, f = random_inputs(meta["cases"][name]["seed"], nx, ny, nz)
    lu = run_rhs(corder, box, nz, [0, 0, 0, 0, *os45], coefs, to_layout(f["u"], box, corder), f, h=meta["h"])
    assert relerr(lu, to_layout(gold[name], box, corder)) < 1e-13


@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4_matches_golden(corder):
    meta = json.load(open(os.path.join(GOLD, "kernels.json")))
    gold = np.load(os.path.join(GOLD, "kernels.npz"))
    nx, ny, nz = meta["grid"]
    box, f = random_inputs(meta["cases"]["addsgd4fort_rev"]["seed"], nx, ny, nz)
    up0 = to_layout(f["up"], box, corder)
    up = up0.copy()
    O.addsgd4(corder, box, up, to_layout(f["u"], box, corder), to_layout(f["um"], box, corder), f["rho"],
              f["dcx"], f["dcy"], f["dcz"], f["strx"], f["stry"], f["strz"], f["cox"], f["coy"], f["coz"],
              meta["beta"])
    assert relerr(up - up0, to_layout(gold["addsgd4fort_rev"], box, corder) - up0) < 1e-13


@pytest.mark.skipif(O.reference() is None, reason="reference kernels not built (no reference tree)")
@pytest.mark.parametrize("corder", [0, 1])
@pytest.mark.parametrize("os45", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_rhs4sg_matches_reference_build(coefs, corder, os45, seed):
    nx, ny, nz = 11, 10, 15
    box, f = random_inputs(seed, nx, ny, nz)
    u = to_layout(f["u"], box, corder)
    a = run_rhs(corder, box, nz, [0, 0, 0, 0, *os45], coefs, u, f)
    b = run_rhs(corder, box, nz, [0, 0, 0, 0, *os45], coefs, u, f, lib=O.reference())
    assert relerr(a, b) < 1e-14


@pytest.mark.skipif(O.reference() is None, reason="reference kernels not built (no reference tree)")
@pytest.mark.parametrize("corder", [0, 1])
def test_addsgd4_matches_reference_build(corder):
    box, f = random_inputs(5, 12, 11, 10)
    u, um = to_layout(f["u"], box, corder), to_layout(f["um"], box, corder)
    up0 = to_layout(f["up"], box, corder)
    res = []
    for lib in (None, O.reference()):
        up = up0.copy()
        O.addsgd4(corder, box, up, u, um, f["rho"], f["dcx"], f["dcy"], f["dcz"], f["strx"], f["stry"], f["strz"],
                  f["cox"], f["coy"], f["coz"], 0.02, lib=lib)
        res.append(up - up0)
    assert relerr(res[0], res[1]) < 1e-14


# ---------------------------------------------------------------------------
# mathematics
# ---------------------------------------------------------------------------
def coords(box, h):
    ib, ie, jb, je, kb, ke = box
    k, j, i = np.meshgrid(np.arange(kb, ke + 1), np.arange(jb, je + 1), np.arange(ib, ie + 1), indexing="ij")
    return (i - 1) * h, (j - 1) * h, (k - 1) * h


def interior_mask(box, nz, os_):
    ib, ie, jb, je, kb, ke = box
    m = np.zeros((ke - kb + 1, je - jb + 1, ie - ib + 1), bool)
    m[2:-2, 2:-2, 2:-2] = True
    return m


@pytest.mark.parametrize("os45", [(0, 0), (1, 1)])
def test_rhs_annihilates_constants(coefs, os45):
    nx, ny, nz = 12, 12, 14
    box, f = random_inputs(3, nx, ny, nz)
    n = npts(box)
    u = np.concatenate([np.full(n, 0.3), np.full(n, -1.2), np.full(n, 2.5)])
    lu = run_rhs(1, box, nz, [0, 0, 0, 0, *os45], coefs, u, f)
    assert np.max(np.abs(lu)) < 1e-10


def test_rhs_is_linear(coefs):
    nx, ny, nz = 12, 12, 14
    box, f = random_inputs(4, nx, ny, nz)
    _, g = random_inputs(40, nx, ny, nz)
    os_ = [0, 0, 0, 0, 1, 1]
    a = run_rhs(1, box, nz, os_, coefs, f["u"], f)
    b = run_rhs(1, box, nz, os_, coefs, g["u"], f)
    c = run_rhs(1, box, nz, os_, coefs, 2.0 * f["u"] - 3.0 * g["u"], f)
    assert relerr(c, 2.0 * a - 3.0 * b) < 1e-13


def elastic_operator_poly(X, Y, Z, mu, la):
    """L(u) for u = (y z + x^2, x^2 - z^2 + x y, x z + y^2) with constant mu,
    lambda:  L(u)_i = mu lap u_i + (mu + la) d_i div u (hand-derived)."""
    # u1 = y z + x^2     lap = 2,  u2 = x^2 - z^2 + x y  lap = 0, u3 = x z + y^2  lap = 2
    # div u = 2x + x + x = 4x  -> grad div = (4, 0, 0)
    l1 = mu * 2 + (mu + la) * 4
    l2 = mu * 0 + 0.0
    l3 = mu * 2 + 0.0
    return l1, l2, l3


@pytest.mark.parametrize("os45", [(0, 0), (1, 0), (1, 1)])
def test_rhs_exact_on_quadratics(coefs, os45):
    """4th-order interior stencil and 2nd-order-accurate-or-better SBP rows
    are exact on quadratic displacements with constant coefficients."""
    nx, ny, nz = 14, 13, 16
    h = 0.25
    box = grid_box(nx, ny, nz)
    n = npts(box)
    X, Y, Z = (c.ravel() for c in coords(box, h))
    mu, la = 1.3, 0.7
    u = np.concatenate([Y * Z + X ** 2, X ** 2 - Z ** 2 + X * Y, X * Z + Y ** 2])
    f = {"mu": np.full(n, mu), "la": np.full(n, la), "strx": np.ones(nx + 4), "stry": np.ones(ny + 4),
         "strz": np.ones(nz + 4)}
    lu = run_rhs(1, box, nz, [0, 0, 0, 0, *os45], coefs, u, f, h=h).reshape(3, nz + 4, ny + 4, nx + 4)
    want = elastic_operator_poly(X, Y, Z, mu, la)
    for c in range(3):
        assert np.max(np.abs(lu[c, 2:-2, 2:-2, 2:-2] - want[c])) < 1e-9 * (1 + abs(want[c]))


def test_addsgd4_vanishes_for_u_equal_um():
    box, f = random_inputs(6, 12, 12, 12)
    up = f["up"].copy()
    O.addsgd4(1, box, up, f["u"], f["u"].copy(), f["rho"], f["dcx"], f["dcy"], f["dcz"], f["strx"], f["stry"],
              f["strz"], f["cox"], f["coy"], f["coz"], 0.02)
    assert np.array_equal(up, f["up"])


def test_addsgd4_is_fourth_difference():
    """With unit coefficients the damping of a 1-D field u(x) = x^p - um is
    beta * D+^2 D-^2 u, zero for p <= 3 and 24 h^0 for p = 4."""
    nx, ny, nz = 12, 12, 12
    box = grid_box(nx, ny, nz)
    n = npts(box)
    i = np.tile(np.arange(-1, nx + 3), (nz + 4) * (ny + 4)).astype(float)
    ones = {a: np.ones(m) for a, m in (("x", nx + 4), ("y", ny + 4), ("z", nz + 4))}
    for p, want in ((3, 0.0), (4, 24.0)):
        u = np.concatenate([i ** p, np.zeros(n), np.zeros(n)])
        up = np.zeros(3 * n)
        O.addsgd4(1, box, up, u, np.zeros(3 * n), np.ones(n), ones["x"], ones["y"], ones["z"], ones["x"],
                  ones["y"], ones["z"], ones["x"], ones["y"], ones["z"], 1.0)
        got = -up.reshape(3, nz + 4, ny + 4, nx + 4)[0, 2:-2, 2:-2, 2:-2]
        assert np.allclose(got, want, atol=1e-9)
