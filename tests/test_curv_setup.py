"""Pins the curvilinear-grid set-up and time loop on the CPU:
  * the grid-point forces of a moment source on a curvilinear grid (host
    set-up, Source.C:1605-1800) satisfy the discrete moment conditions of a
    point moment tensor, sum_p F J = 0 and sum_p F_a x_b J = M_ab, exactly
    when the grid depth is a low-degree polynomial and to the discretisation
    error on a grid under hills;
  * on a flat grid the curvilinear set-up (dt from the eigenvalue bound,
    sources, operator, damping, free-surface closure) reproduces the
    Cartesian time loop step for step.
"""
import numpy as np
import pytest

from common import grid_box, hills_depth, npts
from oracle import oracle as O
from oracle.loop import OracleRun
from sw4lite_amd import ew
from sw4lite_amd import lib as L

M6 = [0.3, 0.8, -0.45, -0.6, 0.25, 0.9]  # xx xy xz yy yz zz


def moments(desc, z, x0, y0, z0, h):
    box = grid_box(desc.nx, desc.ny, desc.nz)
    ib, ie, jb, je, kb, ke = box
    ijk, frc = ew.host_moment_source_curv(desc, z, x0, y0, z0, M6)
    n = npts(box)
    met, jac = np.zeros(4 * n), np.zeros(n)
    assert O.metric(1, box, h, z, met, jac) == 0
    ni, nij = ie - ib + 1, (ie - ib + 1) * (je - jb + 1)
    p = (ijk[:, 0] - ib) + ni * (ijk[:, 1] - jb) + nij * (ijk[:, 2] - kb)
    J = jac[p]
    X = np.stack([(ijk[:, 0] - 1) * h, (ijk[:, 1] - 1) * h, z[p]], axis=1)
    zeroth = (frc * J[:, None]).sum(axis=0)
    first = np.einsum("pa,pb,p->ab", frc, X, J)
    M = np.array([[M6[0], M6[1], M6[2]], [M6[1], M6[3], M6[4]], [M6[2], M6[4], M6[5]]])
    return zeroth, first, M, frc


def test_curvilinear_source_moments_exact_on_polynomial_grid():
    nx, ny, nz, h = 24, 22, 20, 0.1
    desc = L.grid_desc(nx, ny, nz, h, [2, 2, 2, 2, 1, 2])
    box = grid_box(nx, ny, nz)
    ib, ie, jb, je, kb, ke = box
    i = np.arange(ib, ie + 1)[None, None, :]
    j = np.arange(jb, je + 1)[None, :, None]
    k = np.arange(kb, ke + 1)[:, None, None]
    # depth quadratic in i, j and linear in k: differenced and interpolated exactly
    z = (k - 1) * h * (1 + 0.002 * (i - 10.0) + 0.001 * (j - 9.0)) + 0.003 * (i - 11.0) * (j - 8.0) * h
    z = np.ascontiguousarray(np.broadcast_to(z, (ke - kb + 1, je - jb + 1, ie - ib + 1)).reshape(-1))
    zeroth, first, M, frc = moments(desc, z, 1.13, 0.97, 1.04, h)
    scale = np.max(np.abs(M))
    assert np.max(np.abs(zeroth)) < 1e-10 * scale / h
    assert np.max(np.abs(first - M)) < 1e-10 * scale


def test_curvilinear_source_moments_under_hills():
    nx, ny, nz, h = 40, 36, 30, 0.05
    desc = L.grid_desc(nx, ny, nz, h, [2, 2, 2, 2, 0, 2])
    _, z = hills_depth(nx, ny, nz, h, 9)
    zeroth, first, M, frc = moments(desc, z, 0.93, 0.88, 0.6, h)
    scale = np.max(np.abs(M))
    assert np.max(np.abs(zeroth)) < 1e-10 * scale / h
    assert np.max(np.abs(first - M)) < 2e-3 * scale


def test_curvilinear_source_rejects_point_above_surface():
    nx, ny, nz, h = 20, 20, 16, 0.1
    desc = L.grid_desc(nx, ny, nz, h, [2, 2, 2, 2, 0, 2])
    _, z = hills_depth(nx, ny, nz, h, 9)
    with pytest.raises(L.Sw4HipError):
        ew.host_moment_source_curv(desc, z, 1.0, 1.0, -5.0, M6)


def test_flat_curvilinear_loop_equals_cartesian():
    nx, ny, nz, h = 26, 24, 22, 0.1
    bc = [2, 2, 2, 2, 0, 1]  # free top, Dirichlet bottom: no stretching along k
    box = grid_box(nx, ny, nz)
    ib, ie, jb, je, kb, ke = box
    k = np.arange(kb, ke + 1)
    z = np.ascontiguousarray(np.broadcast_to(((k - 1) * h)[:, None, None],
                                             (ke - kb + 1, je - jb + 1, ie - ib + 1)).reshape(-1))
    rng = np.random.default_rng(3)
    n = npts(box)
    rho = rng.uniform(1.0, 1.5, n)
    mu = rng.uniform(0.8, 1.2, n)
    la = rng.uniform(0.8, 1.2, n)
    src = [dict(x=1.23, y=1.11, z=0.94, m=M6, t0=0.2, freq=10.0, tfkind=L.TF_GAUSSIAN)]
    cart = OracleRun(nx, ny, nz, h, bc, rho, mu, la, src, sg_gp=6, nsteps=6)
    curv = OracleRun(nx, ny, nz, h, bc, rho, mu, la, src, sg_gp=6, nsteps=6, zgrid=z)
    assert abs(curv.dt - cart.dt) < 1e-14 * cart.dt
    np.testing.assert_allclose(curv.src[1], cart.src[1], rtol=1e-12, atol=1e-12 * np.max(np.abs(cart.src[1])))
    cart.step(6)
    curv.step(6)
    d = np.max(np.abs(curv.U - cart.U)) / np.max(np.abs(cart.U))
    assert d < 1e-12
