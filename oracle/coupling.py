"""The grid interface of a curvilinear grid stacked on the top Cartesian grid
(EW::enforceCartTopo, EW.C:3503-3531), restated with numpy.

TEST INFRASTRUCTURE ONLY (tests/).

Arrays hold 3 components over a box (ib..ie, jb..je, kb..ke): corder 1 is
component-blocked with i fastest (Sarray.C:756-776), corder 0 interleaves
the component fastest.  Both grids cover the same (i, j) box."""
import numpy as np


def _view(u, corder, ni, nj, nk):
    """u as an array indexed [c, k, j, i] (a view of the same memory)."""
    if corder == 1:
        return u.reshape(3, nk, nj, ni)
    return u.reshape(nk, nj, ni, 3).transpose(3, 0, 1, 2)


def enforce_cart_topo(corder, ij, cart_k, ucart, curv_k, ucurv, g=2):
    """In place, in the reference's order for each (i, j) (EW.C:3514-3528):
    the Cartesian planes kb .. kb+g-1 take the curvilinear planes
    ke-2g+q, then the curvilinear planes ke-q (q = 0..g) take the Cartesian
    planes kb+2g-q."""
    ib, ie, jb, je = ij
    ni, nj = ie - ib + 1, je - jb + 1
    kbc, kec_cart = cart_k
    kbg, keg = curv_k
    vc = _view(ucart, corder, ni, nj, kec_cart - kbc + 1)
    vg = _view(ucurv, corder, ni, nj, keg - kbg + 1)
    for q in range(g):
        vc[:, kbc + q - kbc] = vg[:, keg - 2 * g + q - kbg]
    for q in range(g + 1):
        vg[:, keg - q - kbg] = vc[:, kbc + 2 * g - q - kbc]
