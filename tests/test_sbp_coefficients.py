"""Pins the SBP coefficient table of the library (sbp_coefficients.cpp):
  * bit-identical to the reference's GetStencilCoefficients values
    (ew-cfromfort.C:1487-1944), through the rational table extracted by
    tests/golden/make_golden.py and, where present, the reference source;
  * the accuracy conditions of the operators, checked in exact rational
    arithmetic: the one-sided first derivative bope is exact on polynomials
    of degree <= 2 (rows 1-4) and <= 4 (rows 5-6, the centred stencil), the
    variable-coefficient second derivative acof/ghcof is exact on
    (mu u')' for mu of degree <= 1 and u of degree <= 2, and the
    free-surface derivative sbop is exact on cubics at the boundary.
A wrong digit, a transposed index or a dropped entry fails at least one."""
import json
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

from common import ROOT

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "sbp_coefficients.json")))


def tables():
    acof = {}
    for k, q, m, v in GOLD["acof"]:
        acof[(k, q, m)] = Fr(v)
    bope = {(k, q): Fr(v) for k, q, v in GOLD["bope"]}
    return acof, bope, Fr(GOLD["ghcof1"]), [Fr(v) for v in GOLD["sbop"]]


def test_library_table_is_reference_table(coefs):
    acof_l, ghcof_l, bope_l, sbop_l = coefs
    acof, bope, gh, sbop = tables()
    want = np.zeros(384)
    for (k, q, m), v in acof.items():
        want[(k - 1) + 6 * (q - 1) + 48 * (m - 1)] = float(v)
    assert np.array_equal(acof_l, want)
    wb = np.zeros(48)
    for (k, q), v in bope.items():
        wb[(k - 1) + 6 * (q - 1)] = float(v)
    assert np.array_equal(bope_l, wb)
    assert ghcof_l[0] == float(gh) and not np.any(ghcof_l[1:])
    assert np.array_equal(sbop_l, np.array([float(v) for v in sbop]))


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference source tree not present")
def test_golden_table_matches_reference_source():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import sbp_from_reference
    assert sbp_from_reference("/root/reference") == GOLD


def poly(c, x):
    return sum(ci * x ** i for i, ci in enumerate(c))


def dpoly(c, x):
    return sum(i * ci * x ** (i - 1) for i, ci in enumerate(c) if i > 0)


def test_bope_accuracy():
    _, bope, _, _ = tables()
    # grid x_q = q - 1 (h = 1), q = 1..8
    for k in range(1, 7):
        deg = 2 if k <= 4 else 4
        for p in range(deg + 1):
            c = [0] * p + [1]
            got = sum(bope.get((k, q), 0) * poly(c, Fr(q - 1)) for q in range(1, 9))
            assert got == dpoly(c, Fr(k - 1)), (k, p)


def test_acof_accuracy():
    acof, _, gh, _ = tables()
    ghcof = [gh, 0, 0, 0, 0, 0]
    # h^2 (mu u')' at x_k with x_q = q-1, ghost x_0 = -1
    for k in range(1, 7):
        for mc in ([1], [2, 1], [Fr(1, 3), Fr(-1, 5)]):
            for uc in ([0, 1], [0, 0, 1], [1, 2, 3]):
                mu = [poly(mc, Fr(m - 1)) for m in range(1, 9)]
                got = Fr(0)
                for q in range(1, 9):
                    coef = sum(acof.get((k, q, m), 0) * mu[m - 1] for m in range(1, 9))
                    got += coef * poly(uc, Fr(q - 1))
                got += ghcof[k - 1] * mu[0] * poly(uc, Fr(-1))
                # exact: (mu u')' = mu' u' + mu u''
                x = Fr(k - 1)
                want = dpoly(mc, x) * dpoly(uc, x) + poly(mc, x) * sum(
                    i * (i - 1) * ci * x ** (i - 2) for i, ci in enumerate(uc) if i > 1)
                assert got == want, (k, mc, uc)


def test_sbop_accuracy():
    _, _, _, sbop = tables()
    # boundary derivative at x_1 = 0 from the ghost x_0 = -1 and x_1..x_4
    for p in range(4):
        c = [0] * p + [1]
        got = sbop[0] * poly(c, Fr(-1)) + sum(sbop[m] * poly(c, Fr(m - 1)) for m in range(1, 5))
        assert got == dpoly(c, Fr(0)), p
