"""GPU: the reference's topography deck (pytest/reference/topo/curvilinear.in,
tests/golden/decks/curvilinear.in) run unmodified through the deck front
end: a curvilinear grid under a Gaussian hill stacked on the Cartesian grid
below z = 2 (65 x 65 x 103 and 65 x 65 x 151 points), coupled each step by
EW::enforceCartTopo, a moment source 0.2 below the surface and three
stations, two on the surface and one 0.3 below it.  The three receiver
files must hold the samples of the reference's own fixture
(pytest/reference/topo/curvilinear-output/sta0{1,2,3}.txt,
tests/golden/topo_curvilinear/) to 1e-12 of each station's peak; the
reference program compiled here reproduces the fixture to 4e-15 of the
peak."""
import os

import numpy as np
import pytest

from common import ROOT
from sw4lite_amd import ew
from sw4lite_amd import lib as L

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    L.load()


@pytest.mark.parametrize("mode", [1, 0])
def test_topography_deck_matches_reference(tmp_path, mode):
    print("topography deck, mode", mode, flush=True)
    rep = ew.run_deck(os.path.join(GOLDEN, "decks", "curvilinear.in"), path=str(tmp_path), mode=mode)
    assert rep["ngrids"] == 2 and (rep["nz"], rep["nz_curv"]) == (151, 103)
    assert rep["nsteps"] == 63 and rep["cycle"] == 63
    assert rep["dt"] == pytest.approx(0.5 / 63, rel=1e-15)
    for name in ("sta01", "sta02", "sta03"):
        ours = np.loadtxt(str(tmp_path / (name + ".txt")), comments="#")
        ref = np.loadtxt(os.path.join(GOLDEN, "topo_curvilinear", name + ".txt"), comments="#")
        assert ours.shape == ref.shape == (64, 4)
        np.testing.assert_allclose(ours[:, 0], ref[:, 0], rtol=1e-6, atol=0)  # times are printed to 7 digits
        peak = np.abs(ref[:, 1:]).max()
        err = np.abs(ours[:, 1:] - ref[:, 1:]).max()
        print(name, "peak", peak, "max difference", err, flush=True)
        assert peak > 1
        assert err <= TOL * peak
