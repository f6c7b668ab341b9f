"""GPU parity of the subcell limiter through the C ABI: the reference's golden
vectors (order 3) and the oracle (orders 1..9) for both projections, the cell
min/max and the DMP decision.  Bar: <= 1e-12 relative (sum-factorised
contractions on the GPU vs the oracle's direct triple sums)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def test_golden_projections_gpu():
    import exahype_amd as E
    with open(os.path.join(HERE, "golden", "limiter_3d.json")) as f:
        g = {k: np.array(v) for k, v in json.load(f).items()}
    lim = E.projectOnFVLimiterSpace(g["testFromLuhConversion/luh_in"], 3)
    assert rel(lim, g["testFromLuhConversion/lim_out"]) <= 1e-12
    luh = E.projectOnDGSpace(g["testToLuhConversion/lim_in"], 3)
    assert rel(luh, g["testToLuhConversion/luh_out"]) <= 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("ghost", [0, 2])
def test_projections_vs_oracle(N, ghost):
    import exahype_amd as E
    rng = np.random.default_rng(N + 10 * ghost)
    luh = 1.0 + rng.random((N + 1) ** 3 * 5)
    a = E.projectOnFVLimiterSpace(luh, N, ghost)
    b = O.project_on_fv(N, luh, ghost)
    assert rel(a, b) <= 1e-12
    S = 2 * N + 1 + 2 * ghost
    lim = 1.0 + rng.random(S ** 3 * 5)
    assert rel(E.projectOnDGSpace(lim, N, ghost), O.project_on_dg(N, lim, ghost)) <= 1e-12


@pytest.mark.parametrize("N", [2, 3, 5, 7, 9])
def test_dmp_vs_oracle(N):
    import exahype_amd as E
    rng = np.random.default_rng(N)
    for trial in range(4):
        luh = 1.0 + 0.1 * rng.random((N + 1) ** 3 * 5)
        mn, mx = O.cell_min_max(N, luh)
        widen = 0.02 * (trial - 1.5)  # some admissible, some not
        bmin = np.tile(mn + widen, 6) + 0.001 * rng.random(30)
        bmax = np.tile(mx - widen, 6) - 0.001 * rng.random(30)
        ok_g, bmin_g, bmax_g = E.discreteMaximumPrincipleAndMinAndMaxSearch(luh, N, bmin, bmax, 1e-4, 1e-3)
        ok_o, bmin_o, bmax_o = O.dmp(N, luh, bmin, bmax, 1e-4, 1e-3)
        assert ok_g == ok_o
        assert rel(bmin_g, bmin_o) <= 1e-12 and rel(bmax_g, bmax_o) <= 1e-12
