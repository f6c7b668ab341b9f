"""Pin the oracle's subcell limiter against the reference's golden vectors
(E/tests/kernels/testdata/limiter_testdata.cpp, Dim3, order 3: luh -> subcell
averages, luh -> Gauss-Lobatto values, subcell averages -> luh) and against
what the mathematics fixes (projections preserve constants and the cell
mean; projecting to the subcells and back is the identity on DG
polynomials; the DMP with zero tolerance fails exactly when the cell leaves
its neighbours' range)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from exahype_amd.tables import dg_math

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lg():
    with open(os.path.join(HERE, "golden", "limiter_3d.json")) as f:
        return {k: np.array(v) for k, v in json.load(f).items()}


def test_golden_projections(lg):
    N = 3
    lim = O.project_on_fv(N, lg["testFromLuhConversion/luh_in"])
    exp = lg["testFromLuhConversion/lim_out"]
    assert np.abs(lim - exp).max() <= 1e-12 * np.abs(exp).max()
    luh = O.project_on_dg(N, lg["testToLuhConversion/lim_in"])
    exp = lg["testToLuhConversion/luh_out"]
    assert np.abs(luh - exp).max() <= 1e-12 * np.abs(exp).max()


def test_golden_lobatto_values(lg):
    N = 3
    _, u2lob, _ = O.limiter_matrices(N)
    luh = lg["testFromLuhConversion/luh_in"].reshape(4, 4, 4, 5)
    lob = np.einsum("abcv,az,by,cx->zyxv", luh, u2lob, u2lob, u2lob).ravel()
    exp = lg["testFromLuhConversion/lob_out"]
    assert np.abs(lob - exp).max() <= 1e-12 * np.abs(exp).max()


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 9])
def test_projection_properties(N):
    B, L = N + 1, 2 * N + 1
    u2l, u2lob, l2u = O.limiter_matrices(N)
    # subcell averages of phi_k: columns sum over k to 1 (partition of unity)
    np.testing.assert_allclose(u2l.sum(axis=0), 1.0, rtol=0, atol=1e-12)
    # lim2uh . uh2lim = identity on the DG space
    np.testing.assert_allclose(l2u.T @ u2l.T, np.eye(B), atol=1e-10)
    # cell mean is conserved by the back projection: sum_i avg_i / L == sum_k w_k u_k
    w = dg_math.gl_nodes_weights(N)[1]
    rng = np.random.default_rng(N)
    lim = rng.normal(size=L)
    np.testing.assert_allclose(w @ (l2u.T @ lim), lim.mean(), atol=1e-13)
    # Gauss-Lobatto projection: values of the nodal polynomial at the GLL nodes
    xl = np.vectorize(float.fromhex)(np.array(json.load(open(os.path.join(
        HERE, "..", "exahype_amd", "tables", "dg_tables.json")))["gll_nodes"], dtype=object)).astype(float)[N][:B]
    xi = dg_math.gl_nodes_weights(N)[0]
    ref = np.array([[dg_math.lagrange_eval(xi, k, x) for x in xl] for k in range(B)])
    np.testing.assert_allclose(u2lob, ref, atol=1e-12)


def test_dmp_decision():
    N = 3
    rng = np.random.default_rng(0)
    luh = 1.0 + 0.01 * rng.random(64 * 5)
    mn, mx = O.cell_min_max(N, luh)
    lo, hi = np.tile(mn, 6), np.tile(mx, 6)
    ok, bmin, bmax = O.dmp(N, luh, lo, hi, 0.0, 0.0)
    assert ok
    np.testing.assert_array_equal(bmin[:5], mn)
    np.testing.assert_array_equal(bmin[5:10], mn)
    hi2 = np.tile(mx - 1e-3, 6)
    ok, _, _ = O.dmp(N, luh, lo, hi2, 0.0, 0.0)
    assert not ok
    ok, _, _ = O.dmp(N, luh, lo, hi2, 2e-3, 0.0)  # relaxation admits it
    assert ok
