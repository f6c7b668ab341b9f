"""Pins of the oracle's point sources (oracle/linear.c): kernels::aderdg::
generic::c::pointSource (kernels/aderdg/generic/c/3d/pointSource.cpph:28-145)
and the usePointSource arm of spaceTimePredictorLinear
(kernels/aderdg/generic/c/3d/spaceTimePredictorLinear.cpph:78-102, 287-302).
The reference holds no test data for them (no application of the reference
uses a point source, and its linear-predictor test is disabled), so:

* the Lagrange basis: the unique degree-N polynomials with l_i(xi_j) =
  delta_ij; checked at the nodes and by reproducing x^k, k <= N, at off-node
  points (the reference's generated literals, kernels/DGBasisFunctions.cpp,
  need Peano's peano/utils/Loop.h, which the reference does not vendor);
* pointSource: a source on a node projects onto that node alone with
  force / (w^3 dx^3); any source integrates back to its force under the
  cell's quadrature (partition of unity); level n carries the force at
  t + dt xi_{n-1} (n >= 1) or t (n = 0); a source outside the cell gives zero;
* the predictor arm: at zero velocity the Cauchy-Kowalevski levels are
  closed forms of luh and the forces (lQi[0] = luh + dt sum_n w_n PSi[n+1],
  lQi[1] = PSi[0], higher levels zero, lFh[d] = sum_n w_n PSi[n+1] / 3), held
  with random forces so that a wrong weight, level or factor shows.
"""
import numpy as np
import pytest

from exahype_amd.tables import dg_math
from oracle import oracle as O

VEL = (0.3, -0.7, 0.5)
SRC = dict(x0=(0.52, 0.47, 0.53), amp=(1.0, -2.0, 3.0, 0.5, 4.0), f0=5.0, t0=0.1)


def nodes(N):
    xi, w = dg_math.gl_nodes_weights(N)
    return np.asarray(xi), np.asarray(w)


def ricker(t, f0, t0):
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


@pytest.mark.parametrize("N", range(1, 10))
def test_lagrange_basis_is_the_interpolation_basis(N):
    xi, _ = nodes(N)
    L = np.array([[O.lagrange(N, i, x) for x in xi] for i in range(N + 1)])
    assert np.abs(L - np.eye(N + 1)).max() <= 1e-12
    r = np.random.default_rng(N).uniform(0, 1, 7)
    Lr = np.array([[O.lagrange(N, i, x) for x in r] for i in range(N + 1)])
    for k in range(N + 1):
        assert np.abs((xi[:, None] ** k * Lr).sum(0) - r ** k).max() <= 1e-11, k


@pytest.mark.parametrize("N", [1, 3, 5, 9])
def test_point_source_projection(N):
    B = N + 1
    xi, w = nodes(N)
    centre, dx = np.array([0.5, 0.5, 0.5]), np.array([0.1, 0.08, 0.12])
    t, dt = 0.05, 0.01
    lo = centre - dx / 2
    W3 = np.einsum("i,j,k->ijk", w, w, w) * dx.prod()
    # a source on node (z, y, x) = (1, 0, N): that node alone, force / (w^3 dx^3)
    src = dict(SRC, x0=tuple(lo + dx * np.array([xi[N], xi[0], xi[1]])))
    c = O.cfg(N, pde=4, adv=VEL, linear=1, point_source=src)
    psi = O.point_source(c, t, dt, centre, dx).reshape(B + 1, B, B, B, 5)
    amp = np.array(SRC["amp"])
    for n in range(B + 1):
        tn = t if n == 0 else t + dt * xi[n - 1]
        f = amp * ricker(tn, SRC["f0"], SRC["t0"])
        exp = np.zeros((B, B, B, 5))
        exp[1, 0, N] = f / W3[1, 0, N]
        assert np.abs(psi[n] - exp).max() <= 1e-12 * np.abs(exp).max(), n
    # a generic source: the quadrature of every level returns the force
    c = O.cfg(N, pde=4, adv=VEL, linear=1, point_source=SRC)
    psi = O.point_source(c, t, dt, centre, dx).reshape(B + 1, B, B, B, 5)
    for n in range(B + 1):
        tn = t if n == 0 else t + dt * xi[n - 1]
        f = amp * ricker(tn, SRC["f0"], SRC["t0"])
        assert np.abs(np.einsum("ijk,ijkv->v", W3, psi[n]) - f).max() <= 1e-11 * np.abs(f).max(), n
    # outside the cell: nothing
    c = O.cfg(N, pde=4, adv=VEL, linear=1, point_source=dict(SRC, x0=(0.52, 0.47, 0.7)))
    assert not np.any(O.point_source(c, t, dt, centre, dx))


@pytest.mark.parametrize("N", [1, 2, 4, 7])
def test_predictor_point_source_arm_closed_form(N):
    B = N + 1
    _, w = nodes(N)
    c = O.cfg(N, pde=4, adv=(0.0, 0.0, 0.0), linear=1)
    rng = np.random.default_rng(10 + N)
    luh = rng.standard_normal(B ** 3 * 5)
    psi = rng.standard_normal((B + 1) * B ** 3 * 5)
    dx, dt = np.array([0.05, 0.04, 0.03]), 0.003
    o = O.space_time_predictor_linear(c, luh, dx, dt, psi=psi)
    P = psi.reshape(B + 1, -1)
    integ = (w[:, None] * P[1:]).sum(0)
    lQi = o["lQi"].reshape(B + 1, -1)
    assert np.abs(lQi[0] - (luh + dt * integ)).max() <= 1e-14 * 10
    assert np.abs(lQi[1] - P[0]).max() <= 1e-14 * 10
    assert not np.any(lQi[2:])
    assert np.abs(o["lQhi"] - (luh + dt * integ + 0.5 * dt * P[0])).max() <= 1e-13
    lFhi = o["lFhi"].reshape(4, -1)
    for d in range(3):
        assert np.abs(lFhi[d] - integ / 3).max() <= 1e-14 * 10, d
    assert not np.any(lFhi[3])
    # and without the arm nothing of it remains
    o0 = O.space_time_predictor_linear(c, luh, dx, dt)
    assert np.array_equal(o0["lQhi"], luh)


def test_point_source_is_linear_in_the_forces():
    """P(luh, psi) = P(luh, none) + P(0, psi) for the linear advection PDE."""
    N, B = 3, 4
    c = O.cfg(N, pde=4, adv=VEL, linear=1)
    rng = np.random.default_rng(3)
    luh, psi = rng.standard_normal(B ** 3 * 5), rng.standard_normal((B + 1) * B ** 3 * 5)
    dx, dt = np.array([0.05, 0.04, 0.03]), 1e-3
    a = O.space_time_predictor_linear(c, luh, dx, dt, psi=psi)
    b = O.space_time_predictor_linear(c, luh, dx, dt)
    z = O.space_time_predictor_linear(c, np.zeros_like(luh), dx, dt, psi=psi)
    for k in ("lQhi", "lFhi", "lQhbnd", "lFhbnd"):
        assert np.abs(a[k] - b[k] - z[k]).max() <= 1e-12 * np.abs(a[k]).max(), k


def test_oracle_grid_injects_the_source_in_its_cell_only():
    """On a grid the source cell gains the force's time integral; the time
    of the next step advances by dt."""
    N = 2
    c = O.cfg(N, pde=4, adv=(0.0, 0.0, 0.0), linear=1, point_source=dict(SRC, x0=(0.6, 0.3, 0.9)))
    g = O.OracleGrid(c, 3, 3, 3)
    g.init(O.make_ic(kind=2, q0=[1.0, 0.0, 0.0, 0.0, 2.5]))
    before = g.luh().copy()
    g.step(0.01)
    assert g.g.t == pytest.approx(0.01)
    diff = np.abs(g.luh() - before).reshape(27, -1).max(1)
    # cell (x, y, z) = (1, 0, 2)
    assert diff[(2 * 3 + 0) * 3 + 1] > 0 and np.count_nonzero(diff) == 1


def test_source_on_a_shared_face_counts_in_both_cells():
    """The inside test includes the cell boundary (0 <= r <= 1), so a source
    on the face between two cells is injected by both, as in the reference;
    x = 0.5 on 4 cells of width 0.25 gives r = 1 and r = 0 exactly."""
    N = 3
    c = O.cfg(N, pde=4, adv=(0.0, 0.0, 0.0), linear=1, point_source=dict(SRC, x0=(0.5, 0.3, 0.7)))
    g = O.OracleGrid(c, 4, 3, 3)
    g.init(O.make_ic(kind=2, q0=[1.0, 0.0, 0.0, 0.0, 2.5]))
    before = g.luh().copy()
    g.step(0.01)
    diff = np.abs(g.luh() - before).reshape(36, -1).max(1)
    # cells (x, y, z) = (1, 0, 2) and (2, 0, 2)
    changed = sorted(np.nonzero(diff)[0].tolist())
    assert changed == [(2 * 3 + 0) * 4 + 1, (2 * 3 + 0) * 4 + 2], changed
