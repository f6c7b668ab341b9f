"""Pins of the oracle's non-conservative-product arm (useNCP, useFlux=false).

No golden vector of the reference exercises a non-zero non-conservative
product: testSpaceTimePredictorNonlinear runs the <true,true,true> arm with the
product forced to zero (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:675) and
testRiemannSolverNonlinear runs useNCP=false.  What fixes the arm is an
identity of the scheme: for constant-coefficient linear advection the
non-conservative form q_t + a.grad q = 0 (no flux, B_d = a_d I) and the flux
form q_t + div(a q) = 0 give the same discrete solution.  The GL quadrature
integrates the volume terms exactly, and integration by parts turns
  int phi (-a q_x)  into  int phi' a q - [phi a q]
where the cell-trace term [phi a q] plus the fluctuation +-ncp/2 of the
Riemann solver (riemannSolverNonlinear.cpph:95-113) equals the flux form's
Rusanov flux 1/2 a (qL + qR) - 1/2 s (qR - qL).  In the predictor the
gradient (dudx / dx) of the iterate and the Kxi contraction of the flux agree
node by node.  A wrong sign of the fluctuation or of the fused source, a
dropped 1/dx or a transposed dudx breaks the identity.

The identity is exact for exact tables; the reference's Kxi and dudx are
13-digit literals (Kxi[n][l] and w_l dudx[l][n] differ by up to 7e-13), so
the two forms agree to 1e-13..3e-10 after four steps at orders 1..7.  The
tolerance TOL = 1e-8 is far below what a mistake of the arm gives: with the
fluctuation sign flipped the same runs differ by 5e-4..5e-2, with dudx
transposed by 0.6..1.4 (checked on a modified copy of the oracle).
"""
import numpy as np
import pytest

from exahype_amd.tables import dg_math
from oracle import oracle as orc

VEL = (0.3, -0.7, 0.5)
TOL = 1e-8


def _pair(order):
    return (orc.cfg(order, pde=1, adv=VEL), orc.cfg(order, pde=4, adv=VEL))


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _sine_ic():
    return orc.make_ic(kind=3, nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1])


@pytest.mark.parametrize("order", [1, 2, 3, 5, 7])
def test_ncp_advection_matches_flux_form_on_a_grid(order):
    cf, cn = _pair(order)
    gf, gn = orc.OracleGrid(cf, 3, 4, 2), orc.OracleGrid(cn, 3, 4, 2)
    ic = _sine_ic()
    gf.init(ic)
    gn.init(ic)
    dt = 0.5 * gf.stable_dt()
    assert gn.stable_dt() == pytest.approx(2 * dt, rel=1e-15)
    for _ in range(4):
        gf.step(dt)
        gn.step(dt)
    assert _rel(gn.luh(), gf.luh()) < TOL
    assert np.array_equal(gn.iters(), gf.iters())


@pytest.mark.parametrize("order", [2, 4])
def test_ncp_advection_predictor_matches_flux_form(order):
    cf, cn = _pair(order)
    B = order + 1
    rng = np.random.default_rng(7 + order)
    luh = 1.0 + 0.1 * rng.standard_normal(B ** 3 * 5)
    dx = np.array([0.05, 0.08, 0.03])
    dt = 2e-3
    of = orc.space_time_predictor(cf, luh, dx, dt)
    on = orc.space_time_predictor(cn, luh, dx, dt)
    assert on["iters"] == of["iters"]
    assert _rel(on["lQi"], of["lQi"]) < TOL
    assert _rel(on["lQhbnd"], of["lQhbnd"]) < TOL
    assert not np.any(on["lFhbnd"])  # no flux: the reference zeroes lFhbnd
    # the volume terms agree up to the trace term the Riemann solver supplies:
    # lduh_ncp = lduh_flux - sum_faces (trace of a q) * basis value at the face
    lduh_f = orc.volume_integral(cf, of["lFhi"], dx)
    lduh_n = orc.volume_integral(cn, on["lFhi"], dx)
    M = dg_math.operators(order)
    FL, FR, w = np.asarray(M["FLCoeff"]), np.asarray(M["FRCoeff"]), np.asarray(M["gl_weights"])
    qh = on["lQhi"].reshape(B, B, B, 5)
    trace = np.zeros((B, B, B, 5))
    for d, ax in ((0, 2), (1, 1), (2, 0)):  # x is the fastest index of [z][y][x]
        qR = np.tensordot(qh, FR, axes=([ax], [0]))  # value at the + face
        qL = np.tensordot(qh, FL, axes=([ax], [0]))
        shape = [1, 1, 1, 1]
        shape[ax] = B
        wt = [w, w, w]
        wt[ax] = np.ones(B)
        wface = (wt[0][:, None, None] * wt[1][None, :, None] * wt[2][None, None, :])[..., None]
        trace += VEL[d] / dx[d] * wface * (np.expand_dims(qR, ax) * FR.reshape(shape)
                                           - np.expand_dims(qL, ax) * FL.reshape(shape))
    assert _rel(lduh_n, lduh_f - trace.ravel()) < TOL


def test_riemann_fluctuation_of_equal_states_is_the_product():
    """QL == QR: the gradient row is zero and the Rusanov part is the mean of
    the incoming fluxes, so FL - FR is the product at Q (ncp(Q, 0))."""
    order = 3
    B = order + 1
    c = orc.cfg(order, pde=5)
    rng = np.random.default_rng(3)
    Q = np.tile(np.array([1.1, 0.1, -0.2, 0.3, 2.6]), B * B) + 0.01 * rng.standard_normal(B * B * 5)
    FL0 = rng.standard_normal(B * B * 5)
    FR0 = rng.standard_normal(B * B * 5)
    for d in range(3):
        FL, FR = orc.riemann_solver(c, FL0, FR0, Q, Q, 0.0, d)
        mean = 0.5 * (FL0 + FR0)
        ncp = np.stack([orc.ncp(c, Q[5 * i:5 * i + 5], np.zeros(15)) for i in range(B * B)]).ravel()
        assert np.allclose(FL - FR, ncp, rtol=0, atol=1e-14)
        assert np.allclose(0.5 * (FL + FR), mean, rtol=0, atol=1e-14)


def test_riemann_ncp_advection_is_upwind_for_zero_flux():
    """Advection in NCP form with FL = FR = 0 on entry and speed s = |a_d|:
    FL = -s/2 (qR - qL) + a_d/2 (qR - qL) and FR = -s/2 (qR - qL) - a_d/2 (qR - qL),
    i.e. the left cell sees only right-moving and the right cell only
    left-moving fluctuations."""
    order = 2
    B = order + 1
    c = orc.cfg(order, pde=4, adv=VEL)
    rng = np.random.default_rng(5)
    QL = rng.standard_normal(B * B * 5)
    QR = rng.standard_normal(B * B * 5)
    z = np.zeros(B * B * 5)
    for d in range(3):
        FL, FR = orc.riemann_solver(c, z, z, QL, QR, 0.0, d)
        s = abs(VEL[d])
        jump = QR - QL
        assert np.allclose(FL, -0.5 * s * jump + 0.5 * VEL[d] * jump, rtol=0, atol=1e-15)
        assert np.allclose(FR, -0.5 * s * jump - 0.5 * VEL[d] * jump, rtol=0, atol=1e-15)
