"""GPU parity of the linear kernel family (the generated glue's isLinear
branch): Cauchy-Kowalevski predictor with the linear volume integral,
coefficient-matrix Riemann solver, fluctuation-form surface integral, and
whole grid steps -- through the C ABI against the reference's golden vectors
(1e-10) and the oracle (<= 1e-12 relative), which tests/test_oracle_linear.py
pins.  The predictor's parity cases use random nodal data: every Taylor
level is then a derivative of O(1) relative conditioning (a smooth state's
high derivatives are cancellation-dominated and say nothing about parity).
"""
import numpy as np
import pytest

from oracle import oracle as O

from test_gpu_aderdg import TOL, rel_err

pytestmark = pytest.mark.gpu

VEL = (0.3, -0.7, 0.5)
PNPM = (1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015)


def exa():
    import exahype_amd
    return exahype_amd


def solver(pde, N):
    return exa().ADERDGSolver(N, pde=pde, velocity=VEL, linear=True)


def ocfg(pde, N):
    return O.cfg(N, pde=pde, adv=VEL, linear=1)


def test_golden_linear_kernels(golden):
    E = exa()
    s = solver(E.PDE_EULER_NCP_TEST, 3)
    lFhi = np.zeros(960)
    for i in range(0, 320, 5):
        lFhi[i + 1], lFhi[320 + i + 2], lFhi[640 + i + 3] = 1.0, 2.0, 3.0
    assert np.abs(s.volumeIntegral(lFhi, [0.5] * 3) - golden["testVolumeIntegralLinear/lduh"]).max() <= 1e-10
    lFhbnd = np.zeros(480)
    for i in range(0, 80, 5):
        lFhbnd[i + 1] = lFhbnd[80 + i + 1] = 1.0
        lFhbnd[160 + i + 2] = lFhbnd[240 + i + 2] = 1.0
        lFhbnd[320 + i + 3] = lFhbnd[400 + i + 3] = 1.0
    out = s.surfaceIntegral(np.arange(320) / 10.0, lFhbnd, [0.1] * 3)
    assert np.abs(out - golden["testSurfaceIntegralLinear/lduh"]).max() <= 1e-10
    FL, FR = s.riemannSolver(np.zeros(80), np.zeros(80), golden["testRiemannSolver/QL"],
                             golden["testRiemannSolver/QR"], 1.40831757919882352703e-03, 1)
    assert np.abs(FL - golden["testRiemannSolverLinear/FL"]).max() <= 1e-10
    assert np.abs(FR - golden["testRiemannSolverLinear/FR"]).max() <= 1e-10


@pytest.mark.parametrize("pde", [4, 5])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_predictor_linear_vs_oracle(pde, N):
    print("linear predictor pde", pde, "N", N, flush=True)
    s = solver(pde, N)
    c = ocfg(pde, N)
    B = N + 1
    rng = np.random.default_rng(N + 10 * pde)
    luh = np.tile([1.0, 0.1, -0.2, 0.3, 2.5], B ** 3) + 0.1 * rng.standard_normal(B ** 3 * 5)
    dx = np.array([1 / 27, 1 / 30, 1 / 24])
    dt = 0.9 * PNPM[N] / sum(abs(VEL[d]) / dx[d] for d in range(3))
    g = s.spaceTimePredictor(luh, dx, dt)
    o = O.space_time_predictor_linear(c, luh, dx, dt)
    for k in ("lQhi", "lQhbnd", "lFhbnd", "lFhi"):
        assert rel_err(g[k], o[k]) <= TOL, (pde, N, k, rel_err(g[k], o[k]))
    # every time-derivative level against its own scale
    gq, oq = g["lQi"].reshape(B + 1, -1), o["lQi"].reshape(B + 1, -1)
    gf, of = g["lFi"].reshape(B, -1), o["lFi"].reshape(B, -1)
    for i in range(B + 1):
        if np.any(oq[i]):
            assert rel_err(gq[i], oq[i]) <= TOL, (pde, N, "lQi", i)
    for i in range(B):
        if np.any(of[i]):
            assert rel_err(gf[i], of[i]) <= TOL, (pde, N, "lFi", i)


@pytest.mark.parametrize("pde", [4, 5])
@pytest.mark.parametrize("N", [2, 5, 9])
def test_riemann_linear_vs_oracle(pde, N):
    s = solver(pde, N)
    c = ocfg(pde, N)
    B = N + 1
    rng = np.random.default_rng(N)
    base = np.tile([1.0, 0.1, -0.2, 0.3, 2.5], B * B)
    for d in range(3):
        QL = base + 0.1 * rng.standard_normal(B * B * 5)
        QR = base + 0.1 * rng.standard_normal(B * B * 5)
        gL, gR = s.riemannSolver(rng.standard_normal(B * B * 5), rng.standard_normal(B * B * 5), QL, QR, 1e-3, d)
        oL, oR = O.riemann_solver_linear(c, QL, QR, 1e-3, d)
        assert rel_err(gL, oL) <= TOL and rel_err(gR, oR) <= TOL, (pde, N, d)


def _sine_ic():
    return O.make_ic(kind=3, nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1])


@pytest.mark.parametrize("N,dims,steps", [(2, (4, 3, 5), 4), (3, (4, 3, 3), 3), (5, (3, 2, 3), 3),
                                          (9, (2, 2, 1), 2)])
def test_linear_grid_steps_vs_oracle(N, dims, steps):
    print("linear grid N", N, dims, flush=True)
    E = exa()
    s = solver(E.PDE_ADVECTION_NCP, N)
    og = O.OracleGrid(ocfg(4, N), *dims)
    og.init(_sine_ic())
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    dt_o = og.stable_dt()
    dt_g = g.compute_dt()
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o
    for k in range(steps):
        g.step(1)
        dt_o = og.step(dt_o)
        t, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (k, dt_g, dt_o)
    assert rel_err(g.download(), og.luh()) <= TOL


def test_linear_rejects_pdes_without_a_coefficient_matrix():
    E = exa()
    with pytest.raises(RuntimeError, match="coefficient matrix"):
        E.ADERDGSolver(3, pde=E.PDE_ADVECTION, velocity=VEL, linear=True)
    with pytest.raises(RuntimeError, match="coefficient matrix"):
        E.ADERDGSolver(3, linear=True)
