"""GPU parity of the generic PDE arms (flux / non-conservative product):
linear advection in flux form and in non-conservative form, and Euler with
the test non-conservative product of the reference's Euler kernel test.
HIP kernels through the C ABI vs the oracle, FP64, <= 1e-12 relative
(||a-b||_inf / ||b||_inf per variable) and identical Picard iteration counts.
The oracle's arm is pinned by tests/test_oracle_ncp.py (flux form == NCP form
for constant-coefficient advection).
"""
import numpy as np
import pytest

from oracle import oracle as O

from test_gpu_aderdg import TOL, rel_err, smooth_cell_state

pytestmark = pytest.mark.gpu

VEL = (0.3, -0.7, 0.5)
# The source slot of lFi holds S = -ncp(Q, grad Q) of the last Picard sweep.
# The nodal gradient sum_n dudx[l][n] q_n / dx cancels strongly for a smooth
# iterate (the dudx rows sum to zero; |dudx| / dx ~ 1e3 at N = 5, dx = 1/27),
# so the ~1e-15 rounding differences of the iterate (which itself agrees to
# <= 1e-12) reach ~1e-12 of |S| at N = 5 and ~1e-11 at N = 7 (the
# cancellation grows with |dudx| / dx).  That slot and its time average lShi
# (the source block of lFhi) are held to TOL_GRAD; the iterate, the faces and
# the grid states stay at TOL.  A wrong sign or index of the arm shows up at
# O(1) in the slot.
TOL_GRAD = 1e-10
# device PDE id -> oracle pde id
PDES = {3: 1, 4: 4, 5: 5}


def exa():
    import exahype_amd
    return exahype_amd


def solver(pde, N):
    return exa().ADERDGSolver(N, pde=pde, velocity=VEL)


def ocfg(pde, N):
    return O.cfg(N, pde=PDES[pde], adv=VEL)


@pytest.mark.parametrize("pde", [3, 4, 5])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_predictor_vs_oracle(pde, N):
    print("predictor pde", pde, "N", N, flush=True)
    s = solver(pde, N)
    c = ocfg(pde, N)
    for seed in range(2):
        luh = smooth_cell_state(N, seed)
        dx = np.array([1 / 27, 1 / 30, 1 / 24])
        dt = 0.9 * O.stable_time_step_size(c, luh, dx)
        g = s.spaceTimePredictor(luh, dx, dt)
        o = O.space_time_predictor(c, luh, dx, dt)
        assert g["iters"] == o["iters"], (N, seed)
        tol_src = TOL_GRAD if pde != 3 else TOL
        for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd"):
            assert rel_err(g[k], o[k]) <= TOL, (pde, N, k, rel_err(g[k], o[k]))
        nb = (N + 1) ** 3 * 5
        assert rel_err(g["lFhi"][:3 * nb], o["lFhi"][:3 * nb]) <= TOL, (pde, N, "lFhi flux")
        assert rel_err(g["lFhi"][3 * nb:], o["lFhi"][3 * nb:]) <= tol_src, (pde, N, "lShi")
        gFi, oFi = g["lFi"].reshape(-1, 4, 5), o["lFi"].reshape(-1, 4, 5)
        assert rel_err(gFi[:, :3], oFi[:, :3]) <= TOL, (pde, N, "lFi flux")
        assert rel_err(gFi[:, 3], oFi[:, 3]) <= tol_src, (pde, N, "lFi source", rel_err(gFi[:, 3], oFi[:, 3]))
        if pde == 4:
            assert not np.any(g["lFhbnd"])


@pytest.mark.parametrize("pde", [3, 4, 5])
@pytest.mark.parametrize("N", [2, 5])
def test_riemann_vs_oracle(pde, N):
    s = solver(pde, N)
    c = ocfg(pde, N)
    B = N + 1
    rng = np.random.default_rng(N + 10 * pde)
    for d in range(3):
        QL = smooth_cell_state(N, d)[:B * B * 5]
        QR = smooth_cell_state(N, d + 5)[:B * B * 5]
        FL0 = rng.standard_normal(B * B * 5)
        FR0 = rng.standard_normal(B * B * 5)
        gL, gR = s.riemannSolver(FL0, FR0, QL, QR, 1e-3, d)
        oL, oR = O.riemann_solver(c, FL0, FR0, QL, QR, 1e-3, d)
        assert rel_err(gL, oL) <= TOL and rel_err(gR, oR) <= TOL, (pde, N, d)
        if pde != 3:
            assert not np.allclose(gL, gR)  # the fluctuation splits FL and FR


@pytest.mark.parametrize("pde", [3, 4, 5])
def test_volume_integral_and_dt_vs_oracle(pde):
    N = 4
    B = N + 1
    s = solver(pde, N)
    c = ocfg(pde, N)
    rng = np.random.default_rng(pde)
    lFhi = rng.standard_normal(4 * B ** 3 * 5)  # flux blocks are ignored without a flux
    dx = np.array([0.05, 0.04, 0.03])
    assert rel_err(s.volumeIntegral(lFhi, dx), O.volume_integral(c, lFhi, dx)) <= TOL
    luh = smooth_cell_state(N, 4)
    assert abs(s.stableTimeStepSize(luh, dx) - O.stable_time_step_size(c, luh, dx)) <= 1e-14 * O.stable_time_step_size(
        c, luh, dx)


def _grid_pair(pde, N, dims, ic):
    E = exa()
    s = solver(pde, N)
    og = O.OracleGrid(ocfg(pde, N), *dims)
    og.init(ic)
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    return g, og


def _sine_ic():
    return O.make_ic(kind=3, nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1])


def _vortex_ic():
    rng = np.random.default_rng(11)
    return O.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0,
                     amp=1e-2, nmodes=8, k=rng.integers(1, 3, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                     phi=rng.uniform(0, 6.28, 8))


@pytest.mark.parametrize("pde", [3, 4, 5])
@pytest.mark.parametrize("N,dims,steps", [(3, (4, 3, 3), 3), (5, (3, 2, 3), 2), (7, (2, 2, 2), 2)])
def test_grid_steps_vs_oracle(pde, N, dims, steps):
    if pde == 5 and N == 7:
        # the reference test's artificial product drives the pressure of a
        # node negative in the second step (-8.9e-3 by the oracle); that
        # state is checked to be reported in the next test
        steps = 1
    print("grid pde", pde, "N", N, dims, flush=True)
    g, og = _grid_pair(pde, N, dims, _vortex_ic() if pde == 5 else _sine_ic())
    dt_o = og.stable_dt()
    dt_g = g.compute_dt()
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o
    for k in range(steps):
        g.step(1)
        dt_o = og.step(dt_o)
        t, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (k, dt_g, dt_o)
        assert np.array_equal(g.picard_iterations(), og.iters()), k
    assert rel_err(g.download(), og.luh()) <= TOL


def test_negative_pressure_of_the_test_product_is_reported():
    """Euler with the reference test's product at N=7 on 2^3 cells: after two
    steps a node holds a negative pressure (the oracle's state, below), where
    the Euler wave speed sqrt(gamma p / rho) is not a number.  The device dt
    keeps that NaN (std::max(|lambda|, m) as stableTimeStepSize.cpph:45-47
    accumulates it) and the sync call reports the state instead of stepping on."""
    E = exa()
    N, dims = 7, (2, 2, 2)
    g, og = _grid_pair(5, N, dims, _vortex_ic())
    dt_o = og.stable_dt()
    g.compute_dt()
    for _ in range(2):
        dt_o = og.step(dt_o)
    u = og.luh().reshape(-1, 5)
    p = 0.4 * (u[:, 4] - 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0])
    assert p.min() < 0.0
    g.step(2)
    with pytest.raises(E.NumericsError):
        g.sync()


@pytest.mark.parametrize("N", [2, 3, 5])
def test_ncp_form_matches_flux_form_on_the_gpu(N):
    """The identity that pins the oracle's arm, on the device: advection in
    non-conservative form and in flux form give the same grid solution up to
    the 13-digit tables' inconsistency (see tests/test_oracle_ncp.py)."""
    E = exa()
    dims = (3, 4, 2)
    og = O.OracleGrid(O.cfg(N, pde=1, adv=VEL), *dims)
    og.init(_sine_ic())
    grids = []
    for pde in (3, 4):
        g = E.Grid(solver(pde, N), *dims)
        g.upload(og.luh().copy())
        g.compute_dt()
        g.step(4)
        g.sync()
        grids.append(g.download())
    assert rel_err(grids[1], grids[0]) < 1e-8


def test_generic_pdes_reject_unsupported_paths():
    E = exa()
    with pytest.raises(RuntimeError):
        E.ADERDGSolver(8, pde=E.PDE_ADVECTION_NCP, velocity=VEL)
    with pytest.raises(RuntimeError):
        E.ADERDGSolver(3, pde=E.PDE_ADVECTION_NCP, velocity=VEL, precision=E.PREC_FP32)
    s = solver(4, 3)
    B = 4
    with pytest.raises(RuntimeError, match="batched"):
        s.batch_prediction(np.zeros(6 * B * B * 5), np.zeros(6 * B * B * 5), np.zeros(B ** 3 * 5),
                           np.ones(B ** 3 * 5), [0.1] * 3, 1e-3)
