"""GPU parity of the algebraic-source arm (useSource) with Euler under
constant gravity g along -z: S = (0, 0, 0, -rho g, -m_z g) enters the Picard
right-hand side (spaceTimePredictorNonlinear.cpph:227-252), the time-averaged
source block of lFhi and the volume integral (w^3 lShi).  Bar: the oracle
(pinned in tests/test_oracle_math.py::test_source_arm_fixed_point_and_volume_term)
to <= 1e-12 relative per variable, identical Picard counts, for the per-cell
entry points, whole grid steps and the deferred batch form."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
G = 2.5


def rel_err(a, b, nv=5):
    a = np.asarray(a).reshape(-1, nv)
    b = np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    den = np.maximum(np.maximum(scale, 1e-3 * scale.max()), 1e-300)
    return float((np.abs(a - b).max(axis=0) / den).max())


def ic_kw(seed):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.5), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


def cell_state(N, seed):
    c = O.cfg(N, pde=3, grav=G)
    rng = np.random.default_rng(seed)
    return O.solution_adjustment(c, O.make_ic(0, **ic_kw(seed)), rng.uniform(0.3, 0.7, 3), (1 / 27,) * 3)


@pytest.mark.parametrize("N", [2, 3, 5, 7, 9])
def test_predictor_and_volume_integral_with_source(N):
    import exahype_amd as E
    s = E.ADERDGSolver(N, gravity=G)
    c = O.cfg(N, pde=3, grav=G)
    luh = cell_state(N, N)
    dx = np.array([1 / 27, 1 / 30, 1 / 24])
    dt = 0.9 * O.stable_time_step_size(c, luh, dx)
    g = s.spaceTimePredictor(luh, dx, dt)
    o = O.space_time_predictor(c, luh, dx, dt)
    assert g["iters"] == o["iters"]
    for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd", "lFhi", "lFi"):
        assert rel_err(g[k], o[k]) <= TOL, (N, k, rel_err(g[k], o[k]))
    B = N + 1
    assert np.abs(g["lFhi"].reshape(4, B ** 3, 5)[3][:, 3:]).max() > 0  # the source block is populated
    assert rel_err(s.volumeIntegral(o["lFhi"], dx), O.volume_integral(c, o["lFhi"], dx)) <= TOL


@pytest.mark.parametrize("N,dims,steps", [(3, (4, 3, 3), 3), (5, (3, 2, 2), 2), (8, (2, 2, 2), 1)])
def test_grid_steps_with_source(N, dims, steps):
    import exahype_amd as E
    s = E.ADERDGSolver(N, gravity=G)
    c = O.cfg(N, pde=3, grav=G)
    og = O.OracleGrid(c, *dims)
    og.init(O.make_ic(0, **ic_kw(2)))
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    dt_o = og.stable_dt()
    g.compute_dt()
    for k in range(steps):
        g.step(1)
        dt_o = og.step(dt_o)
        t, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o
        assert np.array_equal(g.picard_iterations(), og.iters()), k
    assert rel_err(g.download(), og.luh()) <= TOL
    # the source acted: the gravity-free oracle ends elsewhere
    o0 = O.OracleGrid(O.cfg(N), *dims)
    o0.init(O.make_ic(0, **ic_kw(2)))
    d0 = o0.stable_dt()
    for k in range(steps):
        d0 = o0.step(d0)
    assert rel_err(g.download(), o0.luh()) > 1e-6


def test_deferred_prediction_with_source():
    import exahype_amd as E
    N = 4
    B = N + 1
    s = E.ADERDGSolver(N, gravity=G)
    items = []
    for k in range(4):
        luh = cell_state(N, 10 + k)
        dx = np.array([0.03, 0.02, 0.025])
        dt = 1e-4 * (k + 1)
        outs = (np.zeros(6 * B * B * 5), np.zeros(6 * B * B * 5), np.zeros(B ** 3 * 5))
        s.batch_prediction(*outs, luh, dx, dt)
        items.append((luh, dx, dt, outs))
    s.batch_flush()
    c = O.cfg(N, pde=3, grav=G)
    for luh, dx, dt, (q, f, du) in items:
        o = O.space_time_predictor(c, luh, dx, dt)
        assert rel_err(q, o["lQhbnd"]) <= TOL and rel_err(f, o["lFhbnd"]) <= TOL
        assert rel_err(du, O.volume_integral(c, o["lFhi"], dx)) <= TOL


def test_source_rejected_where_unsupported():
    import exahype_amd as E
    with pytest.raises(RuntimeError, match="gravity"):
        E.ADERDGSolver(5, gravity=G, precision=E.PREC_FP32)
