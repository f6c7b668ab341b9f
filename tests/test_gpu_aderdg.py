"""GPU parity of the ADER-DG path: HIP kernels through the C ABI vs the oracle.

Bar: the plain FP64 path agrees with the oracle to <= 1e-12 relative
(||a-b||_inf / ||b||_inf per variable), Picard iteration counts are equal,
and the reference's golden vectors pass at the reference's own 1e-10.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def exa():
    import exahype_amd
    return exahype_amd


def rel_err(a, b, nv=5):
    """max over variables of ||a_v - b_v||_inf / ||b_v||_inf.  A variable whose
    scale is below 1e-3 of the largest variable's (e.g. a momentum component
    that is zero up to cancellation because the flow has no velocity in that
    direction) is measured against that floor: its rounding differences come
    from O(1) intermediate terms, not from its own size."""
    a = np.asarray(a).reshape(-1, nv)
    b = np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    den = np.maximum(scale, 1e-3 * scale.max())
    den = np.maximum(den, 1e-300)
    return float((np.abs(a - b).max(axis=0) / den).max())


def euler_state(rng, nnodes, amp=0.1):
    rho = 1.0 + amp * rng.random(nnodes)
    u = amp * rng.normal(size=(nnodes, 3))
    p = 1.0 + amp * rng.random(nnodes)
    E = p / 0.4 + 0.5 * rho * (u ** 2).sum(1)
    return np.column_stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2], E]).ravel()


def smooth_cell_state(N, seed):
    """A smooth state at the GL nodes of one cell (Picard converges like in a real run)."""
    c = O.cfg(N)
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 3, size=(8, 3)).astype(float)
    ic = O.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0,
                   amp=1e-2, nmodes=8, k=k, a=rng.uniform(-1, 1, 8), phi=rng.uniform(0, 6.28, 8))
    centre = rng.uniform(0.3, 0.7, 3)
    return O.solution_adjustment(c, ic, centre, (1 / 27, 1 / 27, 1 / 27))


# ---------------------------------------------------------------- goldens (N=3, npar=1)
@pytest.fixture(scope="module")
def s3p1():
    return exa().ADERDGSolver(3, nvar=5, npar=1)


def restride(a, nnodes, nv=5, nd=6):
    out = np.zeros(nnodes * nd)
    out.reshape(nnodes, nd)[:, :nv] = np.asarray(a).reshape(nnodes, nv)
    return out


def unstride(a, nnodes, nv=5, nd=6):
    return np.asarray(a).reshape(nnodes, nd)[:, :nv].ravel()


def test_golden_predictor(s3p1, golden):
    luh = restride(golden["testSpaceTimePredictorNonlinear/luh"], 64)
    out = s3p1.spaceTimePredictor(luh, [0.05] * 3, 1.083937460199773e-3)
    g = "testSpaceTimePredictorNonlinear/"
    assert np.abs(unstride(out["lQi"], 256) - golden[g + "lQi"]).max() <= 1e-10
    lFi = out["lFi"].reshape(4, 4, 4, 4, 4, 5)[:, :, :, :, :3, :].ravel()
    assert np.abs(lFi - golden[g + "lFi"]).max() <= 1e-10
    assert np.abs(unstride(out["lQhi"], 64) - golden[g + "lQhi"]).max() <= 1e-10
    assert np.abs(out["lFhi"][:960] - golden[g + "lFhi"]).max() <= 1e-10
    assert np.abs(unstride(out["lQhbnd"], 96) - golden[g + "lQhbnd"]).max() <= 1e-10
    assert np.abs(out["lFhbnd"] - golden[g + "lFhbnd"]).max() <= 1e-10


def test_golden_volume_surface_riemann_update(s3p1, golden):
    lFhi = np.zeros(1280)
    for i in range(0, 320, 5):
        lFhi[i + 1] = lFhi[320 + i + 2] = lFhi[640 + i + 3] = 1.0
    assert np.abs(s3p1.volumeIntegral(lFhi, [0.05] * 3) - golden["testVolumeIntegralNonlinear/lduh"]).max() <= 1e-10
    lFhbnd = np.zeros(480)
    for i in range(0, 80, 5):
        lFhbnd[i + 1] = lFhbnd[80 + i + 1] = 1.0
        lFhbnd[160 + i + 2] = lFhbnd[240 + i + 2] = 1.0
        lFhbnd[320 + i + 3] = lFhbnd[400 + i + 3] = 1.0
    out = s3p1.surfaceIntegral(np.arange(320) / 10.0, lFhbnd, [0.1] * 3)
    assert np.abs(out - golden["testSurfaceIntegralNonlinear/lduh"]).max() <= 1e-10
    FL, FR = s3p1.riemannSolver(np.arange(80.0) + 1, np.arange(80.0) - 1,
                                restride(golden["testRiemannSolver/QL"], 16),
                                restride(golden["testRiemannSolver/QR"], 16), 1.40831757919882352703e-03, 1)
    assert np.abs(FL - golden["testRiemannSolverNonlinear/FL"]).max() <= 1e-10
    assert np.abs(FR - golden["testRiemannSolverNonlinear/FR"]).max() <= 1e-10
    luh = np.zeros(384)
    luh[0::6] = 1.0
    luh[4::6] = 2.5
    out = s3p1.solutionUpdate(luh, np.arange(320.0), 1.40831757919882352703e-03)
    assert np.abs(unstride(out, 64) - golden["testSolutionUpdate/luh"]).max() <= 1e-10


# ---------------------------------------------------------------- per-cell vs oracle
@pytest.mark.parametrize("contraction", [1, 2, 3])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_predictor_vs_oracle(N, contraction):
    s = exa().ADERDGSolver(N, contraction=contraction)
    c = O.cfg(N)
    B = N + 1
    for seed in range(2):
        luh = smooth_cell_state(N, seed)
        dx = np.array([1 / 27, 1 / 27, 1 / 27])
        dt = 0.9 * O.stable_time_step_size(c, luh, dx)
        g = s.spaceTimePredictor(luh, dx, dt)
        o = O.space_time_predictor(c, luh, dx, dt)
        assert g["iters"] == o["iters"], (N, seed)
        for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd", "lFhi", "lFi"):
            assert rel_err(g[k], o[k]) <= TOL, (N, k, rel_err(g[k], o[k]))


@pytest.mark.parametrize("N", [8, 9])
def test_staged_predictor_vs_oracle(N):
    """Orders 8 and 9 stage the space-time iterate in global memory; same bar."""
    s = exa().ADERDGSolver(N)
    c = O.cfg(N)
    for seed in range(2):
        luh = smooth_cell_state(N, seed)
        dx = np.array([1 / 27, 1 / 27, 1 / 27])
        dt = 0.9 * O.stable_time_step_size(c, luh, dx)
        g = s.spaceTimePredictor(luh, dx, dt)
        o = O.space_time_predictor(c, luh, dx, dt)
        assert g["iters"] == o["iters"], (N, seed)
        for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd", "lFhi", "lFi"):
            assert rel_err(g[k], o[k]) <= TOL, (N, k, rel_err(g[k], o[k]))


@pytest.mark.parametrize("N", [5, 6, 7])
def test_staged_form_at_lower_orders(N):
    """The staged predictor selected explicitly (EXA_CONTRACT_STAGED) at orders
    that also have register-resident forms: per-cell outputs and a grid step."""
    E = exa()
    s = E.ADERDGSolver(N, contraction=E.CONTRACT_STAGED)
    c = O.cfg(N)
    luh = smooth_cell_state(N, 3)
    dx = np.array([1 / 27, 1 / 27, 1 / 27])
    dt = 0.9 * O.stable_time_step_size(c, luh, dx)
    g = s.spaceTimePredictor(luh, dx, dt)
    o = O.space_time_predictor(c, luh, dx, dt)
    assert g["iters"] == o["iters"]
    for k in ("lQi", "lQhi", "lQhbnd", "lFhbnd", "lFhi", "lFi"):
        assert rel_err(g[k], o[k]) <= TOL, (N, k, rel_err(g[k], o[k]))
    dims = (4, 3, 3)
    og = O.OracleGrid(c, *dims)
    og.init(O.make_ic(0, **_ic_kw(5)))
    gr = E.Grid(s, *dims)
    gr.upload(og.luh().copy())
    dt_o = og.stable_dt()
    gr.compute_dt()
    for k in range(2):
        gr.step(1)
        dt_o = og.step(dt_o)
        gr.sync()
        assert np.array_equal(gr.picard_iterations(), og.iters()), k
    assert rel_err(gr.download(), og.luh()) <= TOL


def test_staged_form_rejects_low_orders_and_matrix_core_precisions():
    E = exa()
    with pytest.raises(RuntimeError, match="staged"):
        E.ADERDGSolver(3, contraction=E.CONTRACT_STAGED)
    with pytest.raises(RuntimeError, match="staged"):
        E.ADERDGSolver(5, contraction=E.CONTRACT_STAGED, precision=E.PREC_BF16_MFMA)


def test_staged_orders_reject_matrix_core_precisions():
    E = exa()
    with pytest.raises(RuntimeError, match="FP32"):
        E.ADERDGSolver(9, precision=E.PREC_FP32_MFMA)
    E.ADERDGSolver(9, precision=E.PREC_FP32)  # FP32 sweeps + FP64 correction are built for 8 and 9


@pytest.mark.parametrize("N", [2, 3, 5, 7, 9])
def test_face_and_cell_kernels_vs_oracle(N):
    s = exa().ADERDGSolver(N)
    c = O.cfg(N)
    B = N + 1
    rng = np.random.default_rng(N)
    dx = np.array([0.02, 0.03, 0.025])
    for d in range(3):
        QL, QR = euler_state(rng, B * B), euler_state(rng, B * B)
        FL, FR = rng.normal(size=B * B * 5), rng.normal(size=B * B * 5)
        g = s.riemannSolver(FL, FR, QL, QR, 1e-3, d)
        o = O.riemann_solver(c, FL, FR, QL, QR, 1e-3, d)
        assert rel_err(g[0], o[0]) <= TOL and rel_err(g[1], o[1]) <= TOL
    lduh = rng.normal(size=B ** 3 * 5)
    lFhbnd = rng.normal(size=6 * B * B * 5)
    assert rel_err(s.surfaceIntegral(lduh, lFhbnd, dx), O.surface_integral(c, lduh, lFhbnd, dx)) <= TOL
    lFhi = rng.normal(size=4 * B ** 3 * 5)
    assert rel_err(s.volumeIntegral(lFhi, dx), O.volume_integral(c, lFhi, dx)) <= TOL
    luh = euler_state(rng, B ** 3)
    assert rel_err(s.solutionUpdate(luh, lduh, 1e-3), O.solution_update(c, luh, lduh, 1e-3)) <= TOL
    dg, do = s.stableTimeStepSize(luh, dx), O.stable_time_step_size(c, luh, dx)
    assert abs(dg - do) <= 1e-14 * do


def test_initial_conditions_vs_oracle():
    E = exa()
    for N in (3, 5):
        s = E.ADERDGSolver(N)
        c = O.cfg(N)
        rng = np.random.default_rng(5)
        kw = dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                  nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                  phi=rng.uniform(0, 6.28, 8))
        for centre in ([0.41, 0.52, 0.33], [0.05, 0.95, 0.5]):
            g = s.adjustSolution(centre, [1 / 27] * 3, 0.0, E.make_ic(0, **kw))
            o = O.solution_adjustment(c, O.make_ic(0, **kw), centre, [1 / 27] * 3)
            assert rel_err(g, o) <= 1e-14


# ---------------------------------------------------------------- whole steps vs oracle grid
def _ic_kw(seed):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


@pytest.mark.parametrize("contraction", [1, 2, 3])
@pytest.mark.parametrize("N,n,steps", [(3, 6, 5), (5, 4, 5), (2, 5, 3), (7, 2, 2)])
def test_grid_steps_vs_oracle(N, n, steps, contraction):
    E = exa()
    s = E.ADERDGSolver(N, contraction=contraction)
    c = O.cfg(N)
    og = O.OracleGrid(c, n, n + 1, n - 1)
    og.init(O.make_ic(0, **_ic_kw(11)))
    g = E.Grid(s, n, n + 1, n - 1)
    g.upload(og.luh().copy())
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


@pytest.mark.parametrize("N,dims,steps", [(9, (2, 3, 1), 2), (8, (3, 2, 2), 2)])
def test_staged_grid_steps_vs_oracle(N, dims, steps):
    E = exa()
    s = E.ADERDGSolver(N)
    c = O.cfg(N)
    og = O.OracleGrid(c, *dims)
    og.init(O.make_ic(0, **_ic_kw(11)))
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    dt_o = og.stable_dt()
    assert abs(g.compute_dt() - dt_o) <= 1e-14 * dt_o
    for k in range(steps):
        g.step(1)
        dt_o = og.step(dt_o)
        t, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (k, dt_g, dt_o)
        assert np.array_equal(g.picard_iterations(), og.iters()), k
    assert rel_err(g.download(), og.luh()) <= TOL


def test_staged_grid_more_cells_than_slots():
    """The order-8 grid walks more cells than it has staging slots (one per
    resident workgroup): every cell must match the per-cell predictor path of
    the oracle after one step."""
    E = exa()
    N, dims = 8, (9, 7, 5)  # 315 cells > 256 slots
    s = E.ADERDGSolver(N)
    c = O.cfg(N)
    og = O.OracleGrid(c, *dims)
    og.init(O.make_ic(0, **_ic_kw(3)))
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    dt_o = og.stable_dt()
    g.compute_dt()
    g.step(1)
    og.step(dt_o)
    g.sync()
    assert np.array_equal(g.picard_iterations(), og.iters())
    assert rel_err(g.download(), og.luh()) <= TOL


def test_graph_replay_matches_eager():
    E = exa()
    s = E.ADERDGSolver(3)
    grids = []
    for use_graph in (False, True):
        g = E.Grid(s, 5, 4, 3)
        g.set_initial_conditions(E.make_ic(0, **_ic_kw(2)))
        g.compute_dt()
        g.use_graph(use_graph)
        for _ in range(3):  # several replays with host synchronisation between them
            g.step(2)
            g.sync()
        grids.append(g.download())
    assert np.array_equal(grids[0], grids[1])


def test_conservation_large_grid():
    """Size-independent property at a real size: weighted totals are conserved."""
    from exahype_amd.tables import dg_math
    E = exa()
    N = 5
    s = E.ADERDGSolver(N)
    g = E.Grid(s, 27, 27, 27)
    g.set_initial_conditions(E.make_ic(0, **_ic_kw(4)))
    g.compute_dt()
    w = dg_math.gl_nodes_weights(N)[1]
    w3 = np.einsum("z,y,x->zyx", w, w, w).ravel()

    def tot():
        return np.einsum("n,cnv->v", w3, g.download().reshape(-1, (N + 1) ** 3, 5))
    t0 = tot()
    g.step(4)
    g.sync()
    t1 = tot()
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()


def test_autotune_records_a_variant():
    """The autotuner times the four launch variants and picks the fastest; a
    solver left at the tuned defaults still reproduces the oracle."""
    E = exa()
    w, c, ms = E.autotune(3, E.PREC_FP64, 12)
    assert w in (1, 2) and c in (1, 2, 3)
    assert np.all(np.isfinite(ms)) and np.all(ms > 0)
    assert ms[(w - 1) * 3 + (c - 1)] == ms.min()
    s = E.ADERDGSolver(3)
    c3 = O.cfg(3)
    luh = smooth_cell_state(3, 0)
    dx = np.array([1 / 27] * 3)
    dt = 0.9 * O.stable_time_step_size(c3, luh, dx)
    g = s.spaceTimePredictor(luh, dx, dt)
    o = O.space_time_predictor(c3, luh, dx, dt)
    assert rel_err(g["lQhbnd"], o["lQhbnd"]) <= TOL
