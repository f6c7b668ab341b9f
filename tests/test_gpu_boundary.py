"""Domain boundaries on the GPU against the oracle (oracle/aderdg.c
orc_boundary_conditions and the non-periodic faces of orc_grid_step):

* the per-face boundaryConditions entry, every face and kind, to 1e-13;
* grids closed on some or all axes (outflow, reflecting wall, exact data),
  eager and graphed, several steps, to 1e-12 with equal Picard counts;
* a gas at rest between reflecting walls stays at rest;
* the PDEs without boundary values refuse a non-periodic grid.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b, nv=5):
    a, b = np.asarray(a).reshape(-1, nv), np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    return float((np.abs(a - b).max(axis=0) / np.maximum(scale, 1e-3 * scale.max())).max())


def _wave_kw():
    return dict(rho_inf=1.0, p_inf=1.0, amp=0.2, uinf=(0.7, -0.4, 0.3), nmodes=1, k=[(1, 2, 1)], a=[1.0], phi=[0.4])


def _vortex_kw(seed=3):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 3, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


def _bcs(kind, ic_kind, kw):
    import exahype_amd as E
    if kind == O.BC_EXACT:
        return E.make_bc(kind, E.make_ic(ic_kind, **kw)), O.make_bc(kind, O.make_ic(ic_kind, **kw))
    return E.make_bc(kind), O.make_bc(kind)


@pytest.mark.parametrize("kind", [O.BC_OUTFLOW, O.BC_REFLECTING, O.BC_EXACT])
@pytest.mark.parametrize("N", [2, 5])
def test_face_boundary_conditions_vs_oracle(kind, N):
    import exahype_amd as E
    s = E.ADERDGSolver(N)
    c = O.cfg(N)
    B2 = (N + 1) ** 2
    rng = np.random.default_rng(N + kind)
    ebc, obc = _bcs(kind, 0, _vortex_kw())
    for face in range(6):
        d = face // 2
        rho = 1.0 + 0.1 * rng.random(B2)
        u = 0.3 * rng.standard_normal((B2, 3))
        p = 1.0 + 0.1 * rng.random(B2)
        QIn = np.column_stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2],
                               p / 0.4 + 0.5 * rho * (u ** 2).sum(1)]).ravel()
        FIn = rng.standard_normal(5 * B2)
        args = ((0.41, 0.57, 0.33), (0.05, 0.04, 0.07), 0.3, 2e-3, face, d)
        Fg, Qg = s.boundaryConditions(ebc, FIn, QIn, *args)
        Fo, Qo = O.boundary_conditions(c, obc, FIn, QIn, *args)
        assert rel(Qg, Qo) <= 1e-13, (face, rel(Qg, Qo))
        assert rel(Fg, Fo) <= 1e-13, (face, rel(Fg, Fo))


def _pair(N, dims, periodic, kind, ic_kind, kw, offset=(0.0, 0.0, 0.0), width=(1.0, 1.0, 1.0)):
    import exahype_amd as E
    c = O.cfg(N)
    og = O.OracleGrid(c, *dims)
    for d in range(3):
        og.g.offset[d], og.g.width[d] = offset[d], width[d]
    og.init(O.make_ic(ic_kind, **kw))
    ebc, obc = _bcs(kind, ic_kind, kw)
    og.set_boundary(periodic, obc)
    s = E.ADERDGSolver(N)
    g = E.Grid(s, *dims, offset=offset, width=width)
    g.upload(og.luh().copy())
    g.set_boundary(periodic, ebc)
    return g, og


def _run(g, og, steps, graph):
    g.use_graph(graph)
    dt_o = og.stable_dt()
    dt_g = g.compute_dt()
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o
    for k in range(steps):
        g.step(1)
        dt_o = og.step(dt_o)
        t, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= TOL * dt_o, (k, dt_g, dt_o)
        assert np.array_equal(g.picard_iterations(), og.iters()), k
        assert rel(g.download(), og.luh()) <= TOL, (k, rel(g.download(), og.luh()))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("kind", [O.BC_OUTFLOW, O.BC_REFLECTING, O.BC_EXACT])
def test_closed_grid_vs_oracle(kind, graph):
    print("closed grid, kind", kind, "graph", graph, flush=True)
    g, og = _pair(3, (4, 5, 3), (0, 0, 0), kind, 4, _wave_kw())
    _run(g, og, 3, graph)


@pytest.mark.parametrize("periodic", [(1, 0, 1), (0, 1, 1), (0, 0, 1)])
def test_partly_periodic_grid_vs_oracle(periodic):
    print("partly periodic", periodic, flush=True)
    g, og = _pair(4, (3, 4, 5), periodic, O.BC_EXACT, 0, _vortex_kw(), offset=(-0.2, 0.1, 0.3),
                  width=(1.2, 0.9, 0.8))
    _run(g, og, 3, True)


def test_boundaries_reset_to_periodic():
    """Closing and reopening the axes gives the periodic step again."""
    import exahype_amd as E
    g, og = _pair(3, (3, 3, 3), (0, 0, 0), O.BC_REFLECTING, 0, _vortex_kw())
    g.set_boundary((1, 1, 1))
    og.set_boundary((1, 1, 1), O.make_bc(O.BC_OUTFLOW))
    _run(g, og, 2, True)


def test_gas_at_rest_between_walls_stays_at_rest():
    import exahype_amd as E
    N = 3
    s = E.ADERDGSolver(N)
    g = E.Grid(s, 4, 4, 4)
    g.set_initial_conditions(E.make_ic(2, q0=(1.2, 0.0, 0.0, 0.0, 2.0)))
    g.set_boundary((0, 0, 0), E.make_bc(E.BC_REFLECTING))
    u0 = g.download()
    g.compute_dt()
    g.step(4)
    g.sync()
    assert np.abs(g.download() - u0).max() <= 1e-12


def test_pdes_without_boundary_values_refuse_closed_axes():
    import exahype_amd as E
    s = E.ADERDGSolver(3, pde=E.PDE_ADVECTION, velocity=(1.0, 0.5, 0.25))
    g = E.Grid(s, 2, 2, 2)
    with pytest.raises(RuntimeError):
        g.set_boundary((1, 0, 1), E.make_bc(E.BC_OUTFLOW))
    g.set_boundary((1, 1, 1))  # all periodic stays allowed
    s2 = E.ADERDGSolver(3)
    with pytest.raises(RuntimeError):
        E.Grid(s2, 2, 2, 2).set_boundary((0, 1, 1), None)
