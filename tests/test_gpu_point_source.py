"""GPU parity of the point sources of the linear scheme: pointSource
(kernels/aderdg/generic/c/3d/pointSource.cpph:28-145) through
exa_aderdg_point_source, the usePointSource arm of the linear predictor
(spaceTimePredictorLinear.cpph:78-102, 287-302) through
exa_aderdg_space_time_predictor_ps, and grid steps that evaluate the source
per node at the time of each step -- against the oracle, which
tests/test_oracle_point_source.py pins.  FP64, <= 1e-12 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O

from test_gpu_aderdg import TOL, rel_err

pytestmark = pytest.mark.gpu

VEL = (0.3, -0.7, 0.5)
SRC = dict(x0=(0.52, 0.47, 0.53), amp=(1.0, -2.0, 3.0, 0.5, 4.0), f0=5.0, t0=0.1)
PNPM = (1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015)


def exa():
    import exahype_amd
    return exahype_amd


def solver(pde, N, src=SRC):
    return exa().ADERDGSolver(N, pde=pde, velocity=VEL, linear=True, point_source=src)


@pytest.mark.parametrize("N", range(1, 10))
def test_point_source_vs_oracle(N):
    print("point source N", N, flush=True)
    s = solver(4, N)
    c = O.cfg(N, pde=4, adv=VEL, linear=1, point_source=SRC)
    for centre, dx, t in (((0.5, 0.5, 0.5), (0.1, 0.08, 0.12), 0.05), ((0.55, 0.45, 0.5), (0.1, 0.1, 0.1), 0.13)):
        g = s.pointSource(t, 0.01, centre, dx)
        o = O.point_source(c, t, 0.01, centre, dx)
        assert np.any(o)
        assert rel_err(g, o) <= TOL, (N, centre)
    # outside the cell
    assert not np.any(s.pointSource(0.05, 0.01, (0.2, 0.5, 0.5), (0.1, 0.1, 0.1)))


@pytest.mark.parametrize("pde", [4, 5])
@pytest.mark.parametrize("N", [1, 3, 5, 9])
def test_predictor_with_point_forces_vs_oracle(pde, N):
    print("point-source predictor pde", pde, "N", N, flush=True)
    s = solver(pde, N)
    c = O.cfg(N, pde=pde, adv=VEL, linear=1, point_source=SRC)
    B = N + 1
    rng = np.random.default_rng(N)
    luh = np.tile([1.0, 0.1, -0.2, 0.3, 2.5], B ** 3) + 0.1 * rng.standard_normal(B ** 3 * 5)
    dx = np.array([1 / 27, 1 / 30, 1 / 24])
    dt = 0.9 * PNPM[N] / sum(abs(VEL[d]) / dx[d] for d in range(3))
    psi = O.point_source(c, 0.08, dt, (0.5, 0.5, 0.5), dx * 3)
    assert np.any(psi)
    g = s.spaceTimePredictor(luh, dx, dt, pointForceSources=psi)
    o = O.space_time_predictor_linear(c, luh, dx, dt, psi=psi)
    for k in ("lQhi", "lQhbnd", "lFhbnd", "lFhi"):
        assert rel_err(g[k], o[k]) <= TOL, (pde, N, k, rel_err(g[k], o[k]))
    gq, oq = g["lQi"].reshape(B + 1, -1), o["lQi"].reshape(B + 1, -1)
    for i in range(B + 1):
        if np.any(oq[i]):
            assert rel_err(gq[i], oq[i]) <= TOL, (pde, N, "lQi", i)
    # without forces the arm is off
    g0 = s.spaceTimePredictor(luh, dx, dt)
    o0 = O.space_time_predictor_linear(c, luh, dx, dt)
    assert rel_err(g0["lQhi"], o0["lQhi"]) <= TOL and not np.allclose(g0["lQhi"], g["lQhi"])


@pytest.mark.parametrize("N,dims,steps", [(2, (4, 3, 5), 4), (5, (3, 2, 3), 3), (9, (2, 2, 1), 2)])
def test_grid_steps_with_point_source_vs_oracle(N, dims, steps):
    print("point-source grid N", N, dims, flush=True)
    E = exa()
    # the source sits in one cell of the grid and peaks during the run
    src = dict(SRC, x0=(0.6, 0.3, 0.7), t0=0.01, f0=20.0)
    s = solver(4, N, src)
    og = O.OracleGrid(O.cfg(N, pde=4, adv=VEL, linear=1, point_source=src), *dims)
    ic = O.make_ic(kind=3, nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1])
    og.init(ic)
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
        assert abs(t - og.g.t) <= 1e-14, (k, t, og.g.t)
    assert rel_err(g.download(), og.luh()) <= TOL
    # the source changed the solution against the same run without it
    og2 = O.OracleGrid(O.cfg(N, pde=4, adv=VEL, linear=1), *dims)
    og2.init(ic)
    dt2 = og2.stable_dt()
    for _ in range(steps):
        dt2 = og2.step(dt2)
    assert rel_err(og2.luh(), og.luh()) > 1e-6


def test_point_source_rejections():
    E = exa()
    with pytest.raises(RuntimeError, match="linear kernels"):
        E.ADERDGSolver(3, pde=E.PDE_ADVECTION_NCP, velocity=VEL, point_source=SRC)
    s = E.ADERDGSolver(3, pde=E.PDE_ADVECTION_NCP, velocity=VEL, linear=True)
    with pytest.raises(RuntimeError, match="no point source"):
        s.pointSource(0.0, 0.01, (0.5, 0.5, 0.5), (0.1, 0.1, 0.1))


def test_grid_source_on_a_shared_face_vs_oracle():
    """A source on the face between two cells (x = 0.5 on 4 cells): the
    device's inside test admits it in both cells, as the oracle's does."""
    E = exa()
    N, dims = 3, (4, 3, 3)
    src = dict(SRC, x0=(0.5, 0.3, 0.7), t0=0.01, f0=20.0)
    s = solver(4, N, src)
    og = O.OracleGrid(O.cfg(N, pde=4, adv=VEL, linear=1, point_source=src), *dims)
    og.init(O.make_ic(kind=3, nmodes=2, k=[(1, 0, 1), (0, 1, 1)], a=[0.2, 0.1], phi=[0.3, 1.1]))
    g = E.Grid(s, *dims)
    g.upload(og.luh().copy())
    dt_o = og.stable_dt()
    g.compute_dt()
    for _ in range(3):
        g.step(1)
        dt_o = og.step(dt_o)
        g.sync()
    assert rel_err(g.download(), og.luh()) <= TOL
