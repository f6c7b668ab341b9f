"""Fused time stepping of the device grid (exa_grid_set_fused_time_stepping)
against the oracle grid stepped with the oracle's restatement of the
reference's fused time step data (oracle.FusedTimeStepping): the same dt
sequence (to 1e-12), states to 1e-12 with equal Picard counts, and the same
count of unstable optimistic steps, eager and graphed, including an
optimistic dt forced far above the stable one from the host; switching
fused stepping off returns to the stable dt of each step.
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


def _vortex_kw(seed=2):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 3, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("N,w", [(3, 0.9), (5, 0.99)])
def test_fused_stepping_vs_oracle(N, w, graph):
    import exahype_amd as E
    print("fused stepping N", N, "w", w, "graph", graph, flush=True)
    n = 4
    c = O.cfg(N)
    og = O.OracleGrid(c, n, n, n)
    og.init(O.make_ic(0, **_vortex_kw()))
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.upload(og.luh().copy())
    g.use_graph(graph)
    g.set_fused_time_stepping(True, w)
    fs = O.FusedTimeStepping(w)
    fs.start(og.stable_dt())
    dt_g = g.compute_dt()
    assert abs(dt_g - fs.dt_c) <= 1e-14 * fs.dt_c
    t_o = 0.0
    for k in range(6):
        if k == 3:  # an optimistic dt far above the stable one, set from the host
            big = 1.5 * fs.dt_c
            g.set_dt(big)
            fs.dt_c = fs.dt_p = big
        g.step(1)
        t_o += fs.dt_c
        stable = og.step(fs.dt_c)
        fs.after_step(stable)
        t_g, dt_g = g.sync()
        assert abs(dt_g - fs.dt_c) <= TOL * fs.dt_c, (k, dt_g, fs.dt_c)
        assert abs(t_g - t_o) <= TOL * t_o, (k, t_g, t_o)
        assert np.array_equal(g.picard_iterations(), og.iters()), k
        assert rel(g.download(), og.luh()) <= TOL, (k, rel(g.download(), og.luh()))
    assert g.fused_reruns() == fs.reruns >= 1


def test_fused_stepping_off_returns_to_stable_steps():
    import exahype_amd as E
    N, n = 3, 3
    c = O.cfg(N)
    og = O.OracleGrid(c, n, n, n)
    og.init(O.make_ic(0, **_vortex_kw(5)))
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.upload(og.luh().copy())
    g.set_fused_time_stepping(True, 0.8)
    g.compute_dt()
    g.step(2)
    g.sync()
    g.set_fused_time_stepping(False)
    og.set_luh(g.download())
    dt_o = og.stable_dt()
    assert abs(g.compute_dt() - dt_o) <= 1e-14 * dt_o
    for k in range(2):
        g.step(1)
        dt_o = og.step(dt_o)
        _, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= TOL * dt_o
        assert rel(g.download(), og.luh()) <= TOL


def test_fused_weight_is_checked():
    import exahype_amd as E
    g = E.Grid(E.ADERDGSolver(2), 2, 2, 2)
    for w in (0.0, -0.5, 1.5):
        with pytest.raises(RuntimeError):
            g.set_fused_time_stepping(True, w)
