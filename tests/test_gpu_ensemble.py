"""Ensembles of independent grids through the C ABI.

Every member of an ensemble (orders 2..7, uneven cell counts, gamma per
member) must reproduce a stand-alone device grid of the same configuration
(itself pinned to the oracle in test_gpu_aderdg.py) to 1e-13 relative -- the
shared launches compute the same per-cell arithmetic -- with identical
Picard iteration totals; one small member is also checked directly against
the oracle grid to the plain method's 1e-12.  Graph replay and eager
stepping agree exactly.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(a, b, nv=5):
    a = np.asarray(a).reshape(-1, nv)
    b = np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    den = np.maximum(np.maximum(scale, 1e-3 * scale.max()), 1e-300)
    return float((np.abs(a - b).max(axis=0) / den).max())


def ic_kw(seed):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=rng.uniform(1, 5), uinf=tuple(rng.uniform(-1, 1, 3)), rho_inf=1.0,
                p_inf=1.0, amp=1e-2, nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float),
                a=rng.uniform(-1, 1, 8), phi=rng.uniform(0, 6.28, 8))


MEMBERS = [dict(order=3, cells=(4, 5, 3), gamma=1.4), dict(order=5, cells=(3, 3, 3), gamma=1.3),
           dict(order=3, cells=(6, 2, 4), gamma=1.67), dict(order=2, cells=(5, 4, 4), gamma=1.5),
           dict(order=7, cells=(2, 2, 2), gamma=1.35), dict(order=5, cells=(2, 3, 2), gamma=1.6)]


@pytest.mark.parametrize("prec", [0, 2])
def test_members_match_standalone_grids(prec):
    import exahype_amd as E
    steps = 4
    ens = E.Ensemble(MEMBERS, precision=prec)
    ens.set_initial_conditions([E.make_ic(0, **ic_kw(i)) for i in range(len(MEMBERS))])
    dts0 = ens.compute_dt()
    ens.step(steps)
    t_e, dt_e = ens.sync()
    it_e = ens.total_picard_iterations()
    it_g = 0
    for i, m in enumerate(MEMBERS):
        s = E.ADERDGSolver(m["order"], gamma=m["gamma"], precision=prec)
        g = E.Grid(s, *m["cells"])
        g.set_initial_conditions(E.make_ic(0, **ic_kw(i)))
        dt0 = g.compute_dt()
        assert abs(dt0 - dts0[i]) <= 1e-15 * dt0
        g.step(steps)
        t, dt = g.sync()
        it_g += g.total_picard_iterations()
        assert abs(t - t_e[i]) <= 1e-14 * t and abs(dt - dt_e[i]) <= 1e-14 * dt
        assert rel_err(ens.download(i), g.download()) <= 1e-13, i
    assert it_e == it_g


def test_member_vs_oracle_and_graph_replay():
    import exahype_amd as E
    m = dict(order=3, cells=(4, 3, 5), gamma=1.55)
    ens = E.Ensemble([m, dict(order=4, cells=(2, 2, 2), gamma=1.4)])
    ens.use_graph(False)
    ic = ic_kw(21)
    ens.set_initial_conditions([E.make_ic(0, **ic), E.make_ic(0, **ic_kw(2))])
    ens.compute_dt()
    c = O.cfg(3, gamma=1.55)
    og = O.OracleGrid(c, *m["cells"])
    og.init(O.make_ic(0, **ic))
    dt = og.stable_dt()
    for _ in range(3):
        dt = og.step(dt)
    ens.step(3)
    assert rel_err(ens.download(0), og.luh()) <= 1e-12
    # the same member again, stepped by graph replays
    ens2 = E.Ensemble([m, dict(order=4, cells=(2, 2, 2), gamma=1.4)])
    ens2.set_initial_conditions([E.make_ic(0, **ic), E.make_ic(0, **ic_kw(2))])
    ens2.compute_dt()
    ens2.step(3)
    assert np.array_equal(ens2.download(0), ens.download(0))
    assert np.array_equal(ens2.download(1), ens.download(1))


def test_too_large_is_refused():
    import exahype_amd as E
    with pytest.raises(RuntimeError):
        E.Ensemble([dict(order=7, cells=(400, 400, 400))])
