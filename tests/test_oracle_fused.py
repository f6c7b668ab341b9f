"""Pins of the oracle's fused time stepping (oracle.FusedTimeStepping: the
time step data of the reference's fused algorithmic steps), against what
the recursion fixes independently of its code:

* a state at rest keeps one stable dt s; the dts then follow the closed form
  of the recursion dt_p <- 0.5 (dt_p + w s) from 0.5 (1 + w) s, the step
  after taking the previous optimistic dt: dt_c(1) = dt_c(2) =
  w s + 0.5 (1 - w) s, dt_c(k) = w s + 0.5^(k-1) (1 - w) s, approaching w s
  from above and never unstable (a wrong weight, factor or start fails it);
* on an evolving state every step after the first starts from a state whose
  stable dt is at least the dt it takes (the CFL bound the rerun rule
  protects), also after a forced unstable optimistic dt, which is counted
  once and replaced by w times the stable dt.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _vortex(seed=1):
    rng = np.random.default_rng(seed)
    return O.make_ic(kind=0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0,
                     amp=1e-2, nmodes=8, k=rng.integers(1, 3, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                     phi=rng.uniform(0, 6.28, 8))


@pytest.mark.parametrize("w", [1.0, 0.9, 0.5])
def test_constant_state_follows_the_closed_form(w):
    c = O.cfg(2)
    og = O.OracleGrid(c, 2, 2, 2)
    og.init(O.make_ic(kind=2, q0=(1.0, 0.0, 0.0, 0.0, 2.5)))
    s = og.stable_dt()
    fs = O.FusedTimeStepping(w)
    fs.start(s)
    for k in range(1, 7):
        expect = w * s + 0.5 ** max(1, k - 1) * (1.0 - w) * s
        assert abs(fs.dt_c - expect) <= 1e-12 * s, (k, fs.dt_c, expect)
        stable = og.step(fs.dt_c)
        assert abs(stable - s) <= 1e-13 * s
        fs.after_step(stable)
    # w = 1 keeps the optimistic dt at s itself, which the stable dt of the
    # (rounding-level) evolved state undercuts now and then: the reference
    # would rerun the predictor there
    assert fs.reruns == 0 or w == 1.0


@pytest.mark.parametrize("w", [0.99, 0.7])
def test_every_step_after_the_first_is_stable(w):
    c = O.cfg(3)
    og = O.OracleGrid(c, 3, 3, 3)
    og.init(_vortex())
    fs = O.FusedTimeStepping(w)
    fs.start(og.stable_dt())
    for k in range(8):
        if k == 4:  # an optimistic dt far above the stable one
            fs.dt_p = 3.0 * og.stable_dt()
        start_stable = og.stable_dt()
        if k > 0:
            assert fs.dt_c <= start_stable * (1 + 1e-14), (k, fs.dt_c, start_stable)
        stable = og.step(fs.dt_c)
        before = fs.reruns
        fs.after_step(stable)
        if k == 4:
            assert fs.reruns == before + 1
            assert fs.dt_c == w * stable and fs.dt_p == min(stable, w * stable)
    assert fs.reruns == 1
