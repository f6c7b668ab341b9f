"""A blown-up state is reported, not stepped on.

A cell whose state is not finite or has a negative pressure has no stable
time step (the reference asserts on it, E/solvers/ADERDGSolver.cpp:1948-2001).
The device dt reductions map such a candidate to 0 so that it wins the global
minimum, and the dt / sync calls return EXA_ERR_NUMERICS.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def exa():
    import exahype_amd
    return exahype_amd


def _vortex(N, n):
    og = O.OracleGrid(O.cfg(N), n, n, n)
    og.init(O.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=0.0,
                      nmodes=0))
    return og.luh().copy()


@pytest.mark.parametrize("bad", ["negative_pressure", "nan"])
def test_grid_reports_blown_up_cell(bad):
    E = exa()
    N, n = 3, 4
    luh = _vortex(N, n)
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.upload(luh)
    dt = g.compute_dt()
    assert dt > 0.0
    g.step(1)
    g.sync()
    bad_luh = luh.copy().reshape(n ** 3, (N + 1) ** 3, 5)
    if bad == "nan":
        bad_luh[37, 5, 2] = np.nan
    else:
        bad_luh[37, 5, 4] = -1.0  # energy below the kinetic energy: p < 0
    g.upload(bad_luh.ravel())
    with pytest.raises(E.NumericsError):
        g.compute_dt()
    # a healthy state again: a dt and steps as before
    g.upload(luh)
    assert abs(g.compute_dt() - dt) <= 1e-15 * dt


def test_grid_step_into_blow_up_is_reported_by_sync():
    E = exa()
    N, n = 3, 4
    luh = _vortex(N, n).reshape(n ** 3, (N + 1) ** 3, 5)
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.upload(luh.ravel())
    g.compute_dt()
    luh[11, :, 4] = -1.0  # the whole cell at negative pressure: the update keeps it there
    g.upload(luh.ravel())
    g.step(1)
    with pytest.raises(E.NumericsError):
        g.sync()


def test_per_cell_stable_dt_reports_negative_pressure():
    E = exa()
    N = 3
    s = E.ADERDGSolver(N)
    luh = _vortex(N, 1).reshape((N + 1) ** 3, 5)
    assert s.stableTimeStepSize(luh.ravel(), (0.1, 0.1, 0.1)) > 0.0
    luh[3, 4] = -2.0
    with pytest.raises(E.NumericsError):
        s.stableTimeStepSize(luh.ravel(), (0.1, 0.1, 0.1))


def test_fv_grid_reports_negative_pressure():
    E = exa()
    fs = E.FiniteVolumesSolver(3, 5, scheme=E.FV_MUSCLHANCOCK)
    g = E.FVGrid(fs, 2, 2, 2)
    g.set_initial_conditions(E.make_ic(2, q0=(1.0, 0.1, 0.0, 0.0, 2.5)))
    assert g.compute_dt() > 0.0
    u = g.download().reshape(8, -1)
    u[5, :] = -1.0
    g.upload(u.ravel())
    with pytest.raises(E.NumericsError):
        g.compute_dt()
