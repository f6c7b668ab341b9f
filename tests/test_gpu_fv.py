"""GPU parity of the finite-volume kernels through the C ABI vs the oracle.

Bar: <= 1e-12 relative (per variable, 1e-3 scale floor as in the ADER-DG
tests) for patch updates (interior and the first ghost layer the reference
also writes), admissible and stable dt, Rusanov fluxes, and whole steps of
periodic patch grids (the EulerFV demonstrator setup in 2D, MUSCL-Hancock in
3D).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b, nv=5):
    a = np.asarray(a).reshape(-1, nv)
    b = np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    den = np.maximum(np.maximum(scale, 1e-3 * scale.max()), 1e-300)
    return float((np.abs(a - b).max(axis=0) / den).max())


def euler_state(rng, n, amp=0.2):
    rho = 1.0 + amp * rng.random(n)
    u = amp * rng.normal(size=(n, 3))
    p = 1.0 + amp * rng.random(n)
    E = p / 0.4 + 0.5 * rho * (u ** 2).sum(1)
    return np.column_stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2], E]).ravel()


# dim, patch, scheme, oracle pde (MUSCL-Hancock up to the 15^3 patch limit)
CASES = [(2, 10, 0, 2), (3, 6, 0, 0), (3, 11, 1, 0), (3, 5, 1, 0), (3, 13, 1, 0)]


@pytest.mark.parametrize("dim,patch,scheme,opde", CASES)
def test_patch_update_and_dt_vs_oracle(dim, patch, scheme, opde):
    import exahype_amd as E
    s = E.FiniteVolumesSolver(dim, patch, scheme)
    c = O.cfg(1, pde=opde)
    rng = np.random.default_rng(patch + dim)
    luh = euler_state(rng, s.side ** dim)
    dx = np.array([0.05, 0.04, 0.03])[:3]
    dt = 2e-4
    g_new, g_adm = s.solutionUpdate(luh, luh, dx, dt)
    if scheme == 0:
        o_new, o_adm = O.godunov_update(c, dim, patch, luh, luh, dx, dt, 0.9)
    else:
        o_new, o_adm = O.musclhancock_update(c, patch, luh, luh, dx, dt, 0.99)
    assert rel_err(g_new, o_new) <= TOL
    assert abs(g_adm - o_adm) <= 1e-14 * o_adm
    g_dt = s.stableTimeStepSize(luh, dx)
    o_dt = O.fv_stable_dt(c, dim, patch, s.ghost, luh, dx, 0.9)
    assert abs(g_dt - o_dt) <= 1e-14 * o_dt


@pytest.mark.parametrize("dim,opde", [(2, 2), (3, 0)])
def test_fv_rusanov_vs_oracle(dim, opde):
    import exahype_amd as E
    s = E.FiniteVolumesSolver(dim, 4, 0)
    c = O.cfg(1, pde=opde)
    rng = np.random.default_rng(9)
    q = euler_state(rng, 2).reshape(2, 5)
    for d in range(dim):
        fL, fR, sm = s.riemannSolver(q[0], q[1], d)
        oL, oR, osm = O.fv_rusanov(c, dim, q[0], q[1], d)
        assert rel_err(fL, oL) <= TOL and rel_err(fR, oR) <= TOL
        assert abs(sm - osm) <= 1e-15 * osm


def _pulse_ic(mod):
    # the EulerFV configuration's initial state: rho = 1 + 0.5 exp(-|x-x0|^2/0.01), u = 0.1, p = 1
    return mod.make_ic(5, xc=(0.42, 0.57, 0.5), L=0.01, amp=0.5, rho_inf=1.0, p_inf=1.0, uinf=(0.1, 0.1, 0.1))


@pytest.mark.parametrize("dim,patch,scheme,opde,n,steps", [(2, 10, 0, 2, 6, 20), (3, 5, 1, 0, 3, 6),
                                                           (3, 4, 0, 0, 3, 6)])
def test_fv_grid_steps_vs_oracle(dim, patch, scheme, opde, n, steps):
    import exahype_amd as E
    s = E.FiniteVolumesSolver(dim, patch, scheme)
    c = O.cfg(1, pde=opde)
    nz = n if dim == 3 else 1
    og = O.OracleFVGrid(c, dim, scheme, patch, n, n, nz)
    og.init(_pulse_ic(O))
    g = E.FVGrid(s, n, n, nz)
    g.set_initial_conditions(_pulse_ic(E))
    ghost = s.ghost
    S = patch + 2 * ghost
    sl = (slice(None),) + (slice(ghost, ghost + patch),) * dim

    def interior(a):
        return a.reshape((og.npatch,) + (S,) * dim + (5,))[sl]
    assert rel_err(interior(g.download()), interior(og.luh())) <= 1e-14
    dt_g, dt_o = g.compute_dt(), og.stable_dt()
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o
    for _ in range(steps):
        dt_o, adm_o = og.step(dt_o)
    g.step(steps)
    _, dt_g, adm_g = g.sync()
    assert abs(dt_g - dt_o) <= 1e-12 * dt_o
    assert abs(adm_g - adm_o) <= 1e-12 * adm_o
    assert rel_err(interior(g.download()), interior(og.luh())) <= TOL
