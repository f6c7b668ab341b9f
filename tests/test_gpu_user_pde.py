"""PDEs compiled by a project into the kernels (include/exahype_amd_pde.h,
exa_pde_register) -- the example of tests/user_pde/user_pdes.hip:

* user linear advection against the library's compiled EXA_PDE_ADVECTION
  with the same velocity: initial state, dt and five grid steps bitwise;
* user Euler (reference-style flux / eigenvalues, reflecting-wall
  boundaryValues) against the oracle: per-cell predictor and face
  boundaryConditions, and a grid closed on every axis over three steps, to
  1e-12 with equal Picard counts;
* the boundary kind of a registered PDE and the built-in kinds do not mix.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "user_pde", "_build", "libuser_pdes.so")
VEL = (0.7, -0.4, 0.3)


def rel(a, b, nv=5):
    a, b = np.asarray(a).reshape(-1, nv), np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    return float((np.abs(a - b).max(axis=0) / np.maximum(scale, 1e-3 * scale.max())).max())


@pytest.fixture(scope="module")
def ids():
    import exahype_amd as E
    E.load_library()
    E.ADERDGSolver(1)  # exa_init on device 0
    if not os.path.exists(LIB):
        pytest.fail("example PDE library not built: " + LIB)
    lib = C.CDLL(LIB)
    out = (C.c_int * 2)()
    assert lib.user_pdes_register(out) == 0, E._lib.exa_last_error()
    assert E.pde_lookup("user_advection") == out[0] and E.pde_lookup("user_euler") == out[1]
    assert out[0] >= E.PDE_USER_FIRST and out[1] >= E.PDE_USER_FIRST
    return out[0], out[1]


def test_unknown_name_is_an_error(ids):
    import exahype_amd as E
    with pytest.raises(RuntimeError):
        E.pde_lookup("no_such_pde")


@pytest.mark.parametrize("N", [2, 5])
def test_user_advection_is_the_compiled_advection(ids, N):
    import exahype_amd as E
    dims = (4, 5, 3)
    su = E.ADERDGSolver(N, pde=ids[0])
    sb = E.ADERDGSolver(N, pde=E.PDE_ADVECTION, velocity=VEL)
    gu, gb = E.Grid(su, *dims), E.Grid(sb, *dims)
    gu.set_initial_conditions(None)
    gb.set_initial_conditions(E.make_ic(3, uinf=VEL, nmodes=1, k=[(1.0, 2.0, 1.0)], a=[0.1], phi=[0.4]))
    assert np.array_equal(gu.download(), gb.download())
    assert gu.compute_dt() == gb.compute_dt()
    gu.step(5)
    gb.step(5)
    tu, du = gu.sync()
    tb, db = gb.sync()
    assert tu == tb and du == db
    assert np.array_equal(gu.download(), gb.download())


def test_user_euler_cell_kernels_vs_oracle(ids):
    import exahype_amd as E
    N = 4
    s = E.ADERDGSolver(N, pde=ids[1])
    c = O.cfg(N)
    ic = O.make_ic(4, rho_inf=1.0, p_inf=1.0, amp=0.2, uinf=(0.7, -0.4, 0.3), nmodes=1, k=[(1, 2, 1)], a=[1.0],
                   phi=[0.4])
    luh = O.solution_adjustment(c, ic, (0.3, 0.6, 0.45), (0.1, 0.1, 0.1))
    dx = np.full(3, 0.1)
    dt = 0.9 * O.stable_time_step_size(c, luh, dx)
    assert abs(s.stableTimeStepSize(luh, dx) - dt / 0.9) <= 1e-14 * dt
    g = s.spaceTimePredictor(luh, dx, dt)
    o = O.space_time_predictor(c, luh, dx, dt)
    for k in ("lQhbnd", "lFhbnd", "lQhi", "lFhi"):
        assert rel(g[k], o[k]) <= 1e-12, k
    B2 = (N + 1) ** 2
    for face in range(6):
        d = face // 2
        QIn = g["lQhbnd"][face * B2 * 5:(face + 1) * B2 * 5]
        FIn = g["lFhbnd"][face * B2 * 5:(face + 1) * B2 * 5]
        Fg, Qg = s.boundaryConditions(E.make_bc(E.BC_USER), FIn, QIn, (0.3, 0.6, 0.45), (0.1, 0.1, 0.1), 0.0, dt,
                                      face, d)
        Fo, Qo = O.boundary_conditions(c, O.make_bc(O.BC_REFLECTING), FIn, QIn, (0.3, 0.6, 0.45), (0.1,) * 3, 0.0,
                                       dt, face, d)
        assert rel(Qg, Qo) <= 1e-15 and rel(Fg, Fo) <= 1e-14, face


def test_user_euler_closed_grid_vs_oracle(ids):
    import exahype_amd as E
    N, dims = 3, (4, 3, 5)
    c = O.cfg(N)
    og = O.OracleGrid(c, *dims)
    og.init(O.make_ic(4, rho_inf=1.0, p_inf=1.0, amp=0.2, uinf=(0.7, -0.4, 0.3), nmodes=1, k=[(1, 2, 1)], a=[1.0],
                      phi=[0.4]))
    og.set_boundary((0, 0, 0), O.make_bc(O.BC_REFLECTING))
    s = E.ADERDGSolver(N, pde=ids[1])
    g = E.Grid(s, *dims)
    g.set_initial_conditions(None)
    assert rel(g.download(), og.luh()) <= 1e-14
    g.upload(og.luh().copy())
    g.set_boundary((0, 0, 0), E.make_bc(E.BC_USER))
    g.use_graph(True)
    dt_o = og.stable_dt()
    assert abs(g.compute_dt() - dt_o) <= 1e-14 * dt_o
    for k in range(3):
        g.step(1)
        dt_o = og.step(dt_o)
        _, dt_g = g.sync()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o
        assert np.array_equal(g.picard_iterations(), og.iters()), k
        assert rel(g.download(), og.luh()) <= 1e-12, k


def test_user_boundary_kind_needs_a_user_pde(ids):
    import exahype_amd as E
    g = E.Grid(E.ADERDGSolver(3), 2, 2, 2)
    with pytest.raises(RuntimeError):
        g.set_boundary((0, 1, 1), E.make_bc(E.BC_USER))
    gu = E.Grid(E.ADERDGSolver(3, pde=ids[1]), 2, 2, 2)
    with pytest.raises(RuntimeError):
        gu.set_boundary((0, 1, 1), E.make_bc(E.BC_REFLECTING))
