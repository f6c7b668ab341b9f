"""A project built from a specification file drives the same kernels as the
objects built directly: the grids from exahype_amd.spec.create step to
bitwise-identical states."""
import numpy as np
import pytest

from test_spec import DG_SPEC, FV_SPEC

pytestmark = pytest.mark.gpu


def test_fv_project_from_spec():
    import exahype_amd as E
    from exahype_amd import spec as S
    ic = E.make_ic(5, xc=(0.4, 0.6, 0.5), L=0.01, amp=0.5, rho_inf=1.0, p_inf=1.0, uinf=(0.1, 0.1, 0.1))
    s1, g1 = S.create(S.Spec(FV_SPEC))
    s2 = E.FiniteVolumesSolver(2, 10)
    g2 = E.FVGrid(s2, 81, 81, 1)
    states = []
    for g in (g1, g2):
        g.set_initial_conditions(ic)
        g.compute_dt()
        g.step(3)
        g.sync()
        states.append(g.download())
    assert np.array_equal(states[0], states[1])


def test_aderdg_project_from_spec():
    import exahype_amd as E
    from exahype_amd import spec as S
    sp = S.Spec(DG_SPEC)
    ic = E.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0)
    s1, g1 = S.create(sp, 0)
    s2 = E.ADERDGSolver(5)
    g2 = E.Grid(s2, 27, 27, 27)
    states = []
    for g in (g1, g2):
        g.set_initial_conditions(ic)
        g.compute_dt()
        g.step(2)
        g.sync()
        states.append(g.download())
    assert np.array_equal(states[0], states[1])
    # the linear solver of the same project, with the advection PDE in non-conservative form
    s3, g3 = S.create(sp, 1, pde=E.PDE_ADVECTION_NCP, velocity=(0.3, -0.7, 0.5))
    assert s3.linear and g3.ncells == 9 ** 3
    g3.set_initial_conditions(E.make_ic(3, nmodes=1, k=[(1, 0, 0)], a=[0.1], phi=[0.0]))
    g3.compute_dt()
    g3.step(2)
    g3.sync()
    assert np.all(np.isfinite(g3.download()))


def test_fused_project_from_spec():
    """fuse-algorithmic-steps = on gives the grid fused stepping with the
    spec's factor: bitwise the directly built fused grid, with a dt sequence
    other than the step-by-step stable one."""
    import exahype_amd as E
    from exahype_amd import spec as S
    from test_spec import _with_optimisation
    sp = S.Spec(_with_optimisation("on", "0.8"))
    ic = E.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0)
    s1, g1 = S.create(sp, 0)
    s2 = E.ADERDGSolver(5)
    g2 = E.Grid(s2, 27, 27, 27)
    g2.set_fused_time_stepping(True, 0.8)
    s3 = E.ADERDGSolver(5)
    g3 = E.Grid(s3, 27, 27, 27)
    out = []
    for g in (g1, g2, g3):
        g.set_initial_conditions(ic)
        g.compute_dt()
        g.step(2)
        out.append((g.sync(), g.download()))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert out[0][0][0] < out[2][0][0]  # fused: 0.5 (1 + 0.8) of the stable dt per step here


def test_limiting_project_from_spec():
    """A Limiting-ADER-DG solver gives the adaptive limiting grid with the
    spec's mesh and DMP parameters: bitwise the directly built grid."""
    import exahype_amd as E
    from exahype_amd import spec as S
    from test_spec import LIMITING_SPEC
    sp = S.Spec(LIMITING_SPEC)
    refine = ((0.5, 0.5, 0.5), 0.25, (0.05, 0.05))
    s1, g1 = S.create(sp, 0, refine=refine)
    s2 = E.ADERDGSolver(3)
    g2 = E.AMRGrid(s2, E.amr_config(9, 3, (0.5, 0.5, 0.5), 0.25, [0.05, 0.05], relax=1e-3, scaling=1e-2,
                                    steps_till_cured=0))
    assert g1.ncells == g2.ncells and np.array_equal(g1.cells(), g2.cells())
    ic = E.make_ic(kind=1, xc=(0.5, 0.5, 0.5), r0=0.25, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1)
    out = []
    for g in (g1, g2):
        g.set_initial_conditions(ic)
        g.compute_dt()
        g.step(2)
        out.append((g.sync(), g.download(), g.limiter_status()))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
    with pytest.raises(S.SpecError):
        S.create(S.Spec(LIMITING_SPEC.replace("offset          = 0.0, 0.0, 0.0", "offset          = 0.1, 0.0, 0.0")))
