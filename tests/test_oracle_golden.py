"""Pin the oracle against the reference's own golden vectors.

The setups restate E/tests/kernels/c/GenericEulerKernelTest3D.cpp (order 3,
5 variables, 1 material parameter, 3D Euler with gamma 1.4); the expected
arrays come from E/tests/kernels/testdata/generic_euler_testdata.cpp (Dim3),
extracted to tests/golden/generic_euler_3d.json.  The reference accepts
|a-b| <= 1e-10 (GenericEulerKernelTest.h:72); the bound used here is the
same absolute 1e-10 plus, where the golden data carries enough digits, a
relative bound.
"""
import numpy as np

from oracle import oracle as O

NVAR, NPAR, ORDER = 5, 1, 3
B = ORDER + 1
ND = NVAR + NPAR
EPS = 1e-10


def cfg():
    return O.cfg(ORDER, NVAR, NPAR, gamma=1.4, cfl=0.9)


def restride(a, nnodes, nv=NVAR, nd=ND):
    out = np.zeros(nnodes * nd)
    out.reshape(nnodes, nd)[:, :nv] = np.asarray(a).reshape(nnodes, nv)
    return out


def unstride(a, nnodes, nv=NVAR, nd=ND):
    return np.asarray(a).reshape(nnodes, nd)[:, :nv].ravel()


def test_pde_fluxes(golden):
    # GenericEulerKernelTest3D.cpp:134-160
    c = O.cfg(ORDER, 5, 0)
    F = O.flux(c, [1.0, 0.1, 0.2, 0.3, 3.5])
    for d, key in enumerate("fgh"):
        np.testing.assert_allclose(F[d], golden["testPDEFluxes/" + key], rtol=1e-15, atol=1e-15)


def test_volume_integral_nonlinear(golden):
    # GenericEulerKernelTest3D.cpp:196-229
    lFhi = np.zeros(1280)
    for i in range(0, 320, 5):
        lFhi[i + 1] = 1.0
        lFhi[320 + i + 2] = 1.0
        lFhi[640 + i + 3] = 1.0
    lduh = O.volume_integral(cfg(), lFhi, [0.05, 0.05, 0.05])
    exp = golden["testVolumeIntegralNonlinear/lduh"]
    assert np.max(np.abs(lduh - exp)) <= EPS
    assert np.max(np.abs(lduh - exp)) <= 1e-12 * np.max(np.abs(exp))


def test_surface_integral_nonlinear(golden):
    # GenericEulerKernelTest3D.cpp:275-317
    lFhbnd = np.zeros(480)
    for i in range(0, 80, 5):
        lFhbnd[0 + i + 1] = 1.0
        lFhbnd[80 + i + 1] = 1.0
        lFhbnd[160 + i + 2] = 1.0
        lFhbnd[240 + i + 2] = 1.0
        lFhbnd[320 + i + 3] = 1.0
        lFhbnd[400 + i + 3] = 1.0
    lduh = np.arange(320) / 10.0
    out = O.surface_integral(cfg(), lduh, lFhbnd, [0.1, 0.1, 0.1])
    exp = golden["testSurfaceIntegralNonlinear/lduh"]
    assert np.max(np.abs(out - exp)) <= EPS


def test_riemann_solver_nonlinear(golden):
    # GenericEulerKernelTest3D.cpp:410-486 (random face states, normal y)
    FL = np.arange(80, dtype=float) + 1
    FR = np.arange(80, dtype=float) - 1
    QL = restride(golden["testRiemannSolver/QL"], 16)
    QR = restride(golden["testRiemannSolver/QR"], 16)
    FLo, FRo = O.riemann_solver(cfg(), FL, FR, QL, QR, 1.40831757919882352703e-03, 1)
    eFL = golden["testRiemannSolverNonlinear/FL"]
    eFR = golden["testRiemannSolverNonlinear/FR"]
    assert np.max(np.abs(FLo - eFL)) <= EPS
    assert np.max(np.abs(FRo - eFR)) <= EPS
    # the golden data carries ~20 digits: demand a tight relative agreement
    assert np.max(np.abs(FLo - eFL)) <= 1e-13 * np.max(np.abs(eFL))


def test_riemann_solver_direction_matters(golden):
    """A mistake in the normal index would change the wave speed: direction 0
    and 2 must NOT reproduce the direction-1 golden data."""
    FL = np.arange(80, dtype=float) + 1
    FR = np.arange(80, dtype=float) - 1
    QL = restride(golden["testRiemannSolver/QL"], 16)
    QR = restride(golden["testRiemannSolver/QR"], 16)
    for d in (0, 2):
        FLo, _ = O.riemann_solver(cfg(), FL, FR, QL, QR, 1e-3, d)
        assert np.max(np.abs(FLo - golden["testRiemannSolverNonlinear/FL"])) > 1e-3


def test_solution_update(golden):
    # GenericEulerKernelTest3D.cpp:495-532
    luh = np.zeros(ND * B ** 3)
    luh[0::ND] = 1.0
    luh[4::ND] = 2.5
    lduh = np.arange(NVAR * B ** 3, dtype=float)
    out = O.solution_update(cfg(), luh, lduh, 1.40831757919882352703e-03)
    got = unstride(out, B ** 3)
    exp = golden["testSolutionUpdate/luh"]
    # the reference checks the first basisSize^2 nodes; the fixture holds all
    assert np.max(np.abs(got[:80] - exp[:80])) <= EPS
    assert np.max(np.abs(got - exp)) <= EPS


def test_space_time_predictor_nonlinear(golden):
    # GenericEulerKernelTest3D.cpp:618-773 (constant state input)
    luh = restride(golden["testSpaceTimePredictorNonlinear/luh"], B ** 3)
    out = O.space_time_predictor(cfg(), luh, [0.05, 0.05, 0.05], 1.083937460199773e-3)
    g = "testSpaceTimePredictorNonlinear/"
    assert np.max(np.abs(unstride(out["lQi"], B ** 4) - golden[g + "lQi"])) <= EPS
    lFi = out["lFi"].reshape(B, B, B, B, 4, NVAR)[:, :, :, :, :3, :].ravel()
    assert np.max(np.abs(lFi - golden[g + "lFi"])) <= EPS
    assert np.max(np.abs(unstride(out["lQhi"], B ** 3) - golden[g + "lQhi"])) <= EPS
    assert np.max(np.abs(out["lFhi"][:960] - golden[g + "lFhi"])) <= EPS
    assert np.max(np.abs(unstride(out["lQhbnd"], 6 * B * B) - golden[g + "lQhbnd"])) <= EPS
    assert np.max(np.abs(out["lFhbnd"] - golden[g + "lFhbnd"])) <= EPS
