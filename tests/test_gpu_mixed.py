"""Mixed-precision predictor: low-precision Picard sweeps (FP32 on the vector
ALU, or the spatial contractions on FP32 / BF16 MFMA) + FP64 correction
sweeps.

Bar (floating point, stated here): after 50 steps of the 3D Euler vortex
setup the L2 difference to the FP64 path is <= 1e-6 for every variable
(L2 over the domain with the Gauss-Legendre quadrature of each cell); the
per-cell predictor outputs differ from the FP64 oracle by at most 5e-6 of the
output's scale, for the FP32 forms with one FP64 correction sweep.  BF16
contraction operands leave a low-precision fixed point only ~1e-3 accurate;
with two FP64 sweeps its 50-step L2 is 1e-5..1e-4 (three sweeps: ~2e-6), so
BF16 misses the 1e-6 bar and is tested at 1e-3 (per-cell outputs likewise):
the measured accuracy of that path, not the acceptance bar of the headline.
Orders 8 and 9 (and the staged form at 5..7) take FP32 sweeps on FP32
staging arrays before the FP64 correction, to the same bars.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ic_kw(seed):
    rng = np.random.default_rng(seed)
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


def l2_per_variable(a, b, N, ncells, h3):
    from exahype_amd.tables import dg_math
    w = dg_math.gl_nodes_weights(N)[1]
    w3 = np.einsum("z,y,x->zyx", w, w, w).ravel()
    d = (np.asarray(a) - np.asarray(b)).reshape(ncells, (N + 1) ** 3, 5)
    return np.sqrt(np.einsum("n,cnv->v", w3, d ** 2) * h3)


# (precision, FP64 correction sweeps, 50-step L2 bar, per-cell bar)
PRECS = [(1, 1, 1e-6, 5e-6), (2, 1, 1e-6, 5e-6), (3, 2, 1e-3, 1e-3)]


def test_mfma_operand_layouts():
    """The lane layouts the matrix-core contraction assumes, against numpy."""
    import exahype_amd as E
    rng = np.random.default_rng(3)
    A = rng.normal(size=(16, 16)).astype(np.float32)
    Bm = rng.normal(size=(16, 16)).astype(np.float32)
    D = E.selftest_mfma(0, A, Bm)
    ref = A.astype(np.float64) @ Bm.astype(np.float64)
    assert np.abs(D - ref).max() <= 1e-5 * np.abs(ref).max()

    def bf16(x):
        u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32)
    D = E.selftest_mfma(1, A, Bm)
    ref = bf16(A).astype(np.float64) @ bf16(Bm).astype(np.float64)
    assert np.abs(D - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("prec,sweeps,l2bar,cellbar", PRECS)
@pytest.mark.parametrize("N", [3, 5])
def test_mixed_predictor_vs_oracle(N, prec, sweeps, l2bar, cellbar):
    import exahype_amd as E
    s = E.ADERDGSolver(N, precision=prec, fp64_sweeps=sweeps)
    c = O.cfg(N)
    rng = np.random.default_rng(1)
    ic = O.make_ic(0, **_ic_kw(3))
    for _ in range(3):
        centre = rng.uniform(0.3, 0.7, 3)
        luh = O.solution_adjustment(c, ic, centre, (1 / 27,) * 3)
        dx = np.full(3, 1 / 27)
        dt = 0.9 * O.stable_time_step_size(c, luh, dx)
        g = s.spaceTimePredictor(luh, dx, dt)
        o = O.space_time_predictor(c, luh, dx, dt)
        for k in ("lQhbnd", "lFhbnd", "lQhi"):
            a, b = g[k].reshape(-1, 5), o[k].reshape(-1, 5)
            err = np.abs(a - b).max() / np.abs(b).max()
            # The oracle's own Picard stopping rule (sum (dq)^2 < 1e-14) leaves
            # ~1e-7 absolute iteration error in its outputs, so two differently
            # iterated predictors agree only to a few 1e-7..1e-6 on flux-type
            # outputs; the acceptance bar of the mixed path is the solution L2
            # after many steps (test below).
            assert err <= cellbar, (N, k, err)


@pytest.mark.parametrize("prec,sweeps,l2bar,cellbar", PRECS)
@pytest.mark.parametrize("N,n", [(3, 12), (5, 9), (7, 5)])
def test_mixed_50_steps_l2_vs_fp64(N, n, prec, sweeps, l2bar, cellbar):
    import exahype_amd as E
    out = []
    for p, sw in ((E.PREC_FP64, 1), (prec, sweeps)):
        s = E.ADERDGSolver(N, precision=p, fp64_sweeps=sw)
        g = E.Grid(s, n, n, n)
        g.set_initial_conditions(E.make_ic(0, **_ic_kw(7)))
        g.compute_dt()
        g.step(50)
        t, _ = g.sync()
        out.append((g.download(), t))
    (a, ta), (b, tb) = out
    l2 = l2_per_variable(b, a, N, n ** 3, 1.0 / n ** 3)
    assert l2.max() <= l2bar, l2
    # dt follows the state, so the simulated times differ at the solution's level
    assert abs(ta - tb) <= l2bar * ta


@pytest.mark.parametrize("N,contraction", [(8, 0), (9, 0), (6, 4)])
def test_mixed_staged_predictor_vs_oracle(N, contraction):
    import exahype_amd as E
    s = E.ADERDGSolver(N, precision=E.PREC_FP32, fp64_sweeps=1, contraction=contraction)
    c = O.cfg(N)
    rng = np.random.default_rng(N)
    ic = O.make_ic(0, **_ic_kw(3))
    for _ in range(2):
        centre = rng.uniform(0.3, 0.7, 3)
        luh = O.solution_adjustment(c, ic, centre, (1 / 27,) * 3)
        dx = np.full(3, 1 / 27)
        dt = 0.9 * O.stable_time_step_size(c, luh, dx)
        g = s.spaceTimePredictor(luh, dx, dt)
        o = O.space_time_predictor(c, luh, dx, dt)
        assert g["iters"] >= 2  # at least one FP32 and one FP64 sweep
        for k in ("lQhbnd", "lFhbnd", "lQhi"):
            a, b = g[k].reshape(-1, 5), o[k].reshape(-1, 5)
            err = np.abs(a - b).max() / np.abs(b).max()
            assert err <= 5e-6, (N, k, err)


@pytest.mark.parametrize("N,n,contraction", [(8, 3, 0), (9, 3, 0), (7, 4, 4)])
def test_mixed_staged_50_steps_l2_vs_fp64(N, n, contraction):
    import exahype_amd as E
    print("mixed staged 50 steps N", N, flush=True)
    out = []
    for p in (E.PREC_FP64, E.PREC_FP32):
        s = E.ADERDGSolver(N, precision=p, fp64_sweeps=1, contraction=contraction)
        g = E.Grid(s, n, n, n)
        g.set_initial_conditions(E.make_ic(0, **_ic_kw(7)))
        g.compute_dt()
        g.step(50)
        t, _ = g.sync()
        out.append((g.download(), t))
    (a, ta), (b, tb) = out
    l2 = l2_per_variable(b, a, N, n ** 3, 1.0 / n ** 3)
    assert l2.max() <= 1e-6, l2
    assert abs(ta - tb) <= 1e-6 * ta
