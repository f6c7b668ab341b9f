"""Troubled-cell detection and batched FV projection on a device grid vs the
oracle (E/solvers/LimitingADERDGSolver.cpp:838-1048 with
kernels/limiter/generic/c/Limiter.cpp).

Setup: the Sod sphere (rho, p = 1, 1 inside r < 0.25, 0.125, 0.1 outside) on
an 8^3 periodic grid at N = 3 (also 6^3 at N = 7 and 4^3 at N = 9, the
staged predictor).  The snapshot extrema of the initial state are
merged over each cell and its six face neighbours, one ADER-DG step is taken,
and the relaxed DMP plus positivity flags cells.  Bar: the troubled sets are
identical to the oracle's (restated with orc_cell_min_max / orc_dmp on the
oracle grid's step), and each troubled cell's subcell projection matches
orc_project_on_fv to 1e-12 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RELAX, SCALING = 1e-4, 1e-3


def sod(mod):
    return mod.make_ic(1, xc=(0.5, 0.5, 0.5), r0=0.25, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1)


def neighbours(c, n):
    ix, iy, iz = c % n, (c // n) % n, c // (n * n)
    out = []
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        out.append(((iz + dz) % n * n + (iy + dy) % n) * n + (ix + dx) % n)
    return out


@pytest.mark.parametrize("N,n", [(3, 8), (7, 6), (9, 4)])
def test_troubled_cells_and_projections_vs_oracle(N, n):
    import exahype_amd as E
    B = N + 1
    per = B ** 3 * 5
    s = E.ADERDGSolver(N)
    g = E.Grid(s, n, n, n)
    g.set_initial_conditions(sod(E))
    g.compute_dt()
    g.limiter_snapshot(5)
    g.step(1)
    ntr, flags = g.limiter_check(RELAX, SCALING)

    c = O.cfg(N)
    og = O.OracleGrid(c, n, n, n)
    og.init(sod(O))
    prev = og.luh().copy().reshape(n ** 3, per)
    mm = [O.cell_min_max(N, prev[k]) for k in range(n ** 3)]
    dt = og.stable_dt()
    og.step(dt)
    new = og.luh().copy().reshape(n ** 3, per)
    ref = np.zeros(n ** 3, np.int32)
    for k in range(n ** 3):
        nb = neighbours(k, n)
        bmin = np.array([np.minimum(mm[k][0], mm[j][0]) for j in nb])
        bmax = np.array([np.maximum(mm[k][1], mm[j][1]) for j in nb])
        ok, _, _ = O.dmp(N, new[k], bmin.ravel(), bmax.ravel(), RELAX, SCALING)
        lo, _ = O.cell_min_max(N, new[k])
        ok = ok and lo[0] > 0 and lo[4] > 0
        ref[k] = 0 if ok else 1
    assert ref.sum() > 0 and (N != 3 or ref.sum() < n ** 3), "the setup must flag some cells (not all at N=3)"
    assert np.array_equal(flags, ref)
    assert ntr == ref.sum()

    ghost = 2
    S = 2 * N + 1 + 2 * ghost
    cells, patches = g.limiter_project(ghost, ntr)
    assert sorted(cells.tolist()) == sorted(np.nonzero(ref)[0].tolist())
    L = 2 * N + 1
    for k, cell in enumerate(cells):
        o = O.project_on_fv(N, new[cell], ghost).reshape(S, S, S, 5)[ghost:ghost + L, ghost:ghost + L, ghost:ghost + L]
        p = patches[k].reshape(5, S, S, S)[:, ghost:ghost + L, ghost:ghost + L, ghost:ghost + L].transpose(1, 2, 3, 0)
        scale = np.abs(o).reshape(-1, 5).max(axis=0)
        err = (np.abs(p - o).reshape(-1, 5).max(axis=0) / scale).max()
        assert err <= 1e-12, (cell, err)


def test_check_requires_snapshot():
    import exahype_amd as E
    s = E.ADERDGSolver(3)
    g = E.Grid(s, 2, 2, 2)
    with pytest.raises(RuntimeError):
        g.limiter_check()
