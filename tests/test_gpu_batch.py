"""The deferred per-cell entry points (exa_aderdg_batch_*) driven the way the
reference's mesh traversals call the solver: a prediction sweep
(performPredictionAndVolumeIntegral), a merging sweep (one Riemann solve per
face, FL/FR/QL/QR pointing into the neighbours' face arrays) and an update
sweep (surfaceIntegral + solutionUpdate), each ended by a flush.  Bar: after
each step the periodic grid equals the oracle's grid step to 1e-12
(relative, per variable), with the reference's heap layouts throughout."""
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
    return dict(xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0, amp=1e-2,
                nmodes=8, k=rng.integers(1, 4, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                phi=rng.uniform(0, 6.28, 8))


@pytest.mark.parametrize("N,dims,steps", [(3, (4, 3, 2), 2), (9, (2, 2, 1), 1)])
def test_deferred_sweeps_match_the_oracle_grid(N, dims, steps):
    import exahype_amd as E
    nx, ny, nz = dims
    nc = nx * ny * nz
    B = N + 1
    c = O.cfg(N)
    og = O.OracleGrid(c, nx, ny, nz)
    og.init(O.make_ic(0, **ic_kw(7)))
    dx = np.array([1.0 / nx, 1.0 / ny, 1.0 / nz])
    s = E.ADERDGSolver(N)
    luh = og.luh().copy().reshape(nc, B ** 3 * 5)
    lduh = np.zeros((nc, B ** 3 * 5))
    lQhbnd = np.zeros((nc, 6, B * B * 5))
    lFhbnd = np.zeros((nc, 6, B * B * 5))

    def nbr(cell, d):
        ix, iy, iz = cell % nx, (cell // nx) % ny, cell // (nx * ny)
        if d == 0:
            ix = (ix + 1) % nx
        elif d == 1:
            iy = (iy + 1) % ny
        else:
            iz = (iz + 1) % nz
        return (iz * ny + iy) * nx + ix

    dt = og.stable_dt()
    for step in range(steps):
        for k in range(nc):
            s.batch_prediction(lQhbnd[k], lFhbnd[k], lduh[k], luh[k], dx, dt)
        assert s.batch_pending() == (nc, 0, 0)
        s.batch_flush()
        for k in range(nc):
            for d in range(3):
                r = nbr(k, d)
                s.batch_riemann(lFhbnd[k, 2 * d + 1], lFhbnd[r, 2 * d], lQhbnd[k, 2 * d + 1], lQhbnd[r, 2 * d], dt, d)
        assert s.batch_pending() == (0, 3 * nc, 0)
        s.batch_flush()
        for k in range(nc):
            s.batch_update(luh[k], lduh[k], lFhbnd[k], dx, dt)
        s.batch_flush()
        assert s.batch_pending() == (0, 0, 0)
        dt = og.step(dt)
        assert rel_err(luh, og.luh()) <= 1e-12, (N, step, rel_err(luh, og.luh()))


def test_deferred_prediction_with_material_parameters():
    """npar > 0: the parameters ride along (constant in time) as in the
    per-cell drop-in; compare every output with the immediate calls."""
    import exahype_amd as E
    N, npar = 3, 2
    B = N + 1
    nd = 5 + npar
    s = E.ADERDGSolver(N, npar=npar)
    rng = np.random.default_rng(4)
    cells = []
    for k in range(5):
        rho = 1.0 + 0.1 * rng.random(B ** 3)
        m = 0.1 * rng.normal(size=(B ** 3, 3))
        p = 1.0 + 0.1 * rng.random(B ** 3)
        E_ = p / 0.4 + 0.5 * (m ** 2).sum(1) / rho
        u = np.column_stack([rho, m, E_, rng.random((B ** 3, npar))]).ravel()
        dx = np.array([0.02, 0.03, 0.025]) * (1 + 0.1 * k)
        dt = 1e-4 * (1 + k)
        outs = (np.zeros(6 * B * B * nd), np.zeros(6 * B * B * 5), np.zeros(B ** 3 * 5))
        s.batch_prediction(*outs, u, dx, dt)
        cells.append((u, dx, dt, outs))
    s.batch_flush()
    for u, dx, dt, (q, f, du) in cells:
        g = s.spaceTimePredictor(u, dx, dt)
        assert np.abs(q - g["lQhbnd"]).max() <= 1e-12 * np.abs(g["lQhbnd"]).max()
        assert np.abs(f - g["lFhbnd"]).max() <= 1e-12 * np.abs(g["lFhbnd"]).max()
        vi = s.volumeIntegral(g["lFhi"], dx)
        assert np.abs(du - vi).max() <= 1e-12 * np.abs(vi).max()
