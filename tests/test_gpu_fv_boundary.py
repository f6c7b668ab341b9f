"""FV boundary hooks and closed FV grids on the GPU against the oracle
(oracle/fv.c): boundaryLayerExtraction, ghostLayerFillingAtBoundary and
ghostLayerFilling bitwise for every face; boundaryConditions for every kind
(exact data to 1e-14); grids with closed axes (the EulerFV demonstrator in
2D, Godunov and MUSCL-Hancock in 3D) over several steps to 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b, nv=5):
    a, b = np.asarray(a).reshape(-1, nv), np.asarray(b).reshape(-1, nv)
    scale = np.abs(b).max(axis=0)
    den = np.maximum(np.maximum(scale, 1e-3 * scale.max()), 1e-300)
    return float((np.abs(a - b).max(axis=0) / den).max())


# dim, patch, scheme, oracle pde
CASES = [(2, 6, 0, 2), (3, 4, 0, 0), (3, 3, 1, 0)]


def _exact(mod):
    return mod.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 0.5, 0.2), rho_inf=1.0, p_inf=1.0,
                       amp=1e-2, nmodes=1, k=[(1.0, 2.0, 1.0)], a=[1.0], phi=[0.3])


@pytest.mark.parametrize("dim,patch,scheme,opde", CASES)
def test_patch_hooks_vs_oracle(dim, patch, scheme, opde):
    import exahype_amd as E
    s = E.FiniteVolumesSolver(dim, patch, scheme)
    c = O.cfg(1, pde=opde)
    G = s.ghost
    rng = np.random.default_rng(dim + patch)
    me = rng.standard_normal(s.patch_size)
    nb = rng.standard_normal(s.patch_size)
    centre, size = (0.31, 0.62, 0.47)[:dim] + (0.0,) * (3 - dim), (0.1, 0.13, 0.07)
    for d in range(dim):
        for sgn in (-1, 1):
            pos = [0, 0, 0]
            pos[d] = sgn
            print("hooks", dim, patch, scheme, pos, flush=True)
            assert np.array_equal(s.ghostLayerFilling(me, nb, pos),
                                  O.fv_ghost_layer_filling(c, dim, patch, G, me, nb, pos))
            bnd = s.boundaryLayerExtraction(me, pos)
            assert np.array_equal(bnd, O.fv_boundary_layer_extraction(c, dim, patch, G, me, pos))
            assert np.array_equal(s.ghostLayerFillingAtBoundary(me, bnd, pos),
                                  O.fv_ghost_layer_filling_at_boundary(c, dim, patch, G, me, bnd, pos))
            face = 2 * d + (sgn > 0)
            for kind in (O.BC_OUTFLOW, O.BC_REFLECTING, O.BC_EXACT):
                ebc = E.make_bc(kind, _exact(E) if kind == O.BC_EXACT else None)
                obc = O.make_bc(kind, _exact(O) if kind == O.BC_EXACT else None)
                got = s.boundaryConditions(ebc, bnd, centre, size, 0.2, 1e-3, face, d)
                exp = O.fv_boundary_conditions(c, obc, dim, patch, G, bnd, centre, size, 0.2, 1e-3, face, d)
                if kind == O.BC_EXACT:
                    assert rel_err(got, exp) <= 1e-14, (kind, rel_err(got, exp))
                else:
                    assert np.array_equal(got, exp), kind


def _pulse_ic(mod):
    return mod.make_ic(5, xc=(0.42, 0.57, 0.5), L=0.01, amp=0.5, rho_inf=1.0, p_inf=1.0, uinf=(0.3, -0.2, 0.1))


@pytest.mark.parametrize("dim,patch,scheme,opde,periodic,kind", [
    (2, 10, 0, 2, (0, 0, 1), O.BC_OUTFLOW),
    (2, 10, 0, 2, (1, 0, 1), O.BC_REFLECTING),
    (3, 4, 0, 0, (0, 0, 0), O.BC_REFLECTING),
    (3, 5, 1, 0, (0, 0, 0), O.BC_OUTFLOW),
    (3, 5, 1, 0, (0, 1, 0), O.BC_EXACT),
])
def test_closed_fv_grid_vs_oracle(dim, patch, scheme, opde, periodic, kind):
    import exahype_amd as E
    print("closed fv grid", dim, patch, scheme, periodic, kind, flush=True)
    s = E.FiniteVolumesSolver(dim, patch, scheme)
    c = O.cfg(1, pde=opde)
    n, steps = (5, 12) if dim == 2 else (3, 5)
    nz = n if dim == 3 else 1
    off, wid = (0.1, -0.2, 0.05), (1.1, 0.9, 1.2)
    og = O.OracleFVGrid(c, dim, scheme, patch, n, n + 1, nz)
    for d in range(3):
        og.g.offset[d], og.g.width[d] = off[d], wid[d]
    og.init(_pulse_ic(O))
    og.set_boundary(periodic, O.make_bc(kind, _exact(O) if kind == O.BC_EXACT else None))
    g = E.FVGrid(s, n, n + 1, nz, offset=off, width=wid)
    g.set_initial_conditions(_pulse_ic(E))
    g.set_boundary(periodic, E.make_bc(kind, _exact(E) if kind == O.BC_EXACT else None))
    S = patch + 2 * s.ghost
    sl = (slice(None),) + (slice(s.ghost, s.ghost + patch),) * dim

    def interior(a):
        return a.reshape((og.npatch,) + (S,) * dim + (5,))[sl]
    dt_g, dt_o = g.compute_dt(), og.stable_dt()
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o
    for _ in range(steps):
        dt_o, adm_o = og.step(dt_o)
    g.step(steps)
    _, dt_g, adm_g = g.sync()
    assert abs(dt_g - dt_o) <= TOL * dt_o
    assert abs(adm_g - adm_o) <= TOL * adm_o
    assert rel_err(interior(g.download()), interior(og.luh())) <= TOL


def test_closing_needs_boundary_values():
    import exahype_amd as E
    s = E.FiniteVolumesSolver(2, 4)
    g = E.FVGrid(s, 2, 2)
    with pytest.raises(RuntimeError):
        g.set_boundary((0, 1, 1))
    g.set_boundary((1, 1, 0))  # z is not an axis of a 2D grid
