"""CPU pins of the oracle's limiting step on a static adaptive mesh
(oracle/amr_limiting.py) and of the library's mesh builder.

* Mesh: the leaves tile the unit cube, no leaf has a face neighbour two
  levels finer, every face is accounted for exactly once; the library's
  host-side builder (exa_amr_mesh_cells, no device needed) gives the same
  leaves as the oracle's independent restatement.
* One level, no troubled cell: the adaptive step is bitwise the oracle's
  plain periodic grid step (oracle/aderdg.c orc_grid_step), an independent
  driver of the same per-cell kernels.
* Constant states are fixed points across two and three levels (face
  prolongation / restriction and Riemann problems at hanging faces).
* Status spreading around one troubled cell on a uniform mesh gives the
  rings max(0, 5 - face distance) of the reference's merge rule.
* A troubled cell's new solution is exactly the MUSCL-Hancock update of its
  patch projected back (the update switch), recomputed here by hand.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from oracle import amr_limiting as A


def _vortex_ic():
    rng = np.random.default_rng(5)
    return O.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(1.0, 1.0, 0.0), rho_inf=1.0, p_inf=1.0,
                     amp=1e-2, nmodes=8, k=rng.integers(1, 3, size=(8, 3)).astype(float), a=rng.uniform(-1, 1, 8),
                     phi=rng.uniform(0, 6.28, 8))


MESHES = [(3, 2, (0.5, 0.5, 0.5), 0.0, [0.01]), (3, 3, (0.5, 0.5, 0.5), 0.0, [0.01, 0.01]),
          (4, 3, (0.41, 0.52, 0.47), 0.2, [0.1, 0.03]), (5, 2, (0.3, 0.6, 0.5), 0.25, [0.05])]


@pytest.mark.parametrize("args", MESHES)
def test_mesh_tiles_and_is_balanced(args):
    m = A.Mesh(*args)
    vol = sum(m.h(l) ** 3 for (l, _) in m.leaves)
    assert abs(vol - 1.0) <= 1e-12
    # every leaf face: one same-level partner, a coarser leaf or nine finer leaves
    nsame = sum(1 for q in range(len(m.leaves)) for f in range(6) if m.nbr[q, f] >= 0)
    assert nsame == 2 * len(m.same)
    assert len(m.same) * 2 + len(m.desc) + len(m.anc) == 6 * len(m.leaves)
    assert len(m.desc) == 9 * len(m.anc)
    for (q, f, kids) in m.anc:
        lq = m.leaves[q][0]
        assert all(m.leaves[k][0] == lq + 1 for (k, _, _) in kids)
    for (q, f, r, fr, sub) in m.desc:
        assert m.leaves[r][0] == m.leaves[q][0] - 1 and fr == f ^ 1


@pytest.mark.parametrize("args", MESHES)
def test_library_mesh_matches_oracle(args):
    import exahype_amd as E
    if not __import__("os").path.exists(E.LIB_PATH):
        pytest.skip("library not built")
    m = A.Mesh(*args)
    lib = E.amr_mesh_cells(E.amr_config(*args))
    assert np.array_equal(lib, np.array([(l,) + c for (l, c) in m.leaves], np.int32))


def test_single_level_is_the_plain_grid_step():
    N, n = 3, 4
    og = A.LimitingGrid(A.Mesh(n, 1, (0.5, 0.5, 0.5), 0.0, []), N, relax=1e3)
    og.init(_vortex_ic())
    grid = O.OracleGrid(O.cfg(N), n, n, n)
    grid.init(_vortex_ic())
    # leaf order (z, y, x) is the grid's cell order
    assert np.array_equal(og.luh.ravel(), grid.luh())
    dt = og.stable_dt()
    assert dt == grid.stable_dt()
    for _ in range(2):
        d1 = og.step(dt)
        d2 = grid.step(dt)
        assert d1 == d2
        dt = d1
    assert np.array_equal(og.luh.ravel(), grid.luh())
    assert (og.status == 0).all() and not og.patch


@pytest.mark.parametrize("levels,widths", [(2, [0.01]), (3, [0.01, 0.01])])
def test_constant_state_is_a_fixed_point(levels, widths):
    N = 3
    og = A.LimitingGrid(A.Mesh(3, levels, (0.5, 0.5, 0.5), 0.0, widths), N)
    q0 = np.array([1.0, 0.1, -0.2, 0.3, 2.5])
    og.luh[:] = np.tile(q0, (N + 1) ** 3)
    og.fmin[:] = q0
    og.fmax[:] = q0
    dt = og.stable_dt()
    for _ in range(2):
        dt = og.step(dt)
    # to the 13-digit Kxi literals, which annihilate constants to ~1e-13 per 1/dx
    assert np.abs(og.luh - np.tile(q0, (N + 1) ** 3)).max() <= 1e-11
    assert (og.status == 0).all()


def test_status_rings_around_one_troubled_cell():
    N, n = 2, 11
    og = A.LimitingGrid(A.Mesh(n, 1, (0.5, 0.5, 0.5), 0.0, []), N)
    c = (5, 5, 5)
    q = og.m.index[(0, c)]
    og.status[q] = A.TROUBLED
    og.fw[q, :] = A.TROUBLED
    og._spread(5)
    for i, (l, cc) in enumerate(og.m.leaves):
        d = sum(min(abs(a - b), n - abs(a - b)) for a, b in zip(cc, c))
        assert og.status[i] == max(0, 5 - d), (cc, og.status[i])


def test_troubled_cell_update_is_the_fv_update_projected():
    """On a uniform mesh with a Sod sphere (troubled cells from the start),
    the step's new solution of a troubled cell equals MUSCL-Hancock on its
    patch (ghost layers from its neighbours' patches) projected onto the DG
    space, evaluated here independently of the driver."""
    N, n = 2, 5
    og = A.LimitingGrid(A.Mesh(n, 1, (0.5, 0.5, 0.5), 0.0, []), N)
    og.init(O.make_ic(kind=1, xc=(0.5, 0.5, 0.5), r0=0.3, rho_in=1.0, p_in=1.0, rho_out=0.125, p_out=0.1))
    T = np.nonzero(og.status == A.TROUBLED)[0]
    assert len(T) > 0
    q = int(T[0])
    S, P, G = og.S, og.P, A.GHOST
    before = {k: v.copy() for k, v in og.patch.items()}
    status0 = og.status.copy()
    dt = og.stable_dt()
    og.step(dt)
    # the patch of q with ghosts from its neighbours (all in the FV range or with patches)
    me = before[q].reshape(S, S, S, 5).copy()
    for d, s in itertools.product(range(3), (0, 1)):
        r = int(og.m.nbr[q, 2 * d + s])
        assert r in before and A.fv_merge(status0[q], status0[r])
        nb = before[r].reshape(S, S, S, 5)
        ax = 2 - d
        for k in range(G):
            dst = [slice(G, G + P)] * 3
            src = [slice(G, G + P)] * 3
            dst[ax] = G + P + k if s else k
            src[ax] = G + k if s else P + k
            me[tuple(dst)] = nb[tuple(src)]
    new, _ = O.musclhancock_update(og.c, P, me.ravel().copy(), me.ravel().copy(), og.dx(q), dt)
    inner = (slice(G, G + P),) * 3
    # the troubled cell's subcells move in this step
    assert np.abs(new.reshape(S, S, S, 5)[inner] - before[q].reshape(S, S, S, 5)[inner]).max() > 1e-3
    expect = O.project_on_dg(N, new, G)
    assert np.array_equal(og.luh[q], expect)
