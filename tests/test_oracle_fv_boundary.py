"""Pins of the oracle's FV boundary hooks (oracle/fv.c: boundaryLayerExtraction,
ghostLayerFillingAtBoundary, ghostLayerFilling, boundaryConditions and the
closed axes of the FV grid driver), each against an independent numpy
restatement by array slicing or against a conservation law:

* ghostLayerFilling copies the neighbour's G interior layers next to the
  shared face into my ghost layers (slices of the [z][y][x][v] patch);
* extraction from the neighbour followed by ghostLayerFillingAtBoundary is
  ghostLayerFilling -- except where the reference's 3D y- and z- extraction
  reads layer patchBegin for every k (commons.cpph:561, :591), checked here
  as that layer repeated;
* outflow boundary values mirror the layers (K = G-1-k), a reflecting wall
  negates the normal momentum, exact data sit at the outside subcell
  centres the reference computes (the 3D x+ face with size[2]);
* a box closed by reflecting walls conserves mass and energy for Godunov
  (2D demonstrator, 3D); MUSCL-Hancock patches are not conservative across
  patch faces even when periodic (each patch predicts its ghost subcells
  from face ghosts only), so there a gas at rest must stay at rest.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _patch(dim, P, G, rng, nd=5):
    S = P + 2 * G
    return rng.standard_normal((S,) * dim + (nd,))


def _slab(dim, G, P, d, lo, hi):
    """index of layers [lo, hi) along axis d (x = last spatial axis), interior elsewhere"""
    idx = [slice(G, G + P)] * dim
    idx[dim - 1 - d] = slice(lo, hi)
    return tuple(idx)


CASES = [(2, 5, 1), (2, 4, 2), (3, 4, 1), (3, 3, 2)]


@pytest.mark.parametrize("dim,P,G", CASES)
def test_ghost_layer_filling_is_a_slab_copy(dim, P, G):
    c = O.cfg(3)
    rng = np.random.default_rng(dim * 10 + P + G)
    me, nb = _patch(dim, P, G, rng), _patch(dim, P, G, rng)
    for d in range(dim):
        for s in (-1, 1):
            pos = [0, 0, 0]
            pos[d] = s
            out = O.fv_ghost_layer_filling(c, dim, P, G, me.ravel(), nb.ravel(), pos).reshape(me.shape)
            exp = me.copy()
            if s < 0:
                exp[_slab(dim, G, P, d, 0, G)] = nb[_slab(dim, G, P, d, P, P + G)]
            else:
                exp[_slab(dim, G, P, d, G + P, 2 * G + P)] = nb[_slab(dim, G, P, d, G, 2 * G)]
            assert np.array_equal(out, exp), (d, s)


@pytest.mark.parametrize("dim,P,G", CASES)
def test_extraction_then_filling_at_boundary_is_ghost_filling(dim, P, G):
    c = O.cfg(3)
    rng = np.random.default_rng(7 + dim + P + G)
    me, nb = _patch(dim, P, G, rng), _patch(dim, P, G, rng)
    for d in range(dim):
        for s in (-1, 1):
            pos, opp = [0, 0, 0], [0, 0, 0]
            pos[d], opp[d] = s, -s
            bnd = O.fv_boundary_layer_extraction(c, dim, P, G, nb.ravel(), opp)
            got = O.fv_ghost_layer_filling_at_boundary(c, dim, P, G, me.ravel(), bnd, pos).reshape(me.shape)
            if dim == 3 and d > 0 and s > 0:
                # the neighbour's lower face gives its first interior layer G times
                exp = me.copy()
                layer = nb[_slab(dim, G, P, d, G, G + 1)]
                exp[_slab(dim, G, P, d, G + P, 2 * G + P)] = np.concatenate([layer] * G, axis=dim - 1 - d)
            else:
                exp = O.fv_ghost_layer_filling(c, dim, P, G, me.ravel(), nb.ravel(), pos).reshape(me.shape)
            assert np.array_equal(got, exp), (d, s)


def _bnd_shape(dim, P, G, d):
    if dim == 2:
        return (P, G, 5) if d == 0 else (G, P, 5)
    return [(P, P, G, 5), (P, G, P, 5), (G, P, P, 5)][d]


@pytest.mark.parametrize("dim,P,G", CASES)
def test_outflow_and_wall_mirror_the_layers(dim, P, G):
    c = O.cfg(3)
    rng = np.random.default_rng(3 + dim + G)
    for d in range(dim):
        shp = _bnd_shape(dim, P, G, d)
        ax = [2, 1, 0][d] if dim == 3 else [1, 0][d]  # the layer axis of the face block
        bin_ = rng.standard_normal(shp)
        for side in (0, 1):
            out = O.fv_boundary_conditions(c, O.make_bc(O.BC_OUTFLOW), dim, P, G, bin_.ravel(), (0.5,) * 3,
                                           (0.1,) * 3, 0.0, 1e-3, 2 * d + side, d).reshape(shp)
            assert np.array_equal(out, np.flip(bin_, axis=ax))
            out = O.fv_boundary_conditions(c, O.make_bc(O.BC_REFLECTING), dim, P, G, bin_.ravel(), (0.5,) * 3,
                                           (0.1,) * 3, 0.0, 1e-3, 2 * d + side, d).reshape(shp)
            exp = np.flip(bin_, axis=ax).copy()
            exp[..., 1 + d] *= -1
            assert np.array_equal(out, exp)


def _entropy_wave():
    return O.make_ic(4, rho_inf=1.0, p_inf=1.0, amp=0.2, uinf=(0.7, -0.4, 0.3), nmodes=1, k=[(1, 2, 3)], a=[1.0],
                     phi=[0.4])


def _rho(x, t):
    u = np.array([0.7, -0.4, 0.3])
    return 1.0 + 0.2 * np.sin(2 * np.pi * np.dot([1, 2, 3], np.asarray(x) - u * t) + 0.4)


@pytest.mark.parametrize("face", range(6))
def test_exact_values_at_the_outside_subcell_centres_3d(face):
    c = O.cfg(3)
    P, G = 3, 2
    d, side = face // 2, face % 2
    centre, size, t = np.array([0.3, 0.55, 0.7]), np.array([0.1, 0.12, 0.08]), 0.2
    shp = _bnd_shape(3, P, G, d)
    out = O.fv_boundary_conditions(c, O.make_bc(O.BC_EXACT, _entropy_wave()), 3, P, G, np.zeros(int(np.prod(shp))),
                                   centre, size, t, 1e-3, face, d).reshape(shp)
    o1, o2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
    for k in range(G):
        K = G - 1 - k
        for a in range(P):
            for b in range(P):
                x = np.zeros(3)
                x[d] = centre[d] + (0.5 * size[d] * (1 + k / (2 * P)) if side else
                                    -0.5 * size[d] * (1 + K / (2 * P)))
                x[o2] = centre[o2] + size[o2] * ((2 * a + 1) / (2 * P) - 0.5)
                x[o1] = centre[o1] + size[2 if (d == 0 and side) else o1] * ((2 * b + 1) / (2 * P) - 0.5)
                e = [(a, b, k), (a, k, b), (k, a, b)][d]
                assert abs(out[e][0] - _rho(x, t)) <= 1e-14, (face, k, a, b)


def _closed_box(dim, scheme, P, n, periodic, kind, steps):
    c = O.cfg(3)
    og = O.OracleFVGrid(c, dim, scheme, P, n, n + 1, n if dim == 3 else 1)
    ic = O.make_ic(0, xc=(0.5, 0.5, 0.5), L=0.1, eps=5.0, uinf=(0.6, -0.3, 0.2), rho_inf=1.0, p_inf=1.0, amp=0.05,
                   nmodes=1, k=[(1, 1, 1)], a=[1.0], phi=[0.3])
    og.init(ic)
    og.set_boundary(periodic, O.make_bc(kind))
    S, G = P + 2 * og.g.ghost, og.g.ghost
    inner = (slice(None),) + (slice(G, G + P),) * dim

    def totals():
        u = og.luh().reshape((-1,) + (S,) * dim + (5,))[inner]
        return u.reshape(-1, 5).sum(axis=0)

    m0 = totals()
    dt = og.stable_dt()
    for _ in range(steps):
        dt, _ = og.step(dt)
    return m0, totals(), og


@pytest.mark.parametrize("dim,scheme,periodic", [(2, 0, (0, 0, 1)), (2, 0, (1, 0, 1)), (3, 0, (0, 0, 0)),
                                                 (3, 0, (1, 0, 1))])
def test_reflecting_box_conserves_mass_and_energy(dim, scheme, periodic):
    m0, m1, og = _closed_box(dim, scheme, 4, 3, periodic, O.BC_REFLECTING, 5)
    for v in (0, 4):
        assert abs(m1[v] - m0[v]) <= 1e-13 * abs(m0[v]), (v, m0[v], m1[v])
    assert og.g.t > 0


@pytest.mark.parametrize("dim,scheme,kind,q0", [(2, 0, O.BC_OUTFLOW, (1.3, 0.2, -0.1, 0.05, 2.4)),
                                               (3, 0, O.BC_OUTFLOW, (1.3, 0.2, -0.1, 0.05, 2.4)),
                                               (3, 1, O.BC_OUTFLOW, (1.3, 0.2, -0.1, 0.05, 2.4)),
                                               (3, 1, O.BC_REFLECTING, (1.3, 0.0, 0.0, 0.0, 2.4)),
                                               (2, 0, O.BC_REFLECTING, (1.3, 0.0, 0.0, 0.0, 2.4))])
def test_constant_state_stays(dim, scheme, kind, q0):
    c = O.cfg(3)
    og = O.OracleFVGrid(c, dim, scheme, 4, 3, 2, 2 if dim == 3 else 1)
    og.init(O.make_ic(2, q0=q0))
    og.set_boundary((0, 0, 0), O.make_bc(kind))
    G, P, S = og.g.ghost, 4, 4 + 2 * og.g.ghost
    inner = (slice(None),) + (slice(G, G + P),) * dim
    u0 = og.luh().reshape((-1,) + (S,) * dim + (5,))[inner].copy()
    dt = og.stable_dt()
    for _ in range(3):
        dt, _ = og.step(dt)
    u1 = og.luh().reshape((-1,) + (S,) * dim + (5,))[inner]
    assert np.abs(u1 - u0).max() <= 1e-13
