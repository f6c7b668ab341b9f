"""Pins of the oracle's AMR level transfers (kernels/aderdg/generic/c/3d/amrRoutines.cpph).

What fixes them, independently of the projector table the oracle reads:
  * the reference's own tests (exahype/tests/kernels/c/GenericEulerKernelTest3D.cpp:775-918):
    constants survive prolongation, and restricting a constant from every
    sub-cell sums to the constant;
  * prolongation is exact for polynomials of degree N: the fine nodal values are
    the coarse polynomial evaluated at (i + xi_j)/3^levels on each axis, with the
    Lagrange basis built here from numpy's Gauss-Legendre nodes.  Anisotropic
    polynomials pin which sub-index digit drives which axis and that the
    digits are read most significant first;
  * restriction is the L2 projection (GL mass is exact and diagonal at degree
    2N): coarse_m = (1/w_m) * integral over the sub-cell of phi_m * f, computed
    here with a finer numpy quadrature;
  * prolongate -> restrict over all 3^(d*levels) descendants returns the coarse
    data.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def gl(B):
    x, w = np.polynomial.legendre.leggauss(B)
    return 0.5 * (x + 1), 0.5 * w


def lagrange(nodes, x):
    """phi_k(x) for all k: shape (len(x), B)."""
    x = np.atleast_1d(x)
    B = len(nodes)
    out = np.ones((len(x), B))
    for k in range(B):
        for j in range(B):
            if j != k:
                out[:, k] *= (x - nodes[j]) / (nodes[k] - nodes[j])
    return out


def poly_field(rng, N, dim, nd):
    """Random tensor polynomial of degree N per axis and variable: coeffs[v][i..]."""
    return rng.normal(size=(nd,) + (N + 1,) * dim)


def eval_poly(c, pts):
    """c: (nd, B, B[, B]) monomial coeffs in (x, y[, z]); pts: list of axis coordinates (x fastest)."""
    nd = c.shape[0]
    dim = c.ndim - 1
    B = c.shape[1]
    V = [np.vander(p, B, increasing=True) for p in pts]  # (npts_axis, B)
    if dim == 2:
        x, y = V
        # out[b][a][v]: a = y index, b = x index -> layout [y][x][v]
        val = np.einsum("vij,yj,xi->yxv", c, y, x)
    else:
        x, y, z = V
        val = np.einsum("vijk,zk,yj,xi->zyxv", c, z, y, x)
    return val.reshape(-1, nd) if dim == 2 else val.reshape(-1, nd)


def fine_points(N, levels, idx):
    nodes, _ = gl(N + 1)
    s = 3 ** levels
    return [(i + nodes) / s for i in idx]


@pytest.mark.parametrize("N", [1, 3, 5, 7, 9])
@pytest.mark.parametrize("levels", [1, 2, 3])
def test_volume_prolongation_is_polynomial_evaluation(N, levels):
    rng = np.random.default_rng(100 * N + levels)
    nodes, _ = gl(N + 1)
    c = poly_field(rng, N, 3, 4)
    coarse = eval_poly(c, [nodes] * 3).ravel()
    n = 3 ** levels
    for idx in [(0, 0, 0), (n - 1, 1 % n, n // 2), tuple(rng.integers(0, n, 3))]:
        fine = O.volume_unknowns_prolongation(N, coarse, 2, 2 + levels, idx, ndata=4)
        ref = eval_poly(c, fine_points(N, levels, idx)).ravel()
        assert np.abs(fine - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("N", [2, 3, 5, 9])
@pytest.mark.parametrize("levels", [1, 2])
def test_face_prolongation_is_polynomial_evaluation(N, levels):
    rng = np.random.default_rng(7 * N + levels)
    nodes, _ = gl(N + 1)
    nvar, npar = 3, 2
    cq = poly_field(rng, N, 2, nvar + npar)
    cf = poly_field(rng, N, 2, nvar)
    qc = eval_poly(cq, [nodes] * 2).ravel()
    fc = eval_poly(cf, [nodes] * 2).ravel()
    n = 3 ** levels
    for idx in [(0, n - 1), (n - 1, 0), tuple(rng.integers(0, n, 2))]:
        qf, ff = O.face_unknowns_prolongation(N, qc, fc, 0, levels, idx, nvar=nvar, npar=npar)
        pts = fine_points(N, levels, idx)
        rq, rf = eval_poly(cq, pts).ravel(), eval_poly(cf, pts).ravel()
        assert np.abs(qf - rq).max() <= 1e-9 * np.abs(rq).max()
        assert np.abs(ff - rf).max() <= 1e-9 * np.abs(rf).max()


def l2_restriction(N, fine_vals, levels, idx, nd):
    """(1/w_m) * integral over the sub-cell of phi_m f, f the fine nodal polynomial."""
    dim = len(idx)
    nodes, w = gl(N + 1)
    qx, qw = gl(N + 3)
    s = 3 ** levels
    f = fine_vals.reshape((N + 1,) * dim + (nd,))
    # fine polynomial at fine-local quadrature points, per axis interpolation
    Lf = lagrange(nodes, qx)                   # (nq, B): fine basis at local points
    phys = [(i + qx) / s for i in idx]          # physical points per axis (x, y[, z])
    Lc = [lagrange(nodes, p) for p in phys]     # coarse basis at those points
    if dim == 2:
        fq = np.einsum("yxv,qy,px->qpv", f, Lf, Lf)  # q: y points, p: x points
        out = np.einsum("qpv,q,p,qm,pn->mnv", fq, qw, qw, Lc[1], Lc[0]) / s ** 2
        out /= (w[:, None] * w[None, :])[:, :, None]
    else:
        fq = np.einsum("zyxv,rz,qy,px->rqpv", f, Lf, Lf, Lf)
        out = np.einsum("rqpv,r,q,p,rl,qm,pn->lmnv", fq, qw, qw, qw, Lc[2], Lc[1], Lc[0]) / s ** 3
        out /= (w[:, None, None] * w[None, :, None] * w[None, None, :])[..., None]
    return out.ravel()


@pytest.mark.parametrize("N", [1, 3, 5])
@pytest.mark.parametrize("levels", [1, 2])
def test_volume_restriction_is_l2_projection(N, levels):
    rng = np.random.default_rng(31 * N + levels)
    nd = 3
    n = 3 ** levels
    for idx in [(0, 0, 0), (n - 1, 0, 1), tuple(rng.integers(0, n, 3))]:
        fine = rng.normal(size=(N + 1) ** 3 * nd)
        base = rng.normal(size=fine.size)
        got = O.volume_unknowns_restriction(N, base, fine, 1, 1 + levels, idx, ndata=nd)
        ref = base + l2_restriction(N, fine, levels, idx, nd)
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("N", [2, 4, 9])
@pytest.mark.parametrize("levels", [1, 2])
def test_face_restriction_is_l2_projection(N, levels):
    rng = np.random.default_rng(5 * N + levels)
    nvar, npar = 5, 1
    B2 = (N + 1) ** 2
    n = 3 ** levels
    for idx in [(0, n - 1), (n - 1, 1), tuple(rng.integers(0, n, 2))]:
        qf, ff = rng.normal(size=B2 * (nvar + npar)), rng.normal(size=B2 * nvar)
        qc0, fc0 = rng.normal(size=qf.size), rng.normal(size=ff.size)
        qc, fc = O.face_unknowns_restriction(N, qc0, fc0, qf, ff, 3, 3 + levels, idx, nvar=nvar, npar=npar)
        rq = qc0 + l2_restriction(N, qf, levels, idx, nvar + npar)
        rf = fc0 + l2_restriction(N, ff, levels, idx, nvar)
        assert np.abs(qc - rq).max() <= 1e-11 * np.abs(rq).max()
        assert np.abs(fc - rf).max() <= 1e-11 * np.abs(rf).max()


@pytest.mark.parametrize("levels", [1, 2])
def test_constants_reference_kernel_test(levels):
    """GenericEulerKernelTest3D.cpp:775-918: nVar = nPar = 1, order 3, constant 1.
    Tolerance 1e-9 as the reference's volume test (:894): the projector table
    holds 13 significant digits, so sums of its rows miss 1 by ~1e-12."""
    N, B = 3, 4
    one_v = np.ones(B ** 3 * 2)
    n = 3 ** levels
    coarse = np.zeros_like(one_v)
    for idx in itertools.product(range(n), repeat=3):
        fine = O.volume_unknowns_prolongation(N, one_v, 0, levels, idx[::-1], ndata=2)
        assert np.abs(fine - 1).max() <= 1e-9
        coarse = O.volume_unknowns_restriction(N, coarse, one_v, 0, levels, idx[::-1], ndata=2)
    assert np.abs(coarse - 1).max() <= 1e-9
    qc, fc = np.zeros(B * B * 2), np.zeros(B * B)
    for idx in itertools.product(range(n), repeat=2):
        qf, ff = O.face_unknowns_prolongation(N, np.ones(B * B * 2), np.ones(B * B), 0, levels, idx, nvar=1, npar=1)
        assert np.abs(qf - 1).max() <= 1e-9 and np.abs(ff - 1).max() <= 1e-9
        qc, fc = O.face_unknowns_restriction(N, qc, fc, qf, ff, 0, levels, idx, nvar=1, npar=1)
    assert np.abs(qc - 1).max() <= 1e-9 and np.abs(fc - 1).max() <= 1e-9


@pytest.mark.parametrize("N", [3, 5])
def test_prolongate_then_restrict_all_children_is_identity(N):
    rng = np.random.default_rng(N)
    nd = 5
    coarse = rng.normal(size=(N + 1) ** 3 * nd)
    acc = np.zeros_like(coarse)
    for idx in itertools.product(range(3), repeat=3):
        fine = O.volume_unknowns_prolongation(N, coarse, 4, 5, idx, ndata=nd)
        acc = O.volume_unknowns_restriction(N, acc, fine, 4, 5, idx, ndata=nd)
    assert np.abs(acc - coarse).max() <= 1e-10 * np.abs(coarse).max()
