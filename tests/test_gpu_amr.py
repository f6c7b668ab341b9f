"""GPU parity of the AMR level transfers through the C ABI vs the oracle
(kernels/aderdg/generic/c/3d/amrRoutines.cpph).

Bar: <= 1e-12 relative to the largest magnitude of the expected block (the GPU
multiplies the per-level 1D operators into one matrix per axis and applies
them sum-factorised, so it differs from the oracle's full tensor products by
rounding only).  Covers orders 0..9, level jumps 1..3, faces with material
parameters, the per-call drop-ins and the batched device forms (including a
run of restrictions of all 27 children into one coarse cell).
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def close(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() <= TOL * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("N", [0, 1, 3, 5, 7, 9])
@pytest.mark.parametrize("levels", [1, 2, 3])
def test_volume_transfers_vs_oracle(N, levels):
    import exahype_amd as E
    rng = np.random.default_rng(10 * N + levels)
    nvar, npar = 5, 1
    nd = nvar + npar
    n = 3 ** levels
    for idx in [(0, 0, 0), (n - 1, n // 2, 1 % n), tuple(rng.integers(0, n, 3))]:
        coarse = rng.normal(size=(N + 1) ** 3 * nd)
        g = E.volumeUnknownsProlongation(coarse, 1, 1 + levels, idx, N, nvar=nvar, npar=npar)
        o = O.volume_unknowns_prolongation(N, coarse, 1, 1 + levels, idx, ndata=nd)
        assert close(g, o), (N, levels, idx)
        fine = rng.normal(size=coarse.size)
        g = E.volumeUnknownsRestriction(coarse, fine, 1, 1 + levels, idx, N, nvar=nvar, npar=npar)
        o = O.volume_unknowns_restriction(N, coarse, fine, 1, 1 + levels, idx, ndata=nd)
        assert close(g, o), (N, levels, idx)


@pytest.mark.parametrize("N", [0, 2, 3, 5, 9])
@pytest.mark.parametrize("levels", [1, 2])
def test_face_transfers_vs_oracle(N, levels):
    import exahype_amd as E
    rng = np.random.default_rng(3 * N + levels)
    nvar, npar = 5, 2
    B2 = (N + 1) ** 2
    n = 3 ** levels
    for idx in [(0, n - 1), (n - 1, 0), tuple(rng.integers(0, n, 2))]:
        qc, fc = rng.normal(size=B2 * (nvar + npar)), rng.normal(size=B2 * nvar)
        gq, gf = E.faceUnknownsProlongation(qc, fc, 0, levels, idx, N, nvar=nvar, npar=npar)
        oq, of = O.face_unknowns_prolongation(N, qc, fc, 0, levels, idx, nvar=nvar, npar=npar)
        assert close(gq, oq) and close(gf, of)
        qf, ff = rng.normal(size=qc.size), rng.normal(size=fc.size)
        gq, gf = E.faceUnknownsRestriction(qc, fc, qf, ff, 0, levels, idx, N, nvar=nvar, npar=npar)
        oq, of = O.face_unknowns_restriction(N, qc, fc, qf, ff, 0, levels, idx, nvar=nvar, npar=npar)
        assert close(gq, oq) and close(gf, of)


def test_invalid_sub_index_is_an_error():
    import exahype_amd as E
    with pytest.raises(RuntimeError):
        E.volumeUnknownsProlongation(np.ones(64 * 5), 0, 1, (3, 0, 0), 3)
    with pytest.raises(RuntimeError):
        E.volumeUnknownsProlongation(np.ones(64 * 5), 2, 2, (0, 0, 0), 3)


@pytest.mark.parametrize("N", [3, 5])
def test_batched_volume_all_children(N):
    """27 prolongations of one coarse cell, then the 27 restrictions back into a
    zeroed coarse cell as one consecutive run: matches the oracle's sequential
    accumulation and returns the coarse data."""
    import exahype_amd as E
    rng = np.random.default_rng(N)
    nd = 5
    blk = (N + 1) ** 3 * nd
    coarse = rng.normal(size=blk)
    kids = list(itertools.product(range(3), repeat=3))
    tr = [(i * blk, 0, 1, (k[2], k[1], k[0])) for i, k in enumerate(kids)]
    d_coarse = E.DeviceArray(coarse)
    d_fine = E.DeviceArray(np.zeros(len(kids) * blk))
    E.amrProlongationBatched(3, N, nd, d_fine, d_coarse, tr)
    fine = d_fine.numpy()
    for i, k in enumerate(kids):
        o = O.volume_unknowns_prolongation(N, coarse, 0, 1, (k[2], k[1], k[0]), ndata=nd)
        assert close(fine[i * blk:(i + 1) * blk], o)
    d_back = E.DeviceArray(np.zeros(blk))
    E.amrRestrictionBatched(3, N, nd, d_back, d_fine, tr)
    acc = np.zeros(blk)
    for i, k in enumerate(kids):
        acc = O.volume_unknowns_restriction(N, acc, fine[i * blk:(i + 1) * blk], 0, 1, (k[2], k[1], k[0]), ndata=nd)
    back = d_back.numpy()
    assert close(back, acc)
    assert np.abs(back - coarse).max() <= 1e-10 * np.abs(coarse).max()


def test_batched_faces_many_pairs_and_checks():
    import exahype_amd as E
    N, nd, levels = 4, 5, 2
    rng = np.random.default_rng(11)
    blk = (N + 1) ** 2 * nd
    ncoarse, nfine = 7, 40
    coarse = rng.normal(size=ncoarse * blk)
    fine_in = rng.normal(size=nfine * blk)
    tr = [(i * blk, (i % ncoarse) * blk, levels, tuple(rng.integers(0, 9, 2))) for i in range(nfine)]
    d_c = E.DeviceArray(coarse)
    d_f = E.DeviceArray(np.zeros(nfine * blk))
    E.amrProlongationBatched(2, N, nd, d_f, d_c, tr)
    f = d_f.numpy()
    for i, (fo, co, lv, ix) in enumerate(tr):
        o, _ = O.face_unknowns_prolongation(N, coarse[co:co + blk], np.zeros(blk),
                                            0, lv, ix, nvar=nd, npar=0)
        assert close(f[fo:fo + blk], o)
    # restrictions grouped by coarse block (consecutive runs), accumulated in list order
    order = sorted(range(nfine), key=lambda i: tr[i][1])
    rtr = [tr[i] for i in order]
    d_f2 = E.DeviceArray(fine_in)
    d_c2 = E.DeviceArray(coarse)
    E.amrRestrictionBatched(2, N, nd, d_c2, d_f2, rtr)
    exp = coarse.copy()
    for fo, co, lv, ix in rtr:
        q, _ = O.face_unknowns_restriction(N, exp[co:co + blk], np.zeros((N + 1) ** 2 * nd), fine_in[fo:fo + blk],
                                           np.zeros((N + 1) ** 2 * nd), 0, lv, ix, nvar=nd, npar=0)
        exp[co:co + blk] = q
    assert close(d_c2.numpy(), exp)
    # a coarse block split into two runs, and a block past the end, are refused
    with pytest.raises(RuntimeError):
        E.amrRestrictionBatched(2, N, nd, d_c2, d_f2, tr)
    with pytest.raises(RuntimeError):
        E.amrProlongationBatched(2, N, nd, d_f, d_c, [((nfine - 1) * blk + 1, 0, 1, (0, 0))])
