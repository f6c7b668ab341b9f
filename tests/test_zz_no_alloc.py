"""The allocation-free batched restriction (exa_amr_restriction_batched_noalloc)
against exa_amr_restriction_batched: the same device inputs give bitwise the
same coarse arrays, on two calls in a row (the second smaller, of another
order and dimension, so stale staging from the first call would show), and
the size limit and the missing-init case are refused with EXA_ERR_ARG."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _restrict_both(E, dim, N, nd, ncoarse, kids_per_coarse, levels, seed):
    rng = np.random.default_rng(seed)
    blk = (N + 1) ** dim * nd
    nfine = ncoarse * kids_per_coarse
    fine = rng.normal(size=nfine * blk)
    coarse0 = rng.normal(size=ncoarse * blk)
    span = 3 ** levels
    tr = []
    for c in range(ncoarse):
        for k in range(kids_per_coarse):
            idx = tuple(int(i) for i in rng.integers(0, span, 3))
            tr.append(((c * kids_per_coarse + k) * blk, c * blk, levels, idx))
    d_f = E.DeviceArray(fine)
    d_a = E.DeviceArray(coarse0)
    d_b = E.DeviceArray(coarse0)
    E.amrRestrictionBatched(dim, N, nd, d_a, d_f, tr)
    E.amrRestrictionBatchedNoAlloc(dim, N, nd, d_b, d_f, tr)
    return d_a.numpy(), d_b.numpy(), coarse0


def test_noalloc_restriction_matches_allocating_form_twice():
    import exahype_amd as E
    print("noalloc: call 1 (volumes, N=5, 40 coarse blocks x 27)", flush=True)
    a, b, c0 = _restrict_both(E, 3, 5, 5, 40, 27, 1, seed=1)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c0)
    print("noalloc: call 2 (faces, N=3, 6 coarse blocks x 9, two levels)", flush=True)
    a2, b2, c2 = _restrict_both(E, 2, 3, 6, 6, 9, 2, seed=2)
    assert np.array_equal(a2, b2)
    assert not np.array_equal(a2, c2)


def test_noalloc_restriction_refuses_oversized_list():
    import exahype_amd as E
    N, nd = 1, 1
    blk = (N + 1) ** 3 * nd
    n = E.AMR_NOALLOC_MAX_TRANSFERS + 1
    d_f = E.DeviceArray(np.zeros(n * blk))
    d_c = E.DeviceArray(np.zeros(blk))
    tr = [(i * blk, 0, 1, (0, 0, 0)) for i in range(n)]
    with pytest.raises(RuntimeError, match="exceed"):
        E.amrRestrictionBatchedNoAlloc(3, N, nd, d_c, d_f, tr)


def test_noalloc_restriction_all_children_round_trip():
    """27 children prolongated from one coarse cell restrict back to it."""
    import exahype_amd as E
    N, nd = 4, 5
    blk = (N + 1) ** 3 * nd
    rng = np.random.default_rng(7)
    coarse = rng.normal(size=blk)
    kids = list(itertools.product(range(3), repeat=3))
    tr = [(i * blk, 0, 1, (k[2], k[1], k[0])) for i, k in enumerate(kids)]
    d_c = E.DeviceArray(coarse)
    d_f = E.DeviceArray(np.zeros(len(kids) * blk))
    E.amrProlongationBatched(3, N, nd, d_f, d_c, tr)
    d_back = E.DeviceArray(np.zeros(blk))
    E.amrRestrictionBatchedNoAlloc(3, N, nd, d_back, d_f, tr)
    assert np.abs(d_back.numpy() - coarse).max() <= 1e-10 * np.abs(coarse).max()
