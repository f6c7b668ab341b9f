"""The mixed-precision predictor's matrix-core contraction, checked on the CPU.

contract_mfma (exahype_amd/csrc/aderdg_kernels.h) pairs the 3*V*B^2 pencils of
a time slice two per row of a 16x16 MFMA against blockdiag(Kxi, Kxi).  This
replays its index arithmetic lane by lane with the documented operand layouts
(A[l%16][4(l/16)+j], B[4(l/16)+j][l%16], D[4(l/16)+i][l%16]; the FP32 form
uses A[l%16][4s + l/16]) and checks, for every order it is built for, that
every LDS read and write stays inside the slice's pencil block, that every
output node is written exactly once, and that the result equals the direct
per-pencil contraction.  The GPU test of the operand layouts themselves is
tests/test_gpu_mixed.py::test_mfma_operand_layouts.
"""
import numpy as np
import pytest

V = 5


def kbd(K, B, k, n):
    if k < B and n < B:
        return K[k, n]
    if B <= k < 2 * B and B <= n < 2 * B:
        return K[k - B, n - B]
    return 0.0


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_contract_mfma_replay(N, bf16):
    B = N + 1
    B3 = B ** 3
    NT = (B3 + 63) // 64 * 64
    NITEMS = 3 * V * B * B
    NTILES = (NITEMS + 31) // 32
    NWAVES = NT // 64
    KS = (2 * B + 3) // 4
    rng = np.random.default_rng(N)
    K = rng.normal(size=(B, B))
    sF = rng.normal(size=NITEMS * B)
    sP = np.full(NITEMS * B, np.nan)
    writes = np.zeros(NITEMS * B, int)
    for wave in range(NWAVES):
        for tile in range(wave, NTILES, NWAVES):
            # gather operands lane by lane
            A = np.zeros((16, 16))
            Bm = np.zeros((16, 16))
            for lane in range(64):
                m, kq = lane & 15, lane >> 4
                row0 = tile * 32 + 2 * m
                lim = (NITEMS - row0) * B
                ks = [4 * kq + j for j in range(4)] if bf16 else [4 * s + kq for s in range(KS)]
                for k in ks:
                    if k < 2 * B and k < lim:
                        idx = row0 * B + k
                        assert 0 <= idx < NITEMS * B
                        A[m, k] = sF[idx]
                    Bm[k, m] = kbd(K, B, k, m)
            D = A @ Bm
            for lane in range(64):
                n, kq = lane & 15, lane >> 4
                if n < 2 * B:
                    for i in range(4):
                        r = tile * 32 + 2 * (4 * kq + i)
                        if r * B + n < NITEMS * B:
                            sP[r * B + n] = D[4 * kq + i, n]
                            writes[r * B + n] += 1
    assert (writes == 1).all()
    ref = (sF.reshape(NITEMS, B) @ K).ravel()
    assert np.allclose(sP, ref, rtol=1e-12, atol=1e-12)
