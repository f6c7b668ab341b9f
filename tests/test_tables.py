"""Pin the DG operator and quadrature tables every kernel reads.

The product (exahype_amd/csrc/dg_tables_data.h, uploaded to constant memory)
and the oracle (oracle/aderdg.c includes the same header) share these tables,
so GPU-vs-oracle parity cannot catch a table error.  They are pinned here
independently:

* bitwise against the reference's own table code
  (kernels/GaussLegendreQuadrature.cpp, GaussLobattoQuadrature.cpp,
  DGMatrices.cpp), compiled unmodified by ``make -C oracle ref`` into
  oracle/_ref/dump_ref_tables (skipped where that binary was not built:
  it needs the reference sources, present only in the build container);
* against tables derived from the mathematics (Gauss-Legendre quadrature,
  Lagrange basis, exahype_amd/tables/dg_math.py) to the precision of the
  reference's literals (13 significant digits for the matrices);
* Gauss-Lobatto nodes and weights against their closed-form definition
  (endpoints plus the roots of P_N', exactness to degree 2N-1).
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from exahype_amd.tables import dg_math

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "exahype_amd", "csrc", "dg_tables_data.h")
DUMP = os.path.join(ROOT, "oracle", "_ref", "dump_ref_tables")

DUMPED = ["gl_weights", "gl_nodes", "gll_weights", "gll_nodes", "F0", "FLCoeff", "FRCoeff", "Kxi", "iK1", "dudx",
          "equidistantGridProjector1d", "fineGridProjector1d"]


def _header_tables():
    txt = open(HEADER).read()
    out = {}
    for m in re.finditer(r"static const double EXA_TAB_(\w+)((?:\[\d+\])+) = (\{.*?\});", txt, re.S):
        shape = [int(x) for x in re.findall(r"\d+", m.group(2))]
        vals = [float.fromhex(v) for v in re.findall(r"[-+]?0x[0-9a-fA-F.]+p[-+]?\d+", m.group(3))]
        assert len(vals) == int(np.prod(shape)), m.group(1)
        out[m.group(1)] = np.array(vals).reshape(shape)
    return out


TABS = _header_tables()


def test_header_holds_every_table():
    assert set(DUMPED) | {"FCoeff"} == set(TABS)


@pytest.mark.skipif(not os.access(DUMP, os.X_OK),
                    reason="oracle/_ref/dump_ref_tables not built (make -C oracle ref needs the reference sources)")
def test_tables_bitwise_equal_to_reference_table_code():
    ref = json.loads(subprocess.check_output([DUMP]))
    assert set(ref) == set(DUMPED)
    for name in DUMPED:
        a = np.vectorize(float.fromhex)(np.array(ref[name]))
        b = TABS[name]
        assert a.shape == b.shape, name
        # bit patterns, so that -0.0 / 0.0 and every last ulp count
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), name


@pytest.mark.parametrize("N", range(10))
def test_tables_match_first_principles(N):
    B = N + 1
    M = dg_math.operators(N)
    # quadrature and extrapolation vectors: 16-digit literals
    for name in ("gl_nodes", "gl_weights", "FLCoeff", "FRCoeff", "F0"):
        np.testing.assert_allclose(TABS[name][N][:B], M[name], rtol=0, atol=1e-15, err_msg=name)
        assert not TABS[name][N][B:].any(), name + " padding"
    # operator matrices: 13-digit literals (measured <= 4.8e-13 of the largest entry)
    for name in ("Kxi", "iK1", "dudx", "equidistantGridProjector1d"):
        ref = M[name]
        err = np.abs(TABS[name][N][:B, :B] - ref).max() / max(1.0, np.abs(ref).max())
        assert err <= 1e-12, (name, err)
        assert not TABS[name][N][B:, :].any() and not TABS[name][N][:, B:].any(), name + " padding"
    err = np.abs(TABS["fineGridProjector1d"][N][:, :B, :B] - M["fineGridProjector1d"]).max()
    assert err <= 1e-12, ("fineGridProjector1d", err)
    # FCoeff = (FLCoeff, FRCoeff) at 13 digits
    np.testing.assert_allclose(TABS["FCoeff"][N][0][:B], TABS["FLCoeff"][N][:B], atol=1e-12)
    np.testing.assert_allclose(TABS["FCoeff"][N][1][:B], TABS["FRCoeff"][N][:B], atol=1e-12)


@pytest.mark.parametrize("N", range(1, 10))
def test_gauss_lobatto_closed_form(N):
    B = N + 1
    # nodes on [0,1]: the endpoints and the roots of P_N'
    inner = np.polynomial.legendre.Legendre.basis(N).deriv().roots()
    nodes = np.sort(np.concatenate([[-1.0], inner.real, [1.0]]))
    x = 0.5 * (nodes + 1.0)
    got_x = np.sort(TABS["gll_nodes"][N][:B])
    np.testing.assert_allclose(got_x, x, atol=1e-14)
    # weights: the quadrature on [0,1] is exact to degree 2N-1
    order = np.argsort(TABS["gll_nodes"][N][:B])
    w = TABS["gll_weights"][N][:B][order]
    for k in range(2 * N):
        assert abs(np.dot(w, got_x ** k) - 1.0 / (k + 1)) <= 1e-14, k


@pytest.mark.parametrize("N", range(10))
def test_table_invariants(N):
    B = N + 1
    w, xi = TABS["gl_weights"][N][:B], TABS["gl_nodes"][N][:B]
    assert abs(w.sum() - 1.0) <= 4e-15
    # F0 is the basis at t = 0 (FLCoeff); FRCoeff is FLCoeff mirrored
    assert np.array_equal(TABS["F0"][N][:B], TABS["FLCoeff"][N][:B])
    np.testing.assert_allclose(TABS["FRCoeff"][N][:B], TABS["FLCoeff"][N][:B][::-1], atol=1e-15)
    # Lagrange basis: partition of unity at 0 and 1, in the equidistant projector
    # and in each of the three fine-grid projectors
    assert abs(TABS["FLCoeff"][N][:B].sum() - 1.0) <= 1e-14
    np.testing.assert_allclose(TABS["equidistantGridProjector1d"][N][:B, :B].sum(0), 1.0, atol=1e-12)
    for s in range(3):
        np.testing.assert_allclose(TABS["fineGridProjector1d"][N][s][:B, :B].sum(0), 1.0, atol=1e-12)
    # the GL quadrature integrates degree 2N+1 exactly
    for k in range(2 * N + 2):
        assert abs(np.dot(w, xi ** k) - 1.0 / (k + 1)) <= 1e-14, k
