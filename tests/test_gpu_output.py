"""Device sampling for the Cartesian-vertices output against the oracle's
restatement of ADERDG2CartesianVTK::plotVertexData, its exactness on
polynomials, and VTU and Peano patch files written from live grids."""
import numpy as np
import pytest

from oracle import oracle as O

from test_gpu_aderdg import TOL, rel_err
from test_output import read_vtu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [1, 3, 5, 9])
def test_sample_cartesian_vs_oracle_and_polynomials(N, tmp_path):
    import exahype_amd as E
    from exahype_amd import output as OUT
    from exahype_amd.tables import dg_math
    B = N + 1
    dims = (3, 2, 2)
    s = E.ADERDGSolver(N)
    g = E.Grid(s, *dims, offset=(0.1, 0.2, 0.3), width=(0.6, 0.4, 0.8))
    # a polynomial of degree N in every coordinate: sampled exactly
    xi = np.asarray(dg_math.gl_nodes_weights(N)[0])
    dx = np.array([0.2, 0.2, 0.4])
    rng = np.random.default_rng(N)
    c = rng.uniform(-1, 1, (5, 3, B))
    c[:, :, 0] += 2.0

    def poly(x, y, z):
        return np.stack([sum(c[v, 0, k] * x ** k + c[v, 1, k] * y ** k + c[v, 2, k] * z ** k for k in range(B))
                         for v in range(5)], -1)

    luh, ref = [], []
    eq = np.arange(B) / N
    for q in range(np.prod(dims)):
        ix, iy, iz = q % dims[0], (q // dims[0]) % dims[1], q // (dims[0] * dims[1])
        o = np.array([0.1, 0.2, 0.3]) + np.array([ix, iy, iz]) * dx
        Z, Y, X = np.meshgrid(xi, xi, xi, indexing="ij")
        luh.append(poly(o[0] + dx[0] * X, o[1] + dx[1] * Y, o[2] + dx[2] * Z).ravel())
        Z, Y, X = np.meshgrid(eq, eq, eq, indexing="ij")
        ref.append(poly(o[0] + dx[0] * X, o[1] + dx[1] * Y, o[2] + dx[2] * Z).reshape(-1, 5))
    g.upload(np.concatenate(luh))
    samp = g.sample_cartesian()
    cfg = O.cfg(N)
    for q in range(np.prod(dims)):
        assert rel_err(samp[q], O.sample_cartesian_vertices(cfg, luh[q])) <= TOL, q
        assert np.abs(samp[q] - ref[q]).max() <= 1e-11 * np.abs(ref[q]).max(), q
    f = tmp_path / "dg.vtu"
    OUT.aderdg_cartesian_vtu(str(f), g, time=0.0, binary=True)
    npts, ncell, arr, t = read_vtu(str(f))
    assert npts == np.prod(dims) * B ** 3 and ncell == np.prod(dims) * N ** 3
    assert np.array_equal(arr["Q"].reshape(-1, 5), samp.reshape(-1, 5))
    # the points are the sampled positions: the polynomial there is the value
    P = arr["Points"].reshape(-1, 3)
    assert np.abs(poly(P[:, 0], P[:, 1], P[:, 2]) - samp.reshape(-1, 5)).max() <= 1e-11 * np.abs(samp).max()


def test_fv_cells_vtu(tmp_path):
    import exahype_amd as E
    from exahype_amd import output as OUT
    s = E.FiniteVolumesSolver(2, 6)
    g = E.FVGrid(s, 4, 3, 1)
    g.set_initial_conditions(E.make_ic(5, xc=(0.4, 0.6, 0.5), L=0.01, amp=0.5, rho_inf=1.0, p_inf=1.0,
                                       uinf=(0.1, 0.1, 0.1)))
    f = tmp_path / "fv.vtu"
    OUT.fv_cells_vtu(str(f), g)
    npts, ncell, arr, _ = read_vtu(str(f))
    assert ncell == 12 * 36 and npts == 12 * 49
    luh = g.download().reshape(12, 8, 8, 5)[:, 1:7, 1:7]
    assert np.array_equal(arr["Q"].reshape(-1, 5), luh.reshape(-1, 5))


def test_aderdg_peano_patch_file(tmp_path):
    """Every cell one patch at its offset and size, its vertex values the
    device sampling, the time stamp at every vertex."""
    import exahype_amd as E
    from exahype_amd import output as OUT
    N, dims = 3, (3, 2, 2)
    s = E.ADERDGSolver(N)
    g = E.Grid(s, *dims, offset=(0.1, 0.2, 0.3), width=(0.6, 0.4, 0.8))
    g.set_initial_conditions(E.make_ic(4, rho_inf=1.0, p_inf=1.0, amp=0.2, uinf=(0.7, -0.4, 0.3), nmodes=1,
                                       k=[(1, 2, 1)], a=[1.0], phi=[0.4]))
    f = tmp_path / "dg.peano-patch-file"
    OUT.aderdg_cartesian_peano(str(f), g, time=0.125)
    r = OUT.read_peano_patch(str(f))
    assert r["dims"] == 3 and r["patch_size"] == N and r["kind"] == "vertex"
    assert np.array_equal(r["data"]["Q"], g.sample_cartesian())
    assert (r["data"]["time"] == 0.125).all()
    assert np.allclose(r["sizes"], [0.2, 0.2, 0.4]) and np.allclose(r["offsets"][4], [0.1 + 0.2, 0.2 + 0.2, 0.3])
