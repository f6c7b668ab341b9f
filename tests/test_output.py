"""VTK XML output (exahype_amd/output.py): files parse back to the points,
cells and values written, in ASCII and in base64 binary; the cell lattices
tile each cell."""
import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from exahype_amd import output as OUT


def _read_array(el):
    tname = el.get("type")
    dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[tname]
    if el.get("format") == "binary":
        raw = base64.b64decode(el.text)
        n = int.from_bytes(raw[:4], "little")
        return np.frombuffer(raw[4:4 + n], dtype=dt)
    return np.array(el.text.split(), dtype=dt)


def read_vtu(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for el in piece.iter("DataArray"):
        arrays[el.get("Name")] = _read_array(el)
    time = root.find("UnstructuredGrid/FieldData/DataArray")
    return (int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays,
            None if time is None else _read_array(time)[0])


@pytest.mark.parametrize("binary", [False, True])
def test_vtu_round_trip(tmp_path, binary):
    pts = np.random.default_rng(0).random((8, 3))
    cells = np.array([[0, 1, 2, 3, 4, 5, 6, 7]])
    q = np.arange(40.0).reshape(8, 5) / 7.0
    f = tmp_path / "t.vtu"
    OUT.write_vtu(str(f), pts, cells, OUT.VTK_HEXAHEDRON, point_data={"Q": q}, time=0.25, binary=binary)
    npts, ncell, arr, t = read_vtu(str(f))
    assert (npts, ncell, t) == (8, 1, 0.25)
    assert np.array_equal(arr["Points"].reshape(-1, 3), pts)
    assert np.array_equal(arr["connectivity"], cells.ravel())
    assert np.array_equal(arr["offsets"], [8]) and np.array_equal(arr["types"], [12])
    assert np.array_equal(arr["Q"].reshape(8, 5), q)


def test_lattice_tiles_each_cell():
    dx = (0.5, 0.25, 1.0)
    pts, cells = OUT._lattice(2, 1, 1, (1.0, 0.0, -1.0), dx, np.array([0.0, 0.3, 1.0]))
    assert pts.shape == (2 * 27, 3) and cells.shape == (2 * 8, 8)
    vol = 0.0
    for c in cells:
        p = pts[c]
        lo, hi = p.min(0), p.max(0)
        # the 8 corners are the corners of an axis-aligned box
        assert len({tuple(x) for x in p}) == 8
        assert np.all((np.isclose(p, lo) | np.isclose(p, hi)))
        vol += np.prod(hi - lo)
    assert vol == pytest.approx(2 * np.prod(dx))
    assert pts[:, 0].min() == 1.0 and pts[:, 0].max() == pytest.approx(2.0)


@pytest.mark.parametrize("kind,dims", [("vertex", 3), ("cell", 2)])
def test_peano_patch_round_trip(tmp_path, kind, dims):
    from exahype_amd import output as OUT
    rng = np.random.default_rng(3)
    P, npatch = 3, 4
    npts = (P + 1) ** dims if kind == "vertex" else P ** dims
    offsets, sizes = rng.uniform(0, 1, (npatch, dims)), rng.uniform(0.1, 0.2, (npatch, dims))
    q, t = rng.standard_normal((npatch, npts, 5)), np.full((npatch, npts, 1), 0.25)
    f = tmp_path / "p.peano-patch-file"
    OUT.write_peano_patch(str(f), offsets, sizes, P, {"Q": q, "time": t}, kind=kind, dims=dims)
    r = OUT.read_peano_patch(str(f))
    assert r["dims"] == dims and r["patch_size"] == P and r["kind"] == kind
    assert r["unknowns"] == {"Q": 5, "time": 1}
    assert np.array_equal(r["offsets"], offsets) and np.array_equal(r["sizes"], sizes)
    assert np.array_equal(r["data"]["Q"], q) and np.array_equal(r["data"]["time"], t)
    head = f.read_text().splitlines()[:8]
    assert "# Peano patch file " in head and "format ASCII " in head
