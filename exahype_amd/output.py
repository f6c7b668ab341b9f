"""VTK XML unstructured-grid files of ADER-DG and finite-volume solutions.

* ``aderdg_cartesian_vtu``: every cell's solution at its (N+1)^3 equidistant
  vertices (sampled on the device, exa_grid_sample_cartesian) on N^3
  hexahedra per cell -- the vertex variant of the reference's
  ``vtu::Cartesian::vertices`` plotter (E/plotters/VTK/ADERDG2CartesianVTK.cpp).
* ``aderdg_legendre_vtu``: the nodal values on the Gauss-Legendre lattice of
  every cell (``vtu::Legendre::vertices``, E/plotters/VTK/ADERDG2LegendreVTK.cpp).
* ``fv_cells_vtu``: the interior volumes of every finite-volume patch as
  cells with cell data (E/plotters/VTK/FiniteVolumes2VTK.cpp).

Each cell (patch) writes its own points and sub-cells, as the reference's
patch writers do.  Cells are numbered x fastest, as the grids store them.
The reference writes through Peano's VTK writers (not vendored); these files
follow the VTK XML format (UnstructuredGrid, ASCII or base64 binary) that the
same readers consume.
"""
import base64
import struct

import numpy as np

VTK_QUAD = 9
VTK_HEXAHEDRON = 12


def _array_xml(name, data, ncomp, binary, dtype):
    data = np.ascontiguousarray(data, dtype=dtype).ravel()
    tname = {np.float64: "Float64", np.int64: "Int64", np.uint8: "UInt8"}[dtype]
    head = '<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="%s">' % (
        tname, name, ncomp, "binary" if binary else "ascii")
    if binary:  # inline binary: base64 of a 4-byte length header followed by the raw data
        raw = data.tobytes()
        body = base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")
    elif dtype == np.float64:
        body = " ".join(repr(float(x)) for x in data)
    else:
        body = " ".join(str(int(x)) for x in data)
    return head + body + "</DataArray>"


def write_vtu(path, points, cells, cell_type, point_data=None, cell_data=None, time=None, binary=False):
    """points [npoints][3]; cells [ncells][nvertices] (point indices);
    point_data / cell_data {name: [n][ncomp]}."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(cells, dtype=np.int64)
    ncells, nv = cells.shape
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">',
             "<UnstructuredGrid>"]
    if time is not None:
        parts.append("<FieldData>" + _array_xml("TIME", [time], 1, binary, np.float64) + "</FieldData>")
    parts.append('<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (len(points), ncells))
    parts.append("<Points>" + _array_xml("Points", points, 3, binary, np.float64) + "</Points>")
    parts.append("<Cells>" + _array_xml("connectivity", cells, 1, binary, np.int64)
                 + _array_xml("offsets", np.arange(1, ncells + 1) * nv, 1, binary, np.int64)
                 + _array_xml("types", np.full(ncells, cell_type), 1, binary, np.uint8) + "</Cells>")
    for tag, data in (("PointData", point_data), ("CellData", cell_data)):
        if data:
            parts.append("<%s>" % tag + "".join(
                _array_xml(k, v, np.asarray(v).reshape(len(v), -1).shape[1], binary, np.float64)
                for k, v in data.items()) + "</%s>" % tag)
    parts += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(parts) + "\n")


def _lattice(nx, ny, nz, offset, dx, coords):
    """Points of every cell's tensor lattice (1D coordinates `coords` in [0, 1],
    x fastest) and the hexahedra joining neighbouring lattice points."""
    m = len(coords)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cz, cy, cx = np.meshgrid(coords, coords, coords, indexing="ij")
    px = offset[0] + (ix.ravel()[:, None] + cx.ravel()[None, :]) * dx[0]
    py = offset[1] + (iy.ravel()[:, None] + cy.ravel()[None, :]) * dx[1]
    pz = offset[2] + (iz.ravel()[:, None] + cz.ravel()[None, :]) * dx[2]
    points = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(m - 1), np.arange(m - 1), np.arange(m - 1), indexing="ij")
    base = ((k * m + j) * m + i).ravel()
    corners = [0, 1, m + 1, m, m * m, m * m + 1, m * m + m + 1, m * m + m]
    hexes = base[:, None] + np.array(corners)[None, :]
    ncell = nx * ny * nz
    cells = (np.arange(ncell)[:, None, None] * m ** 3 + hexes[None, :, :]).reshape(-1, 8)
    return points, cells


def aderdg_cartesian_vtu(path, grid, time=None, binary=False, name="Q"):
    """Cartesian-vertices plot of an ADER-DG grid (exahype_amd.Grid)."""
    N = grid.solver.order
    dx = [grid.width[d] / n for d, n in enumerate((grid.nx, grid.ny, grid.nz))]
    points, cells = _lattice(grid.nx, grid.ny, grid.nz, grid.offset, dx, np.arange(N + 1) / N)
    values = grid.sample_cartesian().reshape(-1, grid.solver.nvar)
    write_vtu(path, points, cells, VTK_HEXAHEDRON, point_data={name: values}, time=time, binary=binary)


def aderdg_legendre_vtu(path, grid, time=None, binary=False, name="Q"):
    """The nodal values on every cell's Gauss-Legendre lattice."""
    from .tables import dg_math
    N = grid.solver.order
    dx = [grid.width[d] / n for d, n in enumerate((grid.nx, grid.ny, grid.nz))]
    xi = np.asarray(dg_math.gl_nodes_weights(N)[0])
    points, cells = _lattice(grid.nx, grid.ny, grid.nz, grid.offset, dx, xi)
    values = grid.download().reshape(-1, grid.solver.nvar)
    write_vtu(path, points, cells, VTK_HEXAHEDRON, point_data={name: values}, time=time, binary=binary)


def fv_cells_vtu(path, fvgrid, time=None, binary=False, name="Q"):
    """The interior volumes of every finite-volume patch as cells (2D: quads
    in the z = offset plane; 3D: hexahedra)."""
    s = fvgrid.solver
    P, G, dim = s.patch, s.ghost, s.dim
    side = P + 2 * G
    luh = fvgrid.download().reshape((fvgrid.npatch,) + (side,) * dim + (5,))
    inner = luh[(slice(None),) + (slice(G, G + P),) * dim]
    nx, ny, nz = fvgrid.nx, fvgrid.ny, (fvgrid.nz if dim == 3 else 1)
    dx = [fvgrid.width[0] / nx, fvgrid.width[1] / ny, (fvgrid.width[2] / nz if dim == 3 else 0.0)]
    coords = np.arange(P + 1) / P
    if dim == 3:
        points, cells = _lattice(nx, ny, nz, fvgrid.offset, dx, coords)
        write_vtu(path, points, cells, VTK_HEXAHEDRON, cell_data={name: inner.reshape(-1, 5)}, time=time,
                  binary=binary)
        return
    m = P + 1
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    cy, cx = np.meshgrid(coords, coords, indexing="ij")
    px = fvgrid.offset[0] + (ix.ravel()[:, None] + cx.ravel()[None, :]) * dx[0]
    py = fvgrid.offset[1] + (iy.ravel()[:, None] + cy.ravel()[None, :]) * dx[1]
    points = np.stack([px.ravel(), py.ravel(), np.full(px.size, fvgrid.offset[2])], axis=1)
    j, i = np.meshgrid(np.arange(P), np.arange(P), indexing="ij")
    base = (j * m + i).ravel()
    quads = base[:, None] + np.array([0, 1, m + 1, m])[None, :]
    cells = (np.arange(nx * ny)[:, None, None] * m * m + quads[None, :, :]).reshape(-1, 4)
    write_vtu(path, points, cells, VTK_QUAD, cell_data={name: inner.reshape(-1, 5)}, time=time, binary=binary)
