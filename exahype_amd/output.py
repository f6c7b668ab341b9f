"""VTK XML unstructured-grid files of ADER-DG and finite-volume solutions.

* ``aderdg_cartesian_vtu``: every cell's solution at its (N+1)^3 equidistant
  vertices (sampled on the device, exa_grid_sample_cartesian) on N^3
  hexahedra per cell -- the vertex variant of the reference's
  ``vtu::Cartesian::vertices`` plotter (E/plotters/VTK/ADERDG2CartesianVTK.cpp).
* ``aderdg_legendre_vtu``: the nodal values on the Gauss-Legendre lattice of
  every cell (``vtu::Legendre::vertices``, E/plotters/VTK/ADERDG2LegendreVTK.cpp).
* ``fv_cells_vtu``: the interior volumes of every finite-volume patch as
  cells with cell data (E/plotters/VTK/FiniteVolumes2VTK.cpp).
* ``aderdg_cartesian_peano`` / ``write_peano_patch``: Peano patch files (see
  the section below).

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


# ---------------------------------------------------------------------------
# Peano patch files (the reference's Peano::Cartesian::{vertices,cells}::ascii
# plotters, E/plotters/ADERDG2CartesianPeanoPatchFileFormat.cpp:136-372).
# Those plotters hand every cell to Peano's text patch writer as one patch
# (its offset and size) with the values at the patch's (N+1)^d equidistant
# vertices (x fastest) or N^d sub-cell centres, a "Q" dataset of the written
# unknowns and a "time" dataset holding the cell's time stamp at every
# vertex / sub-cell.  The writer itself belongs to Peano, which the reference
# does not vendor; the layout below is a plain-text rendering of that
# structure: a header naming the format, dimension and patch size, the
# declaration of every dataset, then one block per patch.

def write_peano_patch(path, offsets, sizes, patch_size, datasets, kind="vertex", dims=3):
    """Writes a Peano text patch file.  offsets, sizes: (npatch, dims);
    patch_size: sub-cells per direction; datasets: {name: array of shape
    (npatch, points per patch, unknowns)} with points = (patch_size+1)^dims
    (kind "vertex") or patch_size^dims (kind "cell"), x fastest."""
    offsets, sizes = np.asarray(offsets, float), np.asarray(sizes, float)
    npatch = offsets.shape[0]
    npts = (patch_size + 1) ** dims if kind == "vertex" else patch_size ** dims
    tag = "%s-values" % kind
    lines = ["# ", "# Peano patch file ", "# Version 0.1 ", "# ", "format ASCII ", "",
             "dimensions %d " % dims, "patch-size " + " ".join([str(patch_size)] * dims) + " ", ""]
    data = {}
    for name, arr in datasets.items():
        a = np.asarray(arr, float).reshape(npatch, npts, -1)
        data[name] = a
        lines += ['begin %s "%s"' % (tag, name), "  number-of-unknowns %d" % a.shape[2], "end %s" % tag, ""]
    for p in range(npatch):
        lines.append("begin patch")
        lines.append("  offset " + " ".join(repr(float(x)) for x in offsets[p]))
        lines.append("  size " + " ".join(repr(float(x)) for x in sizes[p]))
        for name, a in data.items():
            lines.append('  begin %s "%s"' % (tag, name))
            lines.append("    " + " ".join(repr(float(x)) for x in a[p].ravel()))
            lines.append("  end %s" % tag)
        lines.append("end patch")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_peano_patch(path):
    """Parses a file written by write_peano_patch: dict with dims,
    patch_size, kind, unknowns {name: n}, offsets, sizes and the datasets
    {name: (npatch, points, unknowns)}."""
    dims = ps = None
    kind = None
    unknowns, offsets, sizes, values = {}, [], [], {}
    cur, in_patch = None, False
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            words = line.split()
            if words[0] == "dimensions":
                dims = int(words[1])
            elif words[0] == "patch-size":
                ps = int(words[1])
            elif words[0] == "begin" and words[1] == "patch":
                in_patch = True
            elif words[0] == "end" and words[1] == "patch":
                in_patch = False
            elif words[0] == "begin":
                kind = words[1].split("-")[0]
                cur = words[2].strip('"')
                if in_patch:
                    values.setdefault(cur, [])
            elif words[0] == "end":
                cur = None
            elif words[0] == "number-of-unknowns":
                unknowns[cur] = int(words[1])
            elif words[0] == "offset":
                offsets.append([float(x) for x in words[1:]])
            elif words[0] == "size":
                sizes.append([float(x) for x in words[1:]])
            elif cur is not None and in_patch:
                values[cur].append([float(x) for x in words])
    npts = (ps + 1) ** dims if kind == "vertex" else ps ** dims
    data = {k: np.asarray(v).reshape(len(v), npts, unknowns[k]) for k, v in values.items()}
    return dict(dims=dims, patch_size=ps, kind=kind, unknowns=unknowns, offsets=np.asarray(offsets),
                sizes=np.asarray(sizes), data=data)


def aderdg_cartesian_peano(path, grid, time=0.0, name="Q"):
    """Peano::Cartesian::vertices::ascii plot of an ADER-DG grid: every cell
    a patch of N^3 sub-cells with the solution at its (N+1)^3 equidistant
    vertices (sampled on the device) and the time stamp at every vertex."""
    N = grid.solver.order
    n = (grid.nx, grid.ny, grid.nz)
    dx = np.array([grid.width[d] / n[d] for d in range(3)])
    iz, iy, ix = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    offsets = np.asarray(grid.offset)[None, :] + idx * dx[None, :]
    sizes = np.tile(dx, (len(offsets), 1))
    q = grid.sample_cartesian()
    t = np.full(q.shape[:2] + (1,), float(time))
    write_peano_patch(path, offsets, sizes, N, {name: q, "time": t}, kind="vertex", dims=3)
