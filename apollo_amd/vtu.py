"""Per-frame VTU output in the layout Apollo's frames have.

ApSolver::writeData (src/solvers/apsolver.cc:404-419) views the state Vec
with PETSc's VTK viewer: an unstructured grid of the mesh vertices (z = 0)
and its triangles, an Int32 cell array "Rank" first, then one Float64 cell
array per degree of freedom of the cell, component j = node * m + variable
(the order of the cell's section dofs).  Apollo's readers rely on exactly
that: scripts/wxunsdgdata2.py:38-84 takes array j + 1 as dof j and the
first three cone vertices as the element's corners.  Only interior cells
are written (the viewer skips the ghost cells that
DMPlexConstructGhostCells appends).

The file is written as XML with ASCII data arrays, which every
vtkXMLUnstructuredGridReader accepts.
"""
import numpy as np


def frame_name(run_name, frame):
    """`<runName>_<frame>.vtu` (apsolver.cc:409)."""
    return "%s_%d.vtu" % (run_name, frame)


def _array(fh, name, dtype, values, ncomp=1, per_line=6):
    fh.write('        <DataArray type="%s" Name="%s" NumberOfComponents="%d" format="ascii">\n'
             % (dtype, name, ncomp))
    flat = np.asarray(values).ravel()
    fmt = "%d" if dtype.startswith("Int") or dtype.startswith("UInt") else "%.17g"
    for i in range(0, flat.size, per_line):
        fh.write("          " + " ".join(fmt % v for v in flat[i:i + per_line]) + "\n")
    fh.write("        </DataArray>\n")


def write_frame(path, mesh, q, field="dg", rank=0):
    """Write state q [K][Np][m] on `mesh` (VX, VY, EToV) to `path`."""
    q = np.asarray(q, dtype=np.float64)
    K = mesh.EToV.shape[0]
    if q.ndim != 3 or q.shape[0] != K:
        raise ValueError("q must be [K][Np][m] with K = %d" % K)
    ndof = q.shape[1] * q.shape[2]
    dofs = q.reshape(K, ndof)
    nv = len(mesh.VX)
    pts = np.zeros((nv, 3))
    pts[:, 0], pts[:, 1] = mesh.VX, mesh.VY
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n')
        fh.write('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        fh.write("  <UnstructuredGrid>\n")
        fh.write('    <Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (nv, K))
        fh.write("      <Points>\n")
        _array(fh, "Position", "Float64", pts, 3)
        fh.write("      </Points>\n")
        fh.write("      <Cells>\n")
        _array(fh, "connectivity", "Int32", np.asarray(mesh.EToV, dtype=np.int64))
        _array(fh, "offsets", "Int32", 3 * np.arange(1, K + 1))
        _array(fh, "types", "UInt8", np.full(K, 5))          # VTK_TRIANGLE
        fh.write("      </Cells>\n")
        fh.write("      <CellData>\n")
        _array(fh, "Rank", "Int32", np.full(K, rank))
        for j in range(ndof):
            _array(fh, "%s.%d" % (field, j), "Float64", dofs[:, j])
        fh.write("      </CellData>\n")
        fh.write("    </Piece>\n")
        fh.write("  </UnstructuredGrid>\n")
        fh.write("</VTKFile>\n")


def read_frame(path):
    """Read a frame written by write_frame (or any ASCII VTU of that
    layout): returns (points [nv][3], cells [K][3], arrays list in file
    order of (name, values))."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")

    def values(node):
        dt = node.get("type")
        conv = float if dt.startswith("Float") else int
        return np.array([conv(v) for v in node.text.split()])

    pts = values(piece.find("Points/DataArray")).reshape(-1, 3)
    conn = {d.get("Name"): values(d) for d in piece.findall("Cells/DataArray")}
    cells = conn["connectivity"].reshape(-1, 3)
    arrays = [(d.get("Name"), values(d)) for d in piece.findall("CellData/DataArray")]
    return pts, cells, arrays
