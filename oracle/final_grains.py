"""ORACLE -- test infrastructure only.

Restatement of the reference's end-of-run grain files, the checker for the
library's writer (amca_write_final_grains).  Only tests/ load it.

CellularAutomataManager_3D::writeToFile (CellularAutomataManager_3D.cpp:1381-1639):
  * finalGrains.vtk: legacy ASCII rectilinear grid, nodes x0 + i h; cell data
    grain_id (the orientation id, written as an integer under a "double"
    header), number_cells (cells of that id, or the id when negative),
    melted_flag, and EulerAngle (three angles, "0 0 0" without an id);
  * finalResult.txt: a two-line header block, then one row per interior cell
    "id cells melted e1 e2 e3 ".
Doubles are written as a default std::ostream writes them, which is C's "%g"
(six significant digits).  Values go ten to a line with one counter that runs
on from section to section: each section ends its last partial line but keeps
the count (only the z coordinates reset it, :1473), and the coordinate origins
(:1439, :1453, :1466) and the "0 0 0" of a cell without an id (:1586) are not
counted.
"""
import numpy as np


def _g(x):
    return "%g" % float(x)


class _Lines:
    def __init__(self):
        self.parts = []
        self.count = 0

    def put(self, s):
        self.parts.append(s)

    def counted(self):
        self.count += 1
        if self.count % 10 == 0:
            self.parts.append("\n")
            self.count = 0

    def end(self, reset):
        if self.count != 0:
            self.parts.append("\n")
            if reset:
                self.count = 0


def final_grains_text(nx, ny, nz, x0, y0, z0, h, orient, melted, euler3):
    """(finalGrains.vtk, finalResult.txt) contents.  orient and melted are the
    ghost-inclusive arrays ((nz+2, ny+2, nx+2) or flat); melted None = all 1;
    euler3 is (n_orient, 3)."""
    ori = np.asarray(orient).reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].ravel()
    if melted is None:
        mel = np.ones(ori.size, np.int64)
    else:
        mel = np.asarray(melted).reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].ravel()
    eu = np.asarray(euler3, np.float64).reshape(-1, 3)
    cells = {}
    for v in ori:
        if v >= 0:
            cells[int(v)] = cells.get(int(v), 0) + 1

    L = _Lines()
    L.put("# vtk DataFile Version 2.0\nTest Data\nASCII\nDATASET RECTILINEAR_GRID\n")
    L.put("DIMENSIONS %d %d %d\n" % (nx + 1, ny + 1, nz + 1))
    for name, n, o in (("X", nx, x0), ("Y", ny, y0), ("Z", nz, z0)):
        L.put("%s_COORDINATES %d double\n" % (name, n + 1))
        L.put(_g(o) + " ")
        for q in range(n):
            L.put(_g(o + (q + 1) * h) + " ")
            L.counted()
        L.end(reset=(name == "Z"))
        L.put("\n")
    L.put("CELL_DATA %d\nSCALARS grain_id double 1\nLOOKUP_TABLE default\n" % (nx * ny * nz))
    for v in ori:
        L.put("%d " % v)
        L.counted()
    L.end(reset=False)
    L.put("SCALARS number_cells int 1\nLOOKUP_TABLE default\n")
    for v in ori:
        L.put("%d " % (cells[int(v)] if v >= 0 else v))
        L.counted()
    L.end(reset=False)
    L.put("SCALARS melted_flag int 1\nLOOKUP_TABLE default\n")
    for m in mel:
        L.put("%d " % m)
        L.counted()
    L.end(reset=False)
    L.put("VECTORS EulerAngle float\n")
    for v in ori:
        if v >= 0:
            e = eu[int(v)]
            L.put("%s %s %s " % (_g(e[0]), _g(e[1]), _g(e[2])))
            L.counted()
        else:
            L.put("0 0 0 ")
    L.put("\n")
    vtk = "".join(L.parts)

    rows = ["X0  Y0  Z0   Dcell  nx_cell  ny_cell  nz_cell\n",
            "%s %s %s %s %d %d %d\n" % (_g(x0), _g(y0), _g(z0), _g(h), nx, ny, nz),
            "grain_ID  number_cells  melted_flag  EulerAngle1 EulerAngle2 EulerAngle3\n"]
    for v, m in zip(ori, mel):
        if v >= 0:
            e = eu[int(v)]
            rows.append("%d %d %d %s %s %s \n" % (v, cells[int(v)], m, _g(e[0]), _g(e[1]),
                                                  _g(e[2])))
        else:
            rows.append("%d %d %d 0 0 0 \n" % (v, v, m))
    return vtk, "".join(rows)
