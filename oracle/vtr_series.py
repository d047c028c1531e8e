"""ORACLE -- test infrastructure only.

Restatement of the reference's VTR time series for one rank, the checker for
the library's writer (amca_sim_enable_vtr).  Only tests/ load it.

VtrManager.cpp:186-206 (headers), :212-341 (cell data), :347-384 (coordinates),
:390-398 (closing), :404-503 (.pvtr); VtkManager.cpp:90-132 (.pvd), :188-260
(names).  Frame n writes grainsTimeStates/timeStep_<n:05d>/<name>00000.vtr and
<name><n:05d>.pvtr; the .pvd lists every frame so far.  Cell rows run over
(k, j) with i along the row; doubles print as C's "%g"; the temperature prints
truncated to int; the .pvd time stamps print fixed with eight decimals.
"""
import numpy as np

W2, W3, W4, W5, W6 = " " * 2, " " * 4, " " * 6, " " * 8, " " * 10
NAMES = ["temperature", "temperature_rate", "grain_ID", "grain_orientation", "grain_velocity",
         "element_birth", "melted_flag", "phase_ID", "alpha_phase", "phase_orientation",
         "velocity", "debug", "debugMC"]


def _g(x):
    return "%g" % float(x)


def variable_ids(names):
    ids = [NAMES.index(v) for v in names if v in NAMES]
    return ids or [2]


def vtr_text(nx, ny, nz, x0, y0, z0, h, ids, orient, euler3, T=None, melted=None):
    shape = (nz + 2, ny + 2, nx + 2)
    ori = np.asarray(orient).reshape(shape)
    eu = np.asarray(euler3, np.float64).reshape(-1, 3)
    ext = "0 %d 0 %d 0 %d" % (nx, ny, nz)
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n',
           W2 + '<RectilinearGrid WholeExtent="%s">\n' % ext,
           W3 + '<Piece Extent="%s">\n' % ext,
           W4 + '<CellData Scalars="grain_id">\n']

    def rows(cell):
        for k in range(1, nz + 1):
            for j in range(1, ny + 1):
                out.append(W6 + "".join(cell(i, j, k) for i in range(1, nx + 1)) + "\n")

    for v in ids:
        if v == 0 and T is not None:
            t = np.asarray(T).reshape(shape)
            out.append(W5 + '<DataArray type="Float64" Name="cell_temperature" format="ascii">\n')
            rows(lambda i, j, k: "%d " % int(t[k, j, i]))
        elif v == 2:
            out.append(W5 + '<DataArray type="Int64" Name="grain_id" format="ascii">\n')
            rows(lambda i, j, k: "%d " % ori[k, j, i])
        elif v == 3:
            out.append(W5 + '<DataArray type="Float64" Name="grain_orientation" '
                            'NumberOfComponents="3" format="ascii">\n')
            rows(lambda i, j, k: ("%s %s %s " % tuple(_g(a) for a in eu[ori[k, j, i]]))
                 if ori[k, j, i] >= 0 else "0 0 0 ")
        elif v == 6 and melted is not None:
            m = np.asarray(melted).reshape(shape)
            out.append(W5 + '<DataArray type="Int8" Name="melted_flag" format="ascii">\n')
            rows(lambda i, j, k: "%d " % m[k, j, i])
        else:
            continue
        out.append("\n" + W5 + "</DataArray>\n")
    out.append(W4 + "</CellData>\n" + W4 + "<Coordinates>\n")
    for a, n, o in (("X", nx, x0), ("Y", ny, y0), ("Z", nz, z0)):
        out.append(W5 + '<DataArray type="Float64" name="%s_Coordinates" format="ascii">\n' % a)
        out.append(W6 + "".join(_g(o + q * h) + " " for q in range(n + 1)))
        out.append(W5 + "</DataArray>\n")
    out.append(W4 + "</Coordinates>\n\n" + W3 + "</Piece>\n" + W2 + "</RectilinearGrid>\n"
               "</VTKFile>\n")
    return "".join(out)


def pvtr_text(nx, ny, nz, ids, name):
    ext = "0 %d 0 %d 0 %d" % (nx, ny, nz)
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PRectilinearGrid" version="0.1" byte_order="LittleEndian">\n',
           W2 + '<PRectilinearGrid  WholeExtent="%s" \n' % ext,
           " " * 18 + 'GhostLevel = "0">\n',
           W3 + '<PCellData Scalars="grain_id">\n']
    decl = {0: ['<PDataArray type="Float64" Name="cell_temperature"  format="ascii"/>'],
            2: ['<PDataArray type="Int64" Name="grain_id" format="ascii"/>'],
            3: ['<PDataArray type="Float64" Name="grain_orientation" NumberOfComponents="3" '
                'format="ascii"/>'],
            6: ['<PDataArray type="Int8" Name="melted_flag" format="ascii"/>'],
            11: ['<PDataArray type="Int8" Name="growth_flag"  format="ascii"/>',
                 '<PDataArray type="Float64" Name="cooling"  format="ascii"/>'],
            12: ['<PDataArray type="Float64" Name="%s"  format="ascii"/>' % n
                 for n in ("p", "tMCS", "vf_a", "vf_b", "Hardness")]}
    for v in ids:
        out += [W4 + d + "\n" for d in decl.get(v, [])]
    out.append(W3 + "</PCellData>\n" + W3 + "<PCoordinates>\n")
    out += [W4 + '<PDataArray type="Float64" format="ascii"/>\n'] * 3
    out.append(W3 + "</PCoordinates>\n")
    out.append(W3 + '<Piece Extent="%s"\n' % ext)
    out.append(W3 + "      " + 'Source="%s00000.vtr" />\n' % name)
    out.append(W2 + "</PRectilinearGrid>\n</VTKFile>\n")
    return "".join(out)


def pvd_text(times, name):
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n',
           W2 + "<Collection>\n"]
    for c, t in enumerate(times):
        out.append(W3 + '<DataSet timestep="%.8f" group="" part="0" '
                        'file="./grainsTimeStates/timeStep_%05d/%s%05d.pvtr"/>\n' % (t, c, name, c))
    out.append(W2 + "</Collection>\n</VTKFile>\n")
    return "".join(out)
