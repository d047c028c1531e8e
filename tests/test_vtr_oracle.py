"""Pins oracle/vtr_series.py (the checker of the library's VTR time series)
independently of itself.

- A 2 x 1 x 1 frame with every variable the writer knows, its .pvtr and a
  two-frame .pvd, typed out by hand from the reference's stream statements
  (VtrManager.cpp:186-206 headers, :212-341 cell data, :347-384 coordinates,
  :390-398 closing, :404-503 .pvtr; VtkManager.cpp:90-141 .pvd; indentation
  from CafeMacro.h:54-58, widths from VtkManager.cpp:17-21).
- An XML read-back of a random 5 x 4 x 3 frame: the parsed arrays equal the
  fields in (k, j, i) order, the coordinates are the cell faces.
"""
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import vtr_series as vs  # noqa: E402

# cell (1,1,1): orientation 1 (Euler 1.5 2.25 0.125), 1599.9 K, melted;
# cell (2,1,1): liquid (-1), no material (-1 K), not melted
VTR_2x1x1 = (
    '<?xml version="1.0"?>\n'
    '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
    '  <RectilinearGrid WholeExtent="0 2 0 1 0 1">\n'
    '    <Piece Extent="0 2 0 1 0 1">\n'
    '      <CellData Scalars="grain_id">\n'
    '        <DataArray type="Float64" Name="cell_temperature" format="ascii">\n'
    '          1599 -1 \n'
    '\n'
    '        </DataArray>\n'
    '        <DataArray type="Int64" Name="grain_id" format="ascii">\n'
    '          1 -1 \n'
    '\n'
    '        </DataArray>\n'
    '        <DataArray type="Float64" Name="grain_orientation" NumberOfComponents="3" '
    'format="ascii">\n'
    '          1.5 2.25 0.125 0 0 0 \n'
    '\n'
    '        </DataArray>\n'
    '        <DataArray type="Int8" Name="melted_flag" format="ascii">\n'
    '          1 0 \n'
    '\n'
    '        </DataArray>\n'
    '      </CellData>\n'
    '      <Coordinates>\n'
    '        <DataArray type="Float64" name="X_Coordinates" format="ascii">\n'
    '          5e-07 2.5e-06 4.5e-06         </DataArray>\n'
    '        <DataArray type="Float64" name="Y_Coordinates" format="ascii">\n'
    '          0 2e-06         </DataArray>\n'
    '        <DataArray type="Float64" name="Z_Coordinates" format="ascii">\n'
    '          1e-06 3e-06         </DataArray>\n'
    '      </Coordinates>\n'
    '\n'
    '    </Piece>\n'
    '  </RectilinearGrid>\n'
    '</VTKFile>\n')

PVTR_2x1x1 = (
    '<?xml version="1.0"?>\n'
    '<VTKFile type="PRectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
    '  <PRectilinearGrid  WholeExtent="0 2 0 1 0 1" \n'
    '                  GhostLevel = "0">\n'
    '    <PCellData Scalars="grain_id">\n'
    '      <PDataArray type="Float64" Name="cell_temperature"  format="ascii"/>\n'
    '      <PDataArray type="Int64" Name="grain_id" format="ascii"/>\n'
    '      <PDataArray type="Float64" Name="grain_orientation" NumberOfComponents="3" '
    'format="ascii"/>\n'
    '      <PDataArray type="Int8" Name="melted_flag" format="ascii"/>\n'
    '      <PDataArray type="Int8" Name="growth_flag"  format="ascii"/>\n'
    '      <PDataArray type="Float64" Name="cooling"  format="ascii"/>\n'
    '    </PCellData>\n'
    '    <PCoordinates>\n'
    '      <PDataArray type="Float64" format="ascii"/>\n'
    '      <PDataArray type="Float64" format="ascii"/>\n'
    '      <PDataArray type="Float64" format="ascii"/>\n'
    '    </PCoordinates>\n'
    '    <Piece Extent="0 2 0 1 0 1"\n'
    '          Source="grains00000.vtr" />\n'
    '  </PRectilinearGrid>\n'
    '</VTKFile>\n')

PVD_TWO_FRAMES = (
    '<?xml version="1.0"?>\n'
    '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
    '  <Collection>\n'
    '    <DataSet timestep="0.00000000" group="" part="0" '
    'file="./grainsTimeStates/timeStep_00000/grains00000.pvtr"/>\n'
    '    <DataSet timestep="0.00012500" group="" part="0" '
    'file="./grainsTimeStates/timeStep_00001/grains00001.pvtr"/>\n'
    '  </Collection>\n'
    '</VTKFile>\n')


def _two_cells():
    ori = np.full((3, 3, 4), 0, np.int32)    # (nz+2, ny+2, nx+2), ghosts 0
    ori[1, 1, 1], ori[1, 1, 2] = 1, -1
    T = np.zeros((3, 3, 4))
    T[1, 1, 1], T[1, 1, 2] = 1599.9, -1.0
    mel = np.zeros((3, 3, 4), np.uint8)
    mel[1, 1, 1] = 1
    eu = np.array([[0.1, 0.2, 0.3], [1.5, 2.25, 0.125]])
    return ori, T, mel, eu


def test_hand_written_frame():
    ori, T, mel, eu = _two_cells()
    ids = vs.variable_ids(["temperature", "grain_ID", "grain_orientation", "melted_flag"])
    assert ids == [0, 2, 3, 6]
    got = vs.vtr_text(2, 1, 1, 5e-7, 0.0, 1e-6, 2e-6, ids, ori.ravel(), eu, T.ravel(),
                      mel.ravel())
    assert got == VTR_2x1x1


def test_hand_written_pvtr_and_pvd():
    ids = vs.variable_ids(["temperature", "grain_ID", "grain_orientation", "melted_flag",
                           "debug"])
    assert vs.pvtr_text(2, 1, 1, ids, "grains") == PVTR_2x1x1
    assert vs.pvd_text([0.0, 1.25e-4], "grains") == PVD_TWO_FRAMES


def _floats(text):
    return np.array(text.split(), dtype=np.float64)


def test_frame_read_back():
    """A random 5 x 4 x 3 frame parses as XML; every array holds the fields in
    (k, j, i) order and the coordinates are the cell faces."""
    rng = np.random.default_rng(4)
    nx, ny, nz, h, org = 5, 4, 3, 2.5e-6, (1e-5, -2e-5, 0.0)
    shape = (nz + 2, ny + 2, nx + 2)
    ori = rng.integers(-1, 7, shape).astype(np.int32)
    eu = rng.uniform(0, 3, (7, 3))
    T = rng.uniform(-1, 2000, shape)
    mel = rng.integers(0, 2, shape).astype(np.uint8)
    ids = vs.variable_ids(["temperature", "grain_ID", "grain_orientation", "melted_flag"])
    root = ET.fromstring(vs.vtr_text(nx, ny, nz, *org, h, ids, ori.ravel(), eu, T.ravel(),
                                     mel.ravel()))
    assert root.get("type") == "RectilinearGrid"
    piece = root.find("RectilinearGrid/Piece")
    assert piece.get("Extent") == f"0 {nx} 0 {ny} 0 {nz}"
    arrays = {a.get("Name"): a for a in piece.find("CellData")}
    inner = (slice(1, -1),) * 3
    np.testing.assert_array_equal(_floats(arrays["grain_id"].text), ori[inner].ravel())
    np.testing.assert_array_equal(_floats(arrays["cell_temperature"].text),
                                  np.trunc(T[inner]).ravel())
    np.testing.assert_array_equal(_floats(arrays["melted_flag"].text), mel[inner].ravel())
    o = ori[inner].ravel()
    want = np.where(o[:, None] >= 0, eu[np.maximum(o, 0)], 0.0)
    np.testing.assert_allclose(_floats(arrays["grain_orientation"].text).reshape(-1, 3), want,
                               rtol=5e-6, atol=0)
    coords = piece.find("Coordinates")
    for arr, n, o0 in zip(coords, (nx, ny, nz), org):
        np.testing.assert_allclose(_floats(arr.text), o0 + h * np.arange(n + 1), rtol=5e-6,
                                   atol=1e-12)
