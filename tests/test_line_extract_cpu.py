"""line_extract (dsp-weno_amd/line_extract.py) on hand-made files in the
solver's VTK layout (output.cc:40-80): probe points, the cell a point on a
face takes, points outside the grid, systems and scalar files, and the text
output of the command line (enflo/utils/line_extract.py:1-99)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import line_extract as LE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_vtk(path, nx, ny, nz, h, origin, fields, time=0.25):
    L = ["# vtk DataFile Version 3.0", "enflo", "ASCII", "DATASET STRUCTURED_POINTS",
         "FIELD FieldData 1", "TIME 1 1 double", f"{time}",
         f"DIMENSIONS {nx + 1} {ny + 1} {nz + 1}", f"SPACING {h[0]} {h[1]} {h[2]} ",
         f"ORIGIN {origin[0]} {origin[1]} {origin[2]} ", f"CELL_DATA  {nx * ny * nz}"]
    for name, v in fields:
        if v.ndim == 2:
            L.append(f"VECTORS {name} float")
            L += [" ".join(repr(float(c)) for c in row) for row in v]
        else:
            L += [f"SCALARS {name} float 1", "LOOKUP_TABLE default"]
            L += [repr(float(c)) for c in v]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def test_systems_file_cells_along_the_line(tmp_path):
    nx, ny, nz = 8, 4, 2
    n = nx * ny * nz
    ids = np.arange(n, dtype=np.float64)
    vel = np.stack([1000 + ids, 2000 + ids, 3000 + ids], axis=1)
    p = tmp_path / "sol0003.vtk"
    # x in [0, 1] (8 cells), y in [0, 1] (4 cells), z in [0, 1] (2 cells)
    write_vtk(p, nx, ny, nz, (0.125, 0.25, 0.5), (0.0, 0.0, 0.0),
              [("density", ids), ("velocity", vel), ("pressure", -ids),
               ("physical_entropy", ids * 0)])
    data, names = LE.extract(str(p), 0.0, 1.0, 8)
    assert names == ["density", "velocity", "pressure"]
    # y = 0.5 is the face between cell rows 1 and 2 -> row 2; z = 0.5 -> layer 1
    cells = np.arange(8) + (1 * 4 + 2) * 8
    assert np.array_equal(data[:, 0], np.arange(8) / 8)
    assert np.array_equal(data[:, 1], cells.astype(float))
    assert np.array_equal(data[:, 2], 1000 + cells.astype(float))
    assert np.array_equal(data[:, 3], -cells.astype(float))
    # a line running past the domain: the point at x = 1 exactly takes the
    # last cell, points beyond it give 0
    data, _ = LE.extract(str(p), 0.5, 1.5, 4)
    assert list(data[:, 1]) == [cells[4], cells[6], cells[7], 0.0]
    data, _ = LE.extract(str(p), 1.0, 2.0, 2)
    assert data[0, 1] == cells[7] and data[1, 1] == 0.0


def test_scalar_file_and_command_line(tmp_path):
    # binary-exact spacing: the probe points x = -1 + (i/8)*2 fall on faces
    # (cell = floor((x - origin)/spacing), as the reference's probe filter)
    nx = 8
    u = np.sin(np.arange(nx * 2 * 2))
    p = tmp_path / "sol0000.vtk"
    write_vtk(p, nx, 2, 2, (0.25, 0.5, 0.5), (-1.0, 0.0, 0.0), [("solution", u)])
    q = tmp_path / "sol0001.vtk"
    write_vtk(q, nx, 2, 2, (0.25, 0.5, 0.5), (-1.0, 0.0, 0.0), [("solution", 2 * u)])
    r = subprocess.run([sys.executable, os.path.join(ROOT, "dsp-weno_amd", "line_extract.py"),
                        "-1", "1", "8"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a = np.loadtxt(tmp_path / "sol0000.txt")
    b = np.loadtxt(tmp_path / "sol0001.txt")
    assert a.shape == (8, 2)
    cells = np.arange(8) + (1 * 2 + 1) * 8  # y = 0.5 and z = 0.5 are faces -> upper cells
    assert np.array_equal(a[:, 0], -1 + 0.25 * np.arange(8))
    assert np.array_equal(a[:, 1], u[cells])
    assert np.array_equal(b[:, 1], 2 * u[cells])


def test_usage_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "dsp-weno_amd", "line_extract.py")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "x0,x1,Nx" in r.stdout


def test_rejects_other_datasets(tmp_path):
    p = tmp_path / "bad.vtk"
    p.write_text("# vtk DataFile Version 3.0\nx\nASCII\nDATASET POLYDATA\n")
    with pytest.raises(ValueError):
        LE.read_vtk(str(p))
