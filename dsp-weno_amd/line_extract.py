#!/usr/bin/env python3
"""line_extract: values of a solution file along the line y = z = 0.5.

    python line_extract.py x0 x1 Nx [solNNNN.vtk]

Same command line, probe points and output as the reference utility
(enflo/utils/line_extract.py:1-99, enflo_scalar/utils/line_extract.py), with
no VTK library: the ASCII STRUCTURED_POINTS files written by `tecno` (and by
the reference, output.cc:10-90) are parsed here directly.

* Without a file argument every `sol*.vtk` of the working directory is read.
* Probe points: Nx + 1 points x0 + (i/Nx)(x1 - x0) at y = z = 0.5 (the
  line source of the reference, resolution Nx); the first Nx are written.
* Cell data are not interpolated: a point takes the value of the cell
  floor((x - origin)/spacing) that contains it (a point exactly on a face
  takes the upper cell, the domain's upper boundary the last cell); a point
  outside the grid gets 0, as the reference's probe filter gives.  The
  solver writes SPACING with 6 significant digits, so a probe point on a face
  can land in the lower cell (1/24 is written 0.0416667) — the same cell the
  reference's VTK probe picks from the same file.
* Output: `<file stem>.txt` (numpy.savetxt) with columns x, density,
  x-velocity, pressure for systems files, x, u for scalar files (field
  `solution`); and `<file stem>.pdf` with the same plots when matplotlib is
  importable.
"""
import glob
import math
import sys

import numpy as np


def read_vtk(path):
    """Parse an ASCII STRUCTURED_POINTS file with CELL_DATA.  Returns
    (dims, spacing, origin, fields) with fields[name] shaped (ncell,) or
    (ncell, 3), cells x fastest."""
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < 4 or "STRUCTURED_POINTS" not in lines[3]:
        raise ValueError(f"{path}: not a STRUCTURED_POINTS file")
    tok = " ".join(lines[4:]).split()
    i = 0
    dims = spacing = origin = None
    ncell = None
    fields = {}
    while i < len(tok):
        t = tok[i]
        if t == "FIELD":
            narr = int(tok[i + 2])
            i += 3
            for _ in range(narr):
                ncomp, ntup = int(tok[i + 1]), int(tok[i + 2])
                i += 4 + ncomp * ntup
        elif t == "DIMENSIONS":
            dims = [int(v) for v in tok[i + 1:i + 4]]
            i += 4
        elif t in ("SPACING", "ASPECT_RATIO"):
            spacing = [float(v) for v in tok[i + 1:i + 4]]
            i += 4
        elif t == "ORIGIN":
            origin = [float(v) for v in tok[i + 1:i + 4]]
            i += 4
        elif t == "CELL_DATA":
            ncell = int(tok[i + 1])
            i += 2
        elif t == "SCALARS":
            name = tok[i + 1]
            ncomp = 1
            i += 3
            if tok[i] != "LOOKUP_TABLE":
                ncomp = int(tok[i])
                i += 1
            i += 2  # LOOKUP_TABLE <name>
            vals = np.array(tok[i:i + ncomp * ncell], dtype=np.float64)
            fields[name] = vals if ncomp == 1 else vals.reshape(ncell, ncomp)
            i += ncomp * ncell
        elif t == "VECTORS":
            name = tok[i + 1]
            i += 3
            fields[name] = np.array(tok[i:i + 3 * ncell], dtype=np.float64).reshape(ncell, 3)
            i += 3 * ncell
        elif t == "POINT_DATA":
            raise ValueError(f"{path}: POINT_DATA files are not written by the solver")
        else:
            i += 1
    if dims is None or spacing is None or origin is None or ncell is None:
        raise ValueError(f"{path}: incomplete header")
    return dims, spacing, origin, fields


def cell_index(x, origin, spacing, npts):
    """Cell of coordinate x along one axis of a grid of npts points, or -1
    outside it."""
    d = (x - origin) / spacing
    nc = npts - 1
    if d < 0.0 or d > nc:
        # tolerance of one part in 1e12 of the extent at either end
        if -1e-12 * nc <= d < 0.0:
            return 0
        if nc < d <= nc * (1.0 + 1e-12):
            return nc - 1
        return -1
    i = int(math.floor(d))
    return min(i, nc - 1)


def probe(dims, spacing, origin, x0, x1, nx, y=0.5, z=0.5):
    """Flat cell indices (-1 outside) of the first nx line points."""
    out = np.empty(nx, dtype=np.int64)
    xs = np.empty(nx)
    for i in range(nx):
        x = x0 + (i / nx) * (x1 - x0)
        xs[i] = x
        c = [cell_index(v, origin[a], spacing[a], dims[a]) for a, v in enumerate((x, y, z))]
        if min(c) < 0:
            out[i] = -1
        else:
            out[i] = (c[2] * (dims[1] - 1) + c[1]) * (dims[0] - 1) + c[0]
    return xs, out


def extract(path, x0, x1, nx):
    dims, spacing, origin, fields = read_vtk(path)
    xs, cells = probe(dims, spacing, origin, x0, x1, nx)
    ok = cells >= 0
    idx = np.where(ok, cells, 0)

    def take(a):
        return np.where(ok, a[idx], 0.0)

    if "solution" in fields:
        cols = [xs, take(fields["solution"])]
        names = ["u"]
    else:
        cols = [xs, take(fields["density"]), take(fields["velocity"][:, 0]), take(fields["pressure"])]
        names = ["density", "velocity", "pressure"]
    return np.stack(cols, axis=1), names


def main(argv):
    if len(argv) < 4:
        print("Please specify x0,x1,Nx, vtkfile(optional).")
        print("Example:")
        print("  line_extract.py 0 1 256 sol200.vtk")
        print("If vtk file is not specified, all files with the name sol*.vtk will be read. ")
        return 1
    x0, x1, nx = float(argv[1]), float(argv[2]), int(argv[3])
    files = [argv[4]] if len(argv) > 4 else sorted(glob.glob("sol*.vtk"))
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        plt = None
    for path in files:
        data, names = extract(path, x0, x1, nx)
        stem = path[:-4]
        np.savetxt(f"{stem}.txt", data)
        if plt is not None:
            fig, axs = plt.subplots(1, len(names), figsize=(20, 5), squeeze=False)
            for k, nm in enumerate(names):
                axs[0][k].plot(data[:, 0], data[:, k + 1])
                axs[0][k].set_xlabel("x")
                axs[0][k].set_ylabel(nm)
            plt.savefig(f"{stem}.pdf", bbox_inches="tight")
            plt.close(fig)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
