"""Container files in the reference's formats, and their averaging.

Text files are written as the reference's containers write them with
std::scientific at the default precision ("%e"): flux / power density rows
"x y v" (dim 2) or "x y z v" (dim 3) (T3DScalarContainer::WriteToFileText,
T3DScalarContainer.cc:322-349), spectrum rows "energy flux"
(TSpectrumContainer::WriteToFileText, TSpectrumContainer.cc:224-259).  Binary
flux / power-density files hold the same rows as native float64
(T3DScalarContainer::WriteToFileBinary, :351-399).  The reference's binary
spectrum writer terminates the program (TSpectrumContainer.cc:264-266); here it
raises ValueError.

Averaging follows AverageFromFilesText / AverageFromFilesBinary
(T3DScalarContainer.cc:66-320, TSpectrumContainer.cc:318-396): the first file
gives the points and v/N, every further file's v/N is added with the containers'
compensated sum (AddToPoint, T3DScalarContainer.cc:31-47; AddToFlux,
TSpectrumContainer.cc:159-175), point by point in file order; reading stops at the
end of the first file, and a later file that ends earlier is an error.
"""
import numpy as np


def _rows(v, cols):
    return np.asarray(v, dtype=np.float64).reshape(-1, cols)


def write_scalar_text(path, xyz, values, dim):
    """Rows of a flux / power-density container; xyz are the container points
    ((x1, x2, 0) for dim 2, the 3-D points for dim 3)."""
    if dim not in (2, 3):
        raise ValueError("incorrect dimensions")
    xyz = _rows(xyz, 3)
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    with open(path, "w") as f:
        for p, val in zip(xyz[:len(v)].tolist(), v.tolist()):
            if dim == 2:
                f.write("%e %e %e\n" % (p[0], p[1], val))
            else:
                f.write("%e %e %e %e\n" % (p[0], p[1], p[2], val))


def write_scalar_binary(path, xyz, values, dim):
    if dim not in (2, 3):
        raise ValueError("incorrect dimensions")
    xyz = _rows(xyz, 3)
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    cols = [xyz[:len(v), 0], xyz[:len(v), 1]] + ([xyz[:len(v), 2]] if dim == 3 else []) + [v]
    np.ascontiguousarray(np.stack(cols, axis=1), dtype="<f8").tofile(path)


def write_spectrum_text(path, energies, flux):
    with open(path, "w") as f:
        for e, v in zip(np.asarray(energies, dtype=np.float64).tolist(), np.asarray(flux, dtype=np.float64).tolist()):
            f.write("%e %e\n" % (e, v))


def write_spectrum_binary(path, energies, flux):
    raise ValueError("binary spectrum files are not supported (the reference's writer terminates, "
                     "TSpectrumContainer.cc:264-266)")


def _read_text(path, cols):
    with open(path) as f:
        tok = f.read().split()
    n = len(tok) // cols  # an incomplete last record is not read, as with operator>>
    return np.array([float(t) for t in tok[:n * cols]], dtype=np.float64).reshape(n, cols)


def _read_binary(path, cols):
    a = np.fromfile(path, dtype="<f8")
    n = len(a) // cols
    return a[:n * cols].reshape(n, cols)


def _average(tables):
    """Compensated average of the last column over files; points from the first."""
    if len(tables) < 1:
        raise ValueError("No files given.  You need at least one file as input in a list.")
    n = len(tables[0])
    for t in tables[1:]:
        if len(t) < n:
            raise ValueError("input files have fewer points than the first one")
    N = float(len(tables))
    s = tables[0][:, -1] / N
    c = np.zeros(n)
    for t in tables[1:]:
        y = t[:n, -1] / N - c
        tt = s + y
        c = (tt - s) - y
        s = tt
    return tables[0][:, :-1].copy(), s


def average_scalar_files(files, binary, dim):
    """(points [n, 3], values [n]) averaged over flux / power-density files."""
    if dim not in (2, 3):
        raise ValueError("incorrect dimensions")
    cols = dim + 1
    read = _read_binary if binary else _read_text
    pts, v = _average([read(f, cols) for f in files])
    xyz = np.zeros((len(v), 3))
    xyz[:, :dim] = pts
    return xyz, v


def average_spectrum_files(files):
    """(energies, flux) averaged over text spectrum files."""
    e, v = _average([_read_text(f, 2) for f in files])
    return e[:, 0], v
