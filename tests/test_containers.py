"""Container files in the reference's formats and their averaging, against the
reference's own files (fixture files_and_averages: the reference engine wrote the
input files with calculate_*(ofile=...) and averaged them with AverageFromFilesText /
AverageFromFilesBinary)."""
import os

import numpy as np
import pytest

from conftest import golden


def _write(tmp_path, name, arr):
    p = os.path.join(str(tmp_path), name)
    with open(p, "wb") as f:
        f.write(bytes(np.asarray(arr, dtype=np.uint8)))
    return p


def _read(p):
    with open(p, "rb") as f:
        return f.read()


@pytest.fixture(scope="module")
def g():
    return golden("files_and_averages")


def test_writers_reproduce_reference_text_bytes(g, tmp_path):
    from srsoriginal_amd import containers
    # the averaged container, written by our writers, equals the reference's files byte for byte
    c = g["avg_text_container"]
    p = os.path.join(str(tmp_path), "avg.txt")
    containers.write_scalar_text(p, c[:, :3], c[:, 3], 2)
    assert _read(p) == bytes(g["avg_txt"])
    pb = os.path.join(str(tmp_path), "avg.bin")
    containers.write_scalar_binary(pb, c[:, :3], c[:, 3], 2)
    assert _read(pb) == bytes(g["avg_bin"])
    s = g["avg_spectrum"]
    ps = os.path.join(str(tmp_path), "savg.txt")
    containers.write_spectrum_text(ps, s[:, 0], s[:, 1])
    assert _read(ps) == bytes(g["savg_txt"])
    with pytest.raises(ValueError, match="binary spectrum"):
        containers.write_spectrum_binary(ps, s[:, 0], s[:, 1])


def test_average_flux_text_and_binary_match_reference(g, tmp_path):
    from srsoriginal_amd.sr import sr
    osr = sr()
    f1 = _write(tmp_path, "f1.txt", g["f1_txt"])
    f2 = _write(tmp_path, "f2.txt", g["f2_txt"])
    out_t = os.path.join(str(tmp_path), "o.txt")
    out_b = os.path.join(str(tmp_path), "o.bin")
    lst = osr.average_flux(ifiles=[f1, f2], ofile=out_t, bofile=out_b)
    ref = g["avg_text_container"]
    got = np.array([[p[0], p[1], p[2], v] for p, v in lst])
    assert np.array_equal(got, ref)  # Kahan average of parsed text, bit for bit
    assert _read(out_t) == bytes(g["avg_txt"]) and _read(out_b) == bytes(g["avg_bin"])
    b1 = _write(tmp_path, "a.bin", g["avg_bin"])
    b2 = _write(tmp_path, "a2.bin", g["avg2_bin"])
    out_bt = os.path.join(str(tmp_path), "ob.txt")
    lstb = osr.average_power_density(bifiles=[b1, b2], ofile=out_bt)
    gotb = np.array([[p[0], p[1], p[2], v] for p, v in lstb])
    assert np.array_equal(gotb, g["avg_binary_container"])
    assert _read(out_bt) == bytes(g["avgb_txt"])


def test_average_spectra_matches_reference(g, tmp_path):
    from srsoriginal_amd.sr import sr
    osr = sr()
    s1 = _write(tmp_path, "s1.txt", g["s1_txt"])
    s2 = _write(tmp_path, "s2.txt", g["s2_txt"])
    out = os.path.join(str(tmp_path), "s.txt")
    lst = osr.average_spectra(ifiles=[s1, s2], ofile=out)
    assert np.array_equal(np.array(lst), g["avg_spectrum"])
    assert _read(out) == bytes(g["savg_txt"])
    assert osr.average_spectra(bifiles=[s1]) == []  # the reference does not read binary spectra


def test_average_validation(tmp_path, g):
    from srsoriginal_amd.sr import sr
    osr = sr()
    f1 = _write(tmp_path, "f1.txt", g["f1_txt"])
    with pytest.raises(ValueError, match="not both"):
        osr.average_flux(ifiles=[f1], bifiles=[f1])
    with pytest.raises(ValueError, match="No files given"):
        osr.average_flux()
    lines = bytes(g["f1_txt"]).split(b"\n")
    short = _write(tmp_path, "short.txt", np.frombuffer(b"\n".join(lines[:5]) + b"\n", dtype=np.uint8))
    with pytest.raises(ValueError, match="fewer points"):
        osr.average_flux(ifiles=[f1, short])
