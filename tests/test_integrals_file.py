"""The reference's integrals file (and.integrals.bin; src/lib_Stokes.f90:2742-2872,
src/lib_int.f90:136-221, src/lib_util.f90:201-287) written and read by the library's host code.
Pinned by the Fortran runtime itself: oracle/integrals_check (an independent Fortran restatement of
the reference's writer and reader) reads the library's files and writes files the library reads.
Column lengths cover one short record, exactly one full 8192-value record (then an empty one) and
more than one record. CPU only: no device is involved."""
import os
import subprocess

import numpy as np
import pytest

from andromeda_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "integrals_check")
SHAPES = [(7, 5, 4), (8192, 3, 2), (8192 * 2 + 5, 2, 3)]
ORDER = [0, 3, 5, 1, 2, 4]  # file order xx, yy, zz, xy, xz, yz as indices of xx, xy, xz, yy, yz, zz


def _pattern(k, nr, nc):  # matrix k (1-based, file order): 1e6 k + 1e3 i + j, i and j 1-based
    i = np.arange(1, nr + 1)[:, None]
    j = np.arange(1, nc + 1)[None, :]
    return 1.0e6 * k + 1.0e3 * i + j


def _stokes_pattern(nsp, nn, nq):
    T, G = [None] * 6, [None] * 6
    for f, a in enumerate(ORDER):
        T[a] = _pattern(f + 1, nsp, nn)
        G[a] = _pattern(f + 7, nsp, nq)
    prT = [_pattern(13 + k, nsp, nn) for k in range(3)]
    prG = [_pattern(16 + k, nsp, nq) for k in range(3)]
    return T, G, prT, prG


def _check_tool():
    if not os.path.exists(CHECK):
        pytest.skip("oracle/integrals_check not built (no Fortran compiler)")


@pytest.mark.parametrize("shape", SHAPES)
def test_stokes_round_trip_and_fortran_reads_it(tmp_path, shape):
    nsp, nn, nq = shape
    T, G, prT, prG = _stokes_pattern(nsp, nn, nq)
    path = str(tmp_path / "and.integrals.bin")
    api.integrals_write_stokes(path, T, G, prT, prG, err=(0.25, 0.5))
    T2, G2, prT2, prG2, e = api.integrals_read_stokes(path, nsp, nn, nq)
    for a, b in zip(T + G + prT + prG, T2 + G2 + prT2 + prG2):
        assert np.array_equal(a, b)
    assert np.array_equal(e, [0.25, 0.5])
    _check_tool()
    out = subprocess.run([CHECK, "read", "stokes", path, str(nsp), str(nn), str(nq)], capture_output=True, text=True)
    assert out.stdout.strip() == "OK", out.stdout + out.stderr


@pytest.mark.parametrize("shape", SHAPES)
def test_stokes_reads_fortran_file(tmp_path, shape):
    _check_tool()
    nsp, nn, nq = shape
    path = str(tmp_path / "and.integrals.bin")
    out = subprocess.run([CHECK, "write", "stokes", path, str(nsp), str(nn), str(nq)], capture_output=True, text=True)
    assert out.stdout.strip() == "WRITTEN", out.stderr
    T, G, prT, prG, e = api.integrals_read_stokes(path, nsp, nn, nq)
    Te, Ge, prTe, prGe = _stokes_pattern(nsp, nn, nq)
    for a, b in zip(T + G + prT + prG, Te + Ge + prTe + prGe):
        assert np.array_equal(a, b)
    assert np.array_equal(e, [0.25, 0.5])
    # the reference's size check: a file of other sizes is refused
    with pytest.raises(api.AndromedaError):
        api.integrals_read_stokes(path, nsp, nn + 1, nq)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_laplace_both_ways(tmp_path, shape):
    nsp, nn, nq = shape
    H, G = _pattern(1, nsp, nn), _pattern(2, nsp, nq)
    path = str(tmp_path / "and.integrals.bin")
    api.integrals_write_laplace(path, H, G, err=(0.25, 0.5))
    H2, G2, e = api.integrals_read_laplace(path, nsp, nn, nq)
    assert np.array_equal(H, H2) and np.array_equal(G, G2) and np.array_equal(e, [0.25, 0.5])
    with pytest.raises(api.AndromedaError):  # a Laplace file has no Stokes id record
        api.integrals_read_stokes(path, nsp, nn, nq)
    _check_tool()
    out = subprocess.run([CHECK, "read", "laplace", path, str(nsp), str(nn), str(nq)], capture_output=True, text=True)
    assert out.stdout.strip() == "OK", out.stdout + out.stderr
    path2 = str(tmp_path / "fortran.bin")
    subprocess.run([CHECK, "write", "laplace", path2, str(nsp), str(nn), str(nq)], check=True, capture_output=True)
    H3, G3, _ = api.integrals_read_laplace(path2, nsp, nn, nq)
    assert np.array_equal(H3, H) and np.array_equal(G3, G)


def test_stokes_file_byte_layout(tmp_path):
    """The first records spelled out: 4-byte markers around the id, the sizes and the errors."""
    nsp, nn, nq = 3, 2, 2
    T, G, prT, prG = _stokes_pattern(nsp, nn, nq)
    path = str(tmp_path / "x.bin")
    api.integrals_write_stokes(path, T, G, prT, prG, err=(0.25, 0.5))
    raw = open(path, "rb").read()
    i4 = np.frombuffer(raw[:32], dtype="<i4")
    assert list(i4[:3]) == [4, 1407, 4] and list(i4[3:8]) == [12, 3, 2, 2, 12]
    assert np.frombuffer(raw[36:52], dtype="<f8").tolist() == [0.25, 0.5]
    # 18 matrices, one record per column: 4 + 8 nsp + 4 bytes each
    assert len(raw) == 12 + 20 + 24 + (12 * nn + 6 * nq) * 0 + (9 * nn + 9 * nq) * (8 + 8 * nsp)
