"""Wavefunction files (WFNFL, bertha-scf.f:5578-5583 written, :597-604 read):
NDIM list-directed records EIGN(I), (COEF(J,I), J=1,NDIM).  scf.write_wfn and
scf.read_wfn are checked against each other, against a Fortran runtime doing
the reference's own READ and WRITE statements (flang, when present), and
against hand-written list-directed text."""
import importlib
import os
import subprocess

import numpy as np
import pytest

FLANG = "/opt/rocm/llvm/bin/flang"

# the reference's READ and WRITE of WFNFL, as a program: read NDIM records
# from the first file, write them to the second
RW_F = """      PROGRAM RW
      IMPLICIT NONE
      INTEGER NDIM, I, J
      DOUBLE PRECISION EIGN(64)
      COMPLEX*16 COEF(64,64)
      CHARACTER*256 FIN, FOUT, ARG
      CALL GET_COMMAND_ARGUMENT(1, ARG)
      READ(ARG, *) NDIM
      CALL GET_COMMAND_ARGUMENT(2, FIN)
      CALL GET_COMMAND_ARGUMENT(3, FOUT)
      OPEN(UNIT=8, FILE=FIN, STATUS='UNKNOWN')
      REWIND(UNIT=8)
      DO I=1,NDIM
        READ(8, *) EIGN(I),(COEF(J,I),J=1,NDIM)
      ENDDO
      CLOSE(UNIT=8)
      OPEN(UNIT=8, FILE=FOUT, STATUS='UNKNOWN')
      REWIND(UNIT=8)
      DO I=1,NDIM
        WRITE(8, *) EIGN(I),(COEF(J,I),J=1,NDIM)
      ENDDO
      CLOSE(UNIT=8)
      END
"""


@pytest.fixture(scope="module")
def scf(bdf):
    return importlib.import_module(bdf.__name__ + ".scf")


def _wavefunction(n, seed=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    w = rng.standard_normal(n) * 1e3
    c[0, 0] = 1e-300        # extremes of the exponent range
    c[1, 1] = -2.5e200j
    c[2, 0] = -0.0
    return w, c


def test_round_trip_is_exact(scf, tmp_path):
    w, c = _wavefunction(9)
    p = str(tmp_path / "a.wfn")
    scf.write_wfn(p, w, c)
    w2, c2 = scf.read_wfn(p, 9)
    assert np.array_equal(w, w2) and np.array_equal(c, c2)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
def test_fortran_runtime_reads_and_writes_the_same_file(scf, tmp_path):
    """Our file through the reference's READ, its WRITE back through ours:
    the same numbers (Fortran's list-directed output wraps lines and splits
    complex constants over them)."""
    src = tmp_path / "rw.f"
    src.write_text(RW_F)
    exe = str(tmp_path / "rw")
    subprocess.run([FLANG, "-O1", "-o", exe, str(src)], check=True, capture_output=True)
    n = 11
    w, c = _wavefunction(n, seed=3)
    a, b = str(tmp_path / "a.wfn"), str(tmp_path / "b.wfn")
    scf.write_wfn(a, w, c)
    subprocess.run([exe, str(n), a, b], check=True, capture_output=True)
    assert sum(1 for _ in open(b)) > n  # records span several lines
    w2, c2 = scf.read_wfn(b, n)
    assert np.array_equal(w, w2) and np.array_equal(c, c2)


def test_list_directed_text(scf, tmp_path):
    """D exponents, commas and blanks as separators, a record over three lines
    with a complex constant split between two, and the rest of a record's
    last line ignored (the next READ starts a new record)."""
    text = ("-1.5D+01, (1.0D0,0.0D0) (2.5,-1.25E-1)\n"
            " 3.0E2 (0.5 ,\n -0.5) , (1.,\n"
            "2.) 99.0\n")
    p = tmp_path / "t.wfn"
    p.write_text(text)
    w, c = scf.read_wfn(str(p), 2)
    assert np.array_equal(w, [-15.0, 300.0])
    assert np.array_equal(c, np.array([[1.0, 0.5 - 0.5j], [2.5 - 0.125j, 1.0 + 2.0j]]))


def test_truncated_and_malformed_files(scf, tmp_path):
    p = tmp_path / "t.wfn"
    p.write_text("1.0 (1.0,0.0) (2.0,0.0)\n")
    with pytest.raises(ValueError, match="ends inside record 2"):
        scf.read_wfn(str(p), 2)
    p.write_text("(1.0,0.0) (1.0,0.0)\n")
    with pytest.raises(ValueError, match="starts with a complex"):
        scf.read_wfn(str(p), 1)
    p.write_text("1.0 (x,0.0)\n")
    with pytest.raises(ValueError, match="not a number"):
        scf.read_wfn(str(p), 1)


def test_density_from_coefficients(scf):
    import torch
    w, c = _wavefunction(6, seed=5)
    d = scf.density_from_coefficients(torch.from_numpy(c), 3, 2).numpy()
    occ = c[:, 3:5]
    assert np.allclose(d, occ.conj() @ occ.T, rtol=1e-14, atol=1e-14)
