"""A Fortran program reaches the engine through include/bertha_amd_mod.f90,
as the reference's HFSCF would at its CLMFAST call site (bertha-scf.f:5401):
the module compiles with flang, links libbertha_amd.so, agrees with the C
header on the layout of bdf_system/bdf_options and on error codes, and (on
the GPU) a COMPLEX*16 build with leading dimension NDIM + 3 matches the
oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
PKG = os.path.join(ROOT, "bertha-program_amd")


@pytest.fixture(scope="module")
def caller(tmp_path_factory, bdf):
    if not (os.path.exists(FLANG) or shutil.which(FLANG)):
        pytest.skip("no Fortran compiler")
    bdf.library()  # raises when the engine is not built
    d = tmp_path_factory.mktemp("fortran")
    exe = str(d / "bdf_caller")
    subprocess.run([FLANG, "-module-dir", str(d), "-o", exe,
                    os.path.join(ROOT, "include", "bertha_amd_mod.f90"),
                    os.path.join(ROOT, "tests", "fortran", "bdf_caller.f90"),
                    f"-L{PKG}", "-lbertha_amd", f"-Wl,-rpath,{PKG}"],
                   check=True, cwd=str(d), capture_output=True)
    return exe


def _run(exe, *args):
    p = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_fortran_module_layout_and_error_codes(caller, bdf):
    import ctypes
    from importlib import import_module
    m = import_module(bdf.__name__)
    out = {}
    for line in _run(caller, "abi").splitlines():
        f = line.split()
        if len(f) >= 2 and f[0] != "Fortran" and f[0] != "IEEE":
            out[f[0]] = [int(v) for v in f[1:]]
    assert out["sizes"] == [ctypes.sizeof(m._System), ctypes.sizeof(m._Options)]
    for key in ("coulomb_null_engine", "launch_info_null_engine", "create_bad_hamiltonian",
                "create_bad_precision"):
        assert out[key] == [bdf.BDF_EINVAL], key


@pytest.mark.gpu
def test_fortran_caller_matches_oracle(caller, bdf, orc, tmp_path):
    s = bdf.systems.ne_atom()
    n = s.ndim
    d = bdf.systems.random_density(n, seed=21, real=True)
    np.asfortranarray(d).T.tofile(str(tmp_path / "dens.bin"))  # column-major stream
    print("[fortran] Ne 12s8p, ld = ndim + 3", flush=True)
    assert "ne ok" in _run(caller, "ne", str(tmp_path))
    raw = np.fromfile(str(tmp_path / "gx.bin"), dtype=np.complex128)
    g = raw[:n * n].reshape(n, n).T
    x = raw[n * n:].reshape(n, n).T
    go, xo = orc.Oracle(s).clmfast(d, 3)
    assert np.abs(g - go).max() < 1e-10
    assert np.abs(x - xo).max() < 1e-10
