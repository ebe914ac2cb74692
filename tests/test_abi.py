"""The C-ABI library loads and exports every entry point include/bertha_amd.h
declares; argument errors come back as status codes (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bertha_amd.h")).read()
    return sorted(set(re.findall(r"\b(bdf_[a-z_]+)\s*\(", text)))


def test_header_declares_the_reference_replacements():
    names = declared_symbols()
    for n in ("bdf_create", "bdf_coulomb", "bdf_coulomb_device", "bdf_eri_batch", "bdf_etable",
              "bdf_destroy", "bdf_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(bdf):
    lib = bdf.library()
    for n in declared_symbols():
        assert hasattr(lib, n), n


def test_null_arguments_return_einval(bdf):
    lib = bdf.library()
    assert lib.bdf_create(None, None, None, None, None) == bdf.BDF_EINVAL
    assert b"null" in lib.bdf_last_error()
    assert lib.bdf_coulomb(None, 3, None, None, 0, None, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_coulomb_device(None, 3, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_coulomb1(None, 0, None, None, 0, None, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_coulomb1_device(None, 0, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_racah_info(None, 0, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_breit1(None, 0, None, None, 0, None, None, None, None) == bdf.BDF_EINVAL
    assert lib.bdf_breit1_device(None, 0, None, None, None) == bdf.BDF_EINVAL
    lib.bdf_destroy(None)


def test_fortran_interface_module_names_every_entry_point():
    path = os.path.join(ROOT, "include", "bertha_amd_mod.f90")
    text = open(path).read().lower()
    for n in declared_symbols():
        assert f"name='{n}'" in text or f'name="{n}"' in text, n


def test_schwarz_mode_names(bdf):
    """Engine(schwarz=...) maps onto bdf_options.schwarz and refuses the rest."""
    from importlib import import_module
    m = import_module(bdf.__name__)
    assert [m._schwarz_mode(v) for v in (False, None, 0, True, 1, "reference", 2, "rigorous")] == \
        [0, 0, 0, 1, 1, 1, 2, 2]
    for bad in (3, -1, "strict", 1.5):
        with pytest.raises(ValueError):
            m._schwarz_mode(bad)
