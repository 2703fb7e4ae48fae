"""The C-ABI library loads on the CPU and exports every symbol include/mrlbm.h
declares (no compute calls: this container has no GPU)."""
import os
import re

from mrlbm_2d_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "mrlbm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*|mrlbm_handle \*)\s*(mrlbm_\w+)\s*\(", txt, flags=re.M)))


def test_header_and_binding_agree():
    assert sorted(capi.EXPORTS) == declared_symbols()


def test_library_exports_every_symbol():
    L = capi.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.mrlbm_abi_version() == 1


def test_default_params_match_reference_input():
    p = capi.default_params()
    # input.dat:35-45 defaults
    assert p.y_periodic == 1 and p.x_periodic == 0 and p.rotated_coordinate == 1
    assert p.vel_interp == 1 and p.mom_interp == 1 and p.mass_bc == capi.MASS_STRONG
    assert p.precision == capi.MRLBM_FP64


def test_create_reports_bad_arguments_without_gpu():
    p = capi.default_params()
    p.nx, p.ny = 2, 2
    try:
        capi.Lattice(p)
    except capi.MrlbmError as e:
        assert e.code == capi.E_ARG
    else:
        raise AssertionError("expected an error")
