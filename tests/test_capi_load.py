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
    assert L.mrlbm_abi_version() == 2


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


def test_create_rejects_bad_bulk_options_without_gpu():
    for field, value in (("kernel", 3), ("kernel", 305), ("time_blocking", 3), ("time_blocking", 18), ("time_blocking", -2),
                         ("strip_rows", 65), ("precision", 7), ("mass_bc", 5)):
        p = capi.default_params()
        p.nx, p.ny = 64, 32
        setattr(p, field, value)
        try:
            capi.Lattice(p)
        except capi.MrlbmError as e:
            assert e.code == capi.E_ARG, (field, value, e)
        else:
            raise AssertionError(f"{field}={value} accepted")


def test_create_rejects_grids_beyond_32_bit_node_indices_without_gpu():
    """Node indices i + pitch * j (pitch = nx rounded up to 64) are 32-bit: 46341 x 46340
    nodes fit (2,147,441,940), but with the pitch 46400 x 46340 = 2,150,176,000 does not."""
    assert 46341 * 46340 < 2**31 - 1 < 46400 * 46340
    p = capi.default_params()
    p.nx, p.ny = 46341, 46340
    try:
        capi.Lattice(p)
    except capi.MrlbmError as e:
        assert e.code == capi.E_ARG and "32-bit" in str(e), e
    else:
        raise AssertionError("grid accepted")


def test_default_params_pick_the_measured_variants():
    p = capi.default_params()
    assert p.kernel == 0 and p.time_blocking == -1 and p.block_order == 0
