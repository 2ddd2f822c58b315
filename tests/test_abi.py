"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/psdns_gpu.h and include/psdns_sgs.h, and rejects bad
parameters before it touches the device; the Fortran interface module and
driver are built."""
import ctypes
import pathlib
import re
import subprocess

import pytest

from conftest import REPO, load_package

HEADER = REPO / "include" / "psdns_gpu.h"
SGS_HEADER = REPO / "include" / "psdns_sgs.h"


def declared_functions(header=HEADER):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psdns_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib(P):
    if not P.LIB_PATH.exists():
        P.build()
    return P.load_library()


def test_header_declares_the_boundary():
    fns = declared_functions()
    for f in ("psdns_create", "psdns_stage_nonlinear", "psdns_stage_update", "psdns_put_state",
              "psdns_get_state", "psdns_set_shifts", "psdns_shell_stats", "psdns_last_error"):
        assert f in fns


def test_library_exports_every_declared_symbol(P, lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(P.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (psdns_\w+)", out))
    declared = declared_functions() + declared_functions(SGS_HEADER)
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    for f in declared:
        assert getattr(lib, f) is not None
    assert set(P.EXPORTS) == set(declared_functions())
    assert set(P.SGS_EXPORTS) == set(declared_functions(SGS_HEADER))


def test_sgs_rejects_bad_grids_without_a_device(P, lib):
    h = ctypes.c_void_p()
    assert lib.psdns_sgs_create(48, 0, ctypes.byref(h)) == -1
    assert b"power of two" in lib.psdns_last_error()
    assert lib.psdns_sgs_create(64, -1, ctypes.byref(h)) == -1
    assert lib.psdns_sgs_param_count() == 12164


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=48, ny=48, nz=48), "power of two"),
    (dict(nx=64, ny=32, nz=64), "must be equal"),
    (dict(nc=9), "passive scalars"),
    (dict(nc=2), "Prandtl/Schmidt numbers must be positive"),
    (dict(prec=2), "prec"),
    (dict(mask_kind=1), "mask_kind"),
])
def test_create_rejects_bad_parameters(P, lib, bad, msg):
    kw = dict(nx=64, ny=64, nz=64, nc=0, prec=8, kshift=1, mask_kind=0, viscos=0.01)
    kw.update(bad)
    p = P.Params(**kw)
    h = ctypes.c_void_p()
    rc = lib.psdns_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert h.value is None
    assert msg in lib.psdns_last_error().decode()


def test_null_context_is_an_error(lib):
    assert lib.psdns_phase1(None, 1) == -1
    assert "null context" in lib.psdns_last_error().decode()


def test_params_struct_matches_the_c_layout(P):
    # 7 ints, pad to 8-byte alignment, viscos, pr[8], grad[8][3]
    assert ctypes.sizeof(P.Params) == 32 + 8 + 64 + 192


def test_fortran_interface_and_driver_built(P):
    if not P.DRIVER_PATH.exists():
        P.build()
    assert P.DRIVER_PATH.exists()
    assert (P.PKG / "lib" / "psdns_gpu.mod").exists()
    syms = subprocess.run(["nm", str(P.DRIVER_PATH)], capture_output=True, text=True).stdout
    for f in ("psdns_stage_nonlinear", "psdns_stage_update", "psdns_set_shifts", "psdns_ek_force"):
        assert f in syms
