"""CPU checks of the C ABI boundary: the library loads and exports every symbol include/clmi.h
declares, the Python binding binds each of them, and create() rejects bad configurations without
touching a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "clmi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clmi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("clmi_create", "clmi_destroy", "clmi_lf_evolve", "clmi_lf_sync", "clmi_vv_evolve",
                 "clmi_kick", "clmi_drift", "clmi_set_all_averages", "clmi_hubble_constraint",
                 "clmi_gauss_u1", "clmi_spectrum_bins", "clmi_grf_scalar"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "cosmolattice_amd" / "libclmi.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    src = (ROOT / "cosmolattice_amd" / "clmi.py").read_text()
    missing = [n for n in declared_functions() if f'"{n}"' not in src]
    assert not missing, missing


def test_create_rejects_bad_config_without_gpu():
    from cosmolattice_amd import clmi

    lib = clmi.load_library()
    spec = clmi.lphi4(100.0)
    cfg = clmi.make_config(spec, 7, 0.5, 0.01)  # odd N
    h = ctypes.c_void_p()
    rc = lib.clmi_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1
    assert b"N must be even" in lib.clmi_last_error(None)
    cfg = clmi.make_config(spec, 8, 0.5, 0.01)
    cfg.cs_pow[0] = 3
    cfg.n_cs_terms = 1
    rc = lib.clmi_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1


def test_product_path_does_not_import_oracle():
    """The shipped package never imports, loads or links the oracle."""
    pat = re.compile(r"(import\s+oracle|from\s+oracle|liboracle|oracle/|oracle\.py)")
    for p in (ROOT / "cosmolattice_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".hpp", ".h", ".cpp", ".txt"):
            assert not pat.search(p.read_text()), p


def test_binding_constants_match_the_header():
    """Every plain integer #define of include/clmi.h that the Python binding mirrors has the same value
    there (field kinds under the binding's short names, flags, precisions, limits)."""
    from cosmolattice_amd import clmi

    text = (ROOT / "include" / "clmi.h").read_text()
    defines = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define\s+(CLMI_\w+)\s+(-?\d+)\s*$", text, re.M)}
    short = {"CLMI_FLD_S": "FLD_S", "CLMI_PI_S": "PI_S", "CLMI_FLD_CS": "FLD_CS", "CLMI_PI_CS": "PI_CS",
             "CLMI_FLD_U1": "FLD_U1", "CLMI_PI_U1": "PI_U1", "CLMI_MAG_U1": "MAG_U1"}
    checked = 0
    for name, value in defines.items():
        attr = short.get(name, name)
        if hasattr(clmi, attr):
            assert getattr(clmi, attr) == value, name
            checked += 1
    assert checked >= 12
    assert clmi.CLMI_FLAG_MEMMODE(15) == 15 << 8

