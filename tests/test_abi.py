"""The C-ABI library loads and exports every symbol the header declares."""
import ctypes
import re

from conftest import ROOT


def header_symbols():
    txt = (ROOT / "include" / "peregrine_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pg):
    lib = pg.abi.load()
    names = header_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(pg.abi.SIGNATURES), set(names) ^ set(pg.abi.SIGNATURES)


def test_abi_version_and_error_string(pg):
    lib = pg.abi.load()
    assert lib.pg_abi_version() == 1
    assert b"gfx950" in lib.pg_version()
    # argument errors are reported without touching a device
    rc = lib.pg_solver_create(None, None, None, None, None, 0, None, 0, None)
    assert rc == pg.abi.PG_ERR_ARG
    assert lib.pg_last_error()


def test_struct_layouts_match_header(pg, orc):
    L = orc.lib()
    assert L.orc_sizeof(0) == ctypes.sizeof(pg.abi.PgGrid) == 72
    assert L.orc_sizeof(1) == ctypes.sizeof(pg.abi.PgModel) == 184
    assert L.orc_sizeof(2) == ctypes.sizeof(pg.abi.PgParams) == 72
