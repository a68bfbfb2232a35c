"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""

from __future__ import annotations

import re
from pathlib import Path

from pyloki_amd import _lib

HEADERS = sorted((Path(__file__).resolve().parent.parent / "include").glob("*.h"))


def declared_symbols() -> set[str]:
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(pyl_[a-z0-9_]+)\s*\(", text))
    return names


def test_headers_exist():
    assert HEADERS


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = declared_symbols()
    assert declared, "no declarations parsed"
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared but not exported"


def test_bindings_cover_declarations():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_symbols() == bound


def test_abi_version_and_error_path_without_gpu():
    lib = _lib.load()
    assert lib.pyl_abi_version() == 100
    # An invalid argument is rejected on the host before anything touches the GPU.
    rc = lib.pyl_ffa_resolve(0, None, None, None, None, 1, 1.0, 64, None, None, None)
    assert rc == _lib.PYL_ERR_INVALID
    assert b"nparams" in lib.pyl_last_error()
