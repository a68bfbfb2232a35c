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
    assert lib.pyl_abi_version() == 107
    # An invalid argument is rejected on the host before anything touches the GPU.
    rc = lib.pyl_ffa_resolve(0, None, None, None, None, 1, 1.0, 64, None, None, None)
    assert rc == _lib.PYL_ERR_INVALID
    assert b"nparams" in lib.pyl_last_error()


def test_noalloc_level_rejects_oversized_levels_on_the_host():
    """The no-allocation FFA level refuses a level beyond its fixed table before it looks
    for a device."""
    import ctypes

    import numpy as np

    lib = _lib.load()
    cmax = int(lib.pyl_ffa_level_noalloc_max_coords())
    counts = np.array([3, cmax], dtype=np.int64)
    p = counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    rc = lib.pyl_ffa_level_noalloc(2, None, p, p, None, 1, 1.0, 64, None, 2, None, None)
    assert rc == _lib.PYL_ERR_INVALID
    assert b"maximum without allocation" in lib.pyl_last_error()


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the C compiler's size and field offsets."""
    import shutil
    import subprocess

    from pyloki_amd import prune
    from pyloki_amd.simulation.modulate import PylModulation

    gcc = shutil.which("gcc")
    if gcc is None:
        import pytest

        pytest.skip("gcc not available")
    structs = {"pyl_prune_params": prune.PruneParams, "pyl_prune_coords": prune.PruneCoords,
               "pyl_circ_params": prune.CircParams, "pyl_circ_level": prune.CircLevel,
               "pyl_ep_loop": prune.EpLoop, "pyl_ep_budget": prune.EpBudget,
               "pyl_circ_loop": prune.CircLoop, "pyl_ep_forest": prune.EpForest,
               "pyl_modulation": PylModulation}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pyloki_hip.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run([gcc, "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = structs[cname]
        got = ctypes_size(py) if field == "size" else getattr(py, field).offset
        assert got == int(val), (cname, field, got, val)


def ctypes_size(t) -> int:
    import ctypes

    return ctypes.sizeof(t)


def test_sim_task_dtype_matches_the_header(tmp_path):
    """The numpy record of PylSimTask has the C compiler's size and field offsets."""
    import shutil
    import subprocess

    import pytest

    from pyloki_amd.detection.thresholding import TASK_DTYPE

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pyloki_hip.h"',
             "int main(void) {", 'printf("size %zu\\n", sizeof(PylSimTask));']
    for name in TASK_DTYPE.names:
        lines.append(f'printf("{name} %zu\\n", offsetof(PylSimTask, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "task.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "task"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run([gcc, "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        field, val = line.split()
        got = TASK_DTYPE.itemsize if field == "size" else TASK_DTYPE.fields[field][1]
        assert got == int(val), (field, got, val)
