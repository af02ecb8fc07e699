"""The C-ABI library loads, exports every entry point include/dfsph_abi.h declares, and the ctypes structs of the
Python binding have the C compiler's layout (no GPU needed: nothing is computed)."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from alloy_amd.sph import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dfsph_abi.h")


def _declared_functions() -> list[str]:
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfsph_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.dfsph_abi_version() == _abi.ABI_VERSION


def test_python_binding_covers_the_header():
    lib = _abi.load_library()
    for n in _declared_functions():
        fn = getattr(lib, n)
        assert fn.restype is not None or n == "dfsph_last_error", n


@pytest.mark.parametrize("struct", ["DfsphParams", "DfsphLayout", "DfsphModel", "DfsphReport"])
def test_struct_layout_matches_c(struct):
    cls = getattr(_abi, struct)
    fields = [f[0] for f in cls._fields_]
    body = "\n".join(f'  printf("%zu ", offsetof({struct}, {f}));' for f in fields)
    src = f"""#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
{body}
  printf("%zu\\n", sizeof({struct}));
  return 0;
}}"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        exe = os.path.join(d, "p")
        with open(c, "w") as f:
            f.write(src)
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    c_offsets = [int(x) for x in out[:-1]]
    py_offsets = [getattr(cls, f).offset for f in fields]
    assert c_offsets == py_offsets
    assert int(out[-1]) == ctypes.sizeof(cls)


def test_errors_map_to_python_exceptions():
    lib = _abi.load_library()
    rc = lib.dfsph_create(None, None, None, None)
    assert rc == _abi.DFSPH_EINVAL
    with pytest.raises(ValueError, match="null argument"):
        _abi.check(rc, lib, "create")
