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


@pytest.mark.parametrize("struct", ["DfsphParams", "DfsphLayout", "DfsphModel", "DfsphReport", "DfsphEmitters"])
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


def test_mlp_grad_scratch_size_and_rejection():
    """The gradient's scratch covers one partial per 16-sample wave of every 64-sample block, and a smaller scratch
    is refused before anything is launched (the pointers below are never dereferenced)."""
    lib = _abi.load_library()
    P = lib.dfsph_mlp_param_count()
    for n, waves in [(1, 4), (64, 4), (65, 8), (300, 20), (4096, 256)]:
        assert lib.dfsph_mlp_grad_scratch_floats(n) == waves * (P + 1)
    assert lib.dfsph_mlp_grad_scratch_floats(0) == 0
    fake = ctypes.c_void_p(0x1000)
    need = lib.dfsph_mlp_grad_scratch_floats(300)
    rc = lib.dfsph_mlp_grad(fake, fake, fake, 300, 300, fake, need - 1, fake, None)
    assert rc == _abi.DFSPH_EINVAL
