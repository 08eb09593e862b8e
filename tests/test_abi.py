"""The C-ABI boundary without a GPU: the header compiles as C and C++, its record sizes match the ctypes mirror,
and the built library loads and exports every entry point the header declares."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from x2amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "x2dev.h")
ENTRY = r"\b(x2(?:dev|batch)?_\w+)\s*\("  # the lane's, the batch's and the host calls' (x2_grid_*, x2_adaptive_masks ...)


def declared_entry_points():
    text = "".join(open(os.path.join(ROOT, "include", f)).read() for f in sorted(os.listdir(os.path.join(ROOT, "include")))
                   if f.endswith(".h"))
    return sorted(set(re.findall(ENTRY, text)))


@pytest.mark.parametrize("compiler,ext", [("gcc", "c"), ("g++", "cpp")])
def test_header_compiles_and_sizes_match(compiler, ext):
    names = list(abi.EXPECTED_SIZES)
    body = "\n".join(f'printf("%zu\\n", sizeof({n}));' for n in names)
    src = f'#include <stdio.h>\n#include "{HEADER}"\nint main(void){{ {body} return 0; }}\n'
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t." + ext)
        exe = os.path.join(d, "t")
        open(path, "w").write(src)
        subprocess.run([compiler, "-Wall", "-Werror", "-o", exe, path], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    for name, size in zip(names, out):
        assert int(size) == abi.EXPECTED_SIZES[name][1], name


def test_library_exports_every_declared_entry_point():
    lib = abi.load()
    declared = declared_entry_points()
    assert len(declared) >= 70
    for name in declared:
        assert hasattr(lib, name), name
        assert name in abi.SIGNATURES, f"{name} has no ctypes signature"
    nm = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], check=True, capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (x2\w+)", nm))
    assert set(declared) <= exported


def test_null_arguments_fail_cleanly_without_a_device():
    """Argument checks run before any HIP call: null pointers give X2_ERR_ARGUMENT and a message."""
    lib = abi.load()
    assert lib.x2dev_create(None, None) == -1
    assert b"null" in lib.x2dev_last_error()
    assert lib.x2dev_ready(None) == 0
    assert lib.x2dev_in_flight(None) == 0
    assert lib.x2dev_poll(None, None) == -1
    assert lib.x2dev_launch(None, None, 0, 0) == -1
    assert lib.x2dev_download_film(None) == -1
    assert lib.x2dev_film_ready(None) == 0
    lib.x2dev_destroy(None)
