"""The C-ABI library loads without a GPU and exports every function that
include/arapgs.h declares; the Python binding table covers the same set.
No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

from arapgs_amd import arapgs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "arapgs.h")


def declared():
    with open(HEADER) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(arapgs_\w+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared()
    assert len(names) >= 40
    for must in ("arapgs_create", "arapgs_knn", "arapgs_gn_solve", "arapgs_six_point_update",
                 "arapgs_render", "arapgs_deform_frame", "arapgs_density_sample"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(arapgs.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(arapgs.EXPORTS) == declared()


def test_version_string_without_gpu():
    lib = ctypes.CDLL(arapgs.LIB_PATH)
    lib.arapgs_version.restype = ctypes.c_char_p
    assert lib.arapgs_version().startswith(b"arapgs")


def test_dropin_program_links_against_the_library():
    """tests/dropin/dropin_check (the C++ caller using include/arapgs_dropin.hpp)
    resolves libarapgs.so and libamdhip64 and starts: without arguments it prints
    its usage and exits 2 before any HIP call."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_check")
    assert os.path.exists(exe), "build the C++ caller first (make -C tests/dropin)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr, (r.returncode, r.stderr)
