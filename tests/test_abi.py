"""CPU checks of the C ABI: the HIP library loads and exports every function
that include/kascade_amd.h declares; the Python binding binds all of them.
No compute call is made (no GPU here)."""
import ctypes
import os
import re

import kascade_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "kascade_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["kc_rank_list", "kc_pointer_doubling", "kc_sparse_ruling_set", "kc_root_tree",
                 "kc_rank_via_euler_tour", "kc_rank_lists_segmented", "kc_server_create", "kc_submit"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(kascade_amd.lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(kascade_amd._SIGS)
    kascade_amd.load()


def test_cpp_header_compiles():
    """The C++ drop-in header (reference-shaped spans API) compiles against the C ABI."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "include", "kascade_amd.hpp")
    if not os.path.exists(src):
        return
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "t.cpp")
        open(cpp, "w").write('#include "kascade_amd.hpp"\nint main() { return 0; }\n')
        r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), cpp],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_c_header_is_plain_c():
    """The C ABI header compiles as C11 with warnings as errors (no C++ or
    torch types in the signatures), and a C program links against the library
    with every declared entry point referenced."""
    import subprocess
    import tempfile
    names = declared_functions()
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        refs = "\n".join(f"  (void)&{n};" for n in names)
        open(c, "w").write('#include "kascade_amd.h"\n#include <stddef.h>\n'
                           f"int main(void) {{\n{refs}\n  return 0;\n}}\n")
        exe = os.path.join(d, "t")
        r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic",
                            "-I", os.path.join(ROOT, "include"), c, "-o", exe,
                            "-L", os.path.dirname(kascade_amd.lib_path()), "-lkascade_amd"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
