"""The C ABI boundary (CPU only): the shared library loads, exports every
entry point include/dmrg_hip.h declares, and the Python binding declares
exactly that set.  No compute call is made here (there is no GPU)."""

import ctypes
import os
import re

from dmrgpy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "dmrg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmrg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name


def test_binding_matches_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_struct_layouts():
    """ctypes descriptors against the C layout of include/dmrg_hip.h, as gcc
    lays it out (offset of every field and the struct size)."""
    import os
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    checks = []
    for cname, py in (("dmrg_heff_desc", _lib.HeffDesc), ("dmrg_env_desc", _lib.EnvDesc)):
        checks.append('printf("%s sizeof %zu\\n", sizeof({0}));'.replace("%s", cname).format(cname))
        for f, _ in py._fields_:
            checks.append('printf("{0} {1} %zu\\n", offsetof({0}, {1}));'.format(cname, f))
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"dmrg_hip.h\"\nint main(void) {\n" + \
        "\n".join(checks) + "\nreturn 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "layout.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "layout")
        subprocess.run(["gcc", "-I", os.path.join(root, "include"), os.path.join(d, "layout.c"), "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in (("dmrg_heff_desc", _lib.HeffDesc), ("dmrg_env_desc", _lib.EnvDesc)):
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_version_and_error_are_callable_without_gpu():
    lib = _lib.load()
    assert lib.dmrg_version() >= 100
    assert lib.dmrg_last_error() is not None
