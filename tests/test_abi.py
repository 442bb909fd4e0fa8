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


def _header_struct_fields(cname):
    text = open(os.path.join(ROOT, "include", "dmrg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct " + cname + r"\s*\{(.*?)\}\s*" + cname + ";", text, flags=re.S).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = re.sub(r"^(const\s+)?[a-z_ ]+?\s*\**\s*(?=[A-Za-z_][A-Za-z0-9_]*\s*(,|$))", "", decl)
        names += [n.strip().lstrip("*").strip() for n in decl.split(",")]
    return names


def test_documented_heff_desc_matches_binding_and_header():
    """The HeffDesc snippet in INTEGRATION.md is what a maintainer copies:
    its fields, in order, must be those of _lib.HeffDesc and of struct
    dmrg_heff_desc in include/dmrg_hip.h (a short struct would let the
    library read nproj / proj past its end)."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"class HeffDesc\(ctypes\.Structure\):.*?_fields_ = \[(.*?)\]\n", doc, flags=re.S).group(1)
    documented = re.findall(r'\("([A-Za-z0-9_]+)", ctypes\.(c_[a-z_]+)\)', block)
    binding = [(n, t.__name__) for n, t in _lib.HeffDesc._fields_]
    assert [n for n, _ in documented] == [n for n, _ in binding]
    assert [t for _, t in documented] == [t for _, t in binding]
    assert [n for n, _ in binding] == _header_struct_fields("dmrg_heff_desc")
    assert [n for n, _ in _lib.EnvDesc._fields_] == _header_struct_fields("dmrg_env_desc")
