"""The C ABI boundary (CPU only): the shared library loads, exports every
entry point include/dmrg_hip.h declares, and the Python binding declares
exactly that set, each with the argument list of its C prototype; the
descriptor structs have the C layout.  No compute call is made here (there
is no GPU)."""

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


def header_prototypes():
    """{name: (return type, [parameter types])} of every entry point, as
    written in include/dmrg_hip.h (comments removed, names dropped)."""
    text = open(os.path.join(ROOT, "include", "dmrg_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    protos = {}
    for ret, name, params in re.findall(r"([A-Za-z_][\w\s\*]*?)\b(dmrg_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ps = [] if params.strip() in ("", "void") else [q.strip() for q in params.split(",")]
        # drop the parameter name: the last identifier of each declaration
        types = [re.sub(r"\s*\b[A-Za-z_]\w*$", "", q) if re.search(r"[\w\*]\s+\**[A-Za-z_]\w*$", q) else q
                 for q in ps]
        protos[name] = (" ".join(ret.split()), [" ".join(t.replace("*", " * ").split()) for t in types])
    return protos


def _kind_of_c(t):
    """Coarse class of a C type: the pointee for pointers, else the scalar."""
    t = t.replace("const ", "").strip()
    if t.endswith("*"):
        base = t[:-1].strip()
        if base.endswith("*"):
            return ("ptr", "ptr")
        return ("ptr", {"double": "double", "int": "int", "long long": "ll", "char": "char",
                        "dmrg_heff_desc": "heff", "dmrg_env_desc": "env", "void": "void",
                        "dmrg_ctx": "ctx"}[base])
    return ("val", {"int": "int", "double": "double", "long long": "ll", "void": "void"}[t])


def _kind_of_ctypes(ct):
    if ct is ctypes.c_void_p:
        return ("ptr", None)                 # any pointer (device or opaque)
    if ct is ctypes.c_char_p:
        return ("ptr", "char")
    scal = {ctypes.c_int: "int", ctypes.c_double: "double", ctypes.c_longlong: "ll"}
    if ct in scal:
        return ("val", scal[ct])
    target = getattr(ct, "_type_", None)
    named = {ctypes.c_double: "double", ctypes.c_int: "int", ctypes.c_longlong: "ll", ctypes.c_void_p: "ptr",
             _lib.HeffDesc: "heff", _lib.EnvDesc: "env"}
    return ("ptr", named[target])


def test_binding_argtypes_match_prototypes():
    """Every entry point's ctypes restype / argtypes agree with its C
    prototype: same count and order; ints, doubles and long longs where
    the header has them; a pointer wherever the header has one, and for
    typed ctypes pointers the same pointee (c_void_p stands for device and
    opaque pointers)."""
    protos = header_prototypes()
    assert sorted(protos) == header_functions()
    for name, (res, args) in _lib.SIGNATURES.items():
        ret, params = protos[name]
        assert len(args) == len(params), (name, len(args), params)
        for i, (ct, t) in enumerate(zip(args, params)):
            kc, kt = _kind_of_ctypes(ct), _kind_of_c(t)
            assert kc[0] == kt[0], (name, i, t, ct)
            if kc[1] is not None:
                assert kc[1] == kt[1], (name, i, t, ct)
        kr = _kind_of_c(ret)
        assert _kind_of_ctypes(res) == kr or (kr[0] == "ptr" and _kind_of_ctypes(res)[0] == "ptr"), (name, ret)


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
