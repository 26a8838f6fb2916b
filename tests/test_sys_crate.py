"""The am1g-sys crate (am1-rs_amd/am1g-sys) mirrors include/am1g.h: every #[repr(C)] struct has
the header's fields in the header's order with matching types, and the extern block declares
exactly the header's functions. The Python ctypes structures are checked the same way. There is
no Rust toolchain here, so the check is on the source text (CPU only)."""
import os
import re

import pytest

import am1g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "am1g.h")
CRATE = os.path.join(ROOT, "am1-rs_amd", "am1g-sys", "src", "lib.rs")

C_TO_RUST = {
    "int": "c_int", "double": "c_double", "long long": "c_longlong", "const uint8_t*": "*const u8",
    "const double*": "*const c_double", "double*": "*mut c_double", "const int*": "*const c_int",
}


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def header_structs():
    text = _strip_c_comments(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const \w+\*?|long long|\w+\*?)\s+(.*)", decl)
            ctype, names = m.group(1), m.group(2)
            for nm in names.split(","):
                nm = nm.strip()
                star = ""
                if nm.startswith("*"):
                    star, nm = "*", nm[1:]
                arr = re.match(r"(\w+)\[(\d+)\]", nm)
                t = ctype + star
                if arr:
                    fields.append((arr.group(1), f"[{C_TO_RUST[t]}; {arr.group(2)}]"))
                else:
                    fields.append((nm, C_TO_RUST[t]))
        out[name] = fields
    return out


def crate_structs():
    text = open(CRATE).read()
    text = text[text.index("pub mod raw {"):text.index('extern "C" {')]  # the raw bindings only
    out = {}
    for name, body in re.findall(r"pub struct (\w+) \{(.*?)\n    \}", text, flags=re.S):
        if "_private" in body:
            continue
        out[name] = [(n, " ".join(t.split())) for n, t in re.findall(r"pub (\w+): ([^,]+),", body)]
    return out


def test_crate_structs_match_the_header():
    h = header_structs()
    c = crate_structs()
    assert set(h) == set(c)
    for name in h:
        assert c[name] == h[name], name


def test_crate_declares_every_header_function():
    text = open(CRATE).read()
    block = text[text.index('extern "C" {'):]
    declared = set(re.findall(r"pub fn (am1g_\w+)\(", block))
    assert declared == set(am1g.EXPORTS)


CTYPES = {"c_int": "c_int", "c_double": "c_double", "c_longlong": "c_longlong"}


@pytest.mark.parametrize("cls,name", [(am1g.Request, "am1g_request"), (am1g.Result, "am1g_result"),
                                      (am1g.DcOptions, "am1g_dc_options"), (am1g.DcResult, "am1g_dc_result"),
                                      (am1g.ServeOptions, "am1g_serve_options"),
                                      (am1g.ServeStats, "am1g_serve_stats")])
def test_python_structures_match_the_header(cls, name):
    h = header_structs()[name]
    got = [f[0] for f in cls._fields_]
    assert got == [f[0] for f in h]
