"""The Julia binding written out in INTEGRATION.md mirrors the C structs of
include/simplediffeq_hip.h field for field: same names, same order, same kinds
(Int32 = int32_t, Int64 = int64_t, Float64 = double, Ptr = any pointer).  A binding
whose struct is shorter than the C one lets the library read or write past it."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simplediffeq_hip.h")
BINDING = os.path.join(ROOT, "INTEGRATION.md")

PAIRS = [("sdeh_solve_desc", "SolveDesc"), ("sdeh_io", "SolveIO"), ("sdeh_stats", "Stats")]


def c_fields(src, name):
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, re.S)
    assert m, name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        mt = re.match(r"(const )?(int32_t|int64_t|double|float|void|char)\s*(\**)\s*(.*)", decl)
        assert mt, decl
        base, stars, names = mt.group(2), mt.group(3), mt.group(4)
        for nm in names.split(","):
            nm = nm.strip()
            ptr = stars or nm.startswith("*")
            kind = "ptr" if ptr else {"int32_t": "i32", "int64_t": "i64", "double": "f64"}[base]
            out.append((nm.lstrip("*").strip(), kind))
    return out


def jl_fields(src, name):
    m = re.search(r"mutable struct %s\b[^\n]*\n(.*?)\nend" % name, src, re.S)
    assert m, name
    out = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if "= new(" in line:  # the constructor ends the field list
            break
        if not line:
            continue
        for f in line.split(";"):
            f = f.strip()
            if not f:
                continue
            nm, ty = f.split("::")
            kind = "ptr" if ty.startswith("Ptr") else {"Int32": "i32", "Cint": "i32",
                                                        "Int64": "i64", "Float64": "f64"}[ty]
            out.append((nm.strip(), kind))
    return out


@pytest.mark.parametrize("cname,jname", PAIRS)
def test_julia_mirror_matches_header(cname, jname):
    c = c_fields(open(HEADER).read(), cname)
    j = jl_fields(open(BINDING).read(), jname)
    assert j == c


def test_constructors_fill_every_field():
    src = open(BINDING).read()
    for jname in ("SolveIO", "Stats"):
        m = re.search(r"%s\(\) = new\((.*?)\)\n" % jname, src, re.S)
        assert m, jname
        args = [a for a in m.group(1).replace("\n", " ").split(",") if a.strip()]
        assert len(args) == len(jl_fields(src, jname)), jname
