"""The Julia binding (hypercloudcore.jl_amd/julia/HyperCloudHIP.jl) against the C ABI, without a Julia toolchain:
its descriptor structs must have the fields of include/hcc.h's hcc_domain_desc and hcc_stepper_desc in the same order
with the same types and array shapes (a Julia NTuple{n, T} is a C array T[n]; nested tuples are row-major arrays), and
every libhcc symbol it calls must be declared in the header. The ctypes binding of the Python host is checked the same
way, and the byte sizes of the three views agree."""
import ctypes as C
import os
import re

import pytest

from hcc_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "hypercloudcore.jl_amd", "julia", "HyperCloudHIP.jl")
HDR = os.path.join(ROOT, "include", "hcc.h")

SCALARS = {"Int32": "int32_t", "Float64": "double", "UInt64": "uint64_t", "UInt8": "uint8_t"}
SIZES = {"int32_t": 4, "double": 8, "uint64_t": 8, "uint8_t": 1}


def c_structs():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct \{(.*?)\}\s*(\w+);", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            ctype, rest = decl.split(None, 1)
            for d in rest.split(","):
                m = re.match(r"\s*(\w+)((?:\[\d+\])*)\s*$", d)
                assert m, (name, d)
                fields.append((m.group(1), ctype, [int(x) for x in re.findall(r"\[(\d+)\]", m.group(2))]))
        out[name] = fields
    return out


def jl_type(t):
    t = t.strip()
    m = re.match(r"NTuple\{\s*(\d+)\s*,\s*(.*)\}$", t)
    if m:
        ctype, dims = jl_type(m.group(2))
        return ctype, [int(m.group(1))] + dims
    return SCALARS[t], []


def jl_structs():
    text = open(JL).read()
    out = {}
    for name, body in re.findall(r"^struct (\w+)\n(.*?)^end", text, flags=re.S | re.M):
        fields = []
        for line in body.splitlines():
            line = line.split("#")[0].strip()
            if not line:
                continue
            fname, ftype = line.split("::", 1)
            fields.append((fname.strip(),) + jl_type(ftype))
        out[name] = fields
    return out


def ctypes_fields(cls):
    out = []
    for fname, ft in cls._fields_:
        dims = []
        while hasattr(ft, "_length_"):
            dims.append(ft._length_)
            ft = ft._type_
        ctype = {C.c_int32: "int32_t", C.c_double: "double", C.c_uint64: "uint64_t"}[ft]
        out.append((fname, ctype, dims))
    return out


def c_size(fields):
    """sizeof of a C struct of these fields (natural alignment, as on x86-64)."""
    off, align = 0, 1
    for _, ctype, dims in fields:
        a = SIZES[ctype]
        n = a
        for d in dims:
            n *= d
        off = (off + a - 1) // a * a + n
        align = max(align, a)
    return (off + align - 1) // align * align


@pytest.mark.parametrize("jl_name,c_name,py_cls", [("DomainDesc", "hcc_domain_desc", _lib.DomainDesc),
                                                    ("StepperDesc", "hcc_stepper_desc", _lib.StepperDesc)])
def test_descriptor_layouts_match_the_header(jl_name, c_name, py_cls):
    cs, js = c_structs()[c_name], jl_structs()[jl_name]
    assert [f[0] for f in js] == [f[0] for f in cs], "field names or order differ"
    for jf, cf in zip(js, cs):
        assert jf == (cf[0], cf[1], cf[2]), f"{jl_name}.{jf[0]}: {jf} vs {cf}"
    assert ctypes_fields(py_cls) == [tuple(f) for f in cs]
    assert c_size(cs) == C.sizeof(py_cls)


def test_julia_calls_only_declared_symbols():
    text = open(JL).read()
    called = set(re.findall(r"ccall\(\(:(hcc_\w+),", text))
    declared = set(re.findall(r"\b(hcc_[a-z_0-9]+)\s*\(", open(HDR).read()))
    assert called, "no ccall found"
    assert called <= declared, sorted(called - declared)


def test_julia_binding_keeps_the_stepper_contract():
    """The functor writes pg.rhos back unless resident, get_time_step is defined for the stepper's interpolator, and the
    constructor from the reference's objects exists (a text-level check of the definitions the driver loop needs)."""
    text = open(JL).read()
    assert re.search(r"function \(ts::HIPRKTimeStepper\)\(eq, pg, time::Real, dt::Float64\)", text)
    functor = text[text.index("function (ts::HIPRKTimeStepper)(eq, pg"):]
    functor = functor[:functor.index("\nend")]
    assert "download!(ts, pg)" in functor and "upload!(ts, pg)" in functor
    assert re.search(r"function HyperCloudCore\.get_time_step\(pg::HyperCloudCore\.ParticleGrid, eq::HyperCloudCore"
                     r"\.HyperbolicPDE, d::HIPDivergence", text)
    assert "divergence_interpolator::HIPDivergence" in text
    assert re.search(r"function HIPRKTimeStepper\(eq::HC\.HyperbolicPDE, pg::HC\.ParticleGrid, interp::HC\.MUSCL, "
                     r"mood::HC\.MOOD, tableau;", text)
