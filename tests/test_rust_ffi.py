"""The Rust FFI crate (mcrt_amd/rust/mcrt-hip-sys) agrees with the C header.

No Rust toolchain is in this image, so the crate is checked as text: every
function declared in include/mcrt_abi.h appears in the crate's extern block
with the same parameter and return types, nothing else does, every Pod
struct has the same fields, and the constants have the same values.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcrt_abi.h")
CRATE = os.path.join(ROOT, "mcrt_amd", "rust", "mcrt-hip-sys")

C_SCALARS = {"int": "c_int", "uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "float": "f32",
             "double": "f64", "uint8_t": "u8", "uint16_t": "u16", "char": "c_char", "void": "c_void",
             "mcrt_ctx": "McrtCtx", "mcrt_denoiser": "McrtDenoiser"}


def _strip_c(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_type(decl):
    """'const float *x' -> '*const f32'; 'mcrt_ctx **out' -> '*mut *mut McrtCtx'."""
    decl = " ".join(decl.split())
    m = re.match(r"^(const\s+)?(\w+)\s*((?:\*\s*)*)\s*(\w+)?$", decl)
    assert m, decl
    const, base, stars = m.group(1), m.group(2), m.group(3).replace(" ", "")
    t = C_SCALARS.get(base, base)
    n = len(stars)
    for k in range(n):
        # the innermost pointer carries the const of the pointee
        t = ("*const " if (k == 0 and const) else "*mut ") + t
    return t


def header_functions():
    t = _strip_c(open(HEADER).read())
    out = {}
    for ret, name, params in re.findall(r"^\s*([\w\s\*]+?)\b(mcrt_\w+)\s*\(([^;{]*?)\)\s*;", t, flags=re.M):
        ps = [p.strip() for p in params.split(",")]
        ps = [] if ps == ["void"] else ps
        r = " ".join(ret.split())
        out[name] = ([c_type(p) for p in ps], None if r == "void" else c_type(r + " x"))
    return out


def rust_functions():
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    block = src[src.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    block = re.sub(r"//[^\n]*", "", block)
    out = {}
    for name, params, ret in re.findall(r"pub fn (mcrt_\w+)\s*\(([^)]*)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        ps = []
        for p in [x.strip() for x in params.split(",") if x.strip()]:
            _, ty = p.split(":", 1)
            ps.append(" ".join(ty.split()))
        out[name] = (ps, " ".join(ret.split()) if ret else None)
    return out


def test_crate_files_exist():
    for f in ("Cargo.toml", "build.rs", os.path.join("src", "lib.rs")):
        assert os.path.exists(os.path.join(CRATE, f)), f
    toml = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'links = "mcrt_hip"' in toml
    assert "rustc-link-lib=dylib=mcrt_hip" in open(os.path.join(CRATE, "build.rs")).read()


def test_every_header_function_is_bound_with_the_same_types():
    h, r = header_functions(), rust_functions()
    assert len(h) >= 35
    assert sorted(h) == sorted(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name in h:
        assert h[name] == r[name], (name, h[name], r[name])


def test_library_exports_what_the_crate_binds():
    from mcrt_amd import _lib

    lib = _lib.load()
    for name in rust_functions():
        assert hasattr(lib, name), name


def _c_structs():
    t = _strip_c(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct \w+ \{(.*?)\} (\w+);", t, flags=re.S):
        fields = []
        for line in body.strip().split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            m = re.match(r"^(\w+) (\w+)(?:\[(\w+)\])?$", line)
            assert m, line
            ty, fname, n = m.groups()
            ty = C_SCALARS.get(ty, ty)
            n = {"MCRT_MAX_MATERIALS": "64"}.get(n, n)
            fields.append((fname, f"[{ty}; {n}]" if n else ty))
        out[name] = fields
    return out


def _rust_structs():
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    out = {}
    for name, body in re.findall(r"pub struct (\w+) \{(.*?)\n\}", src, flags=re.S):
        fields = []
        for fname, ty in re.findall(r"pub (\w+): ([^,\n]+),", body):
            ty = " ".join(ty.split()).replace("MCRT_MAX_MATERIALS", "64")
            fields.append((fname, ty))
        out[name] = fields
    return out


@pytest.mark.parametrize("name", ["GpuSphere", "GpuBvhNode", "GpuVertex", "GpuTriangle", "GpuMaterial",
                                  "GpuMaterialData", "CameraUniform", "DenoiseParams"])
def test_pod_structs_have_the_same_fields(name):
    assert _c_structs()[name] == _rust_structs()[name]


def test_constants_have_the_same_values():
    h = _strip_c(open(HEADER).read())
    defs = dict(re.findall(r"#define (MCRT_\w+) ([^\s]+)", h))
    enum = dict(re.findall(r"(MCRT_\w+) = (-?\d+)", h))
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    rust = dict(re.findall(r"pub const (MCRT_\w+): \w+ = ([^;]+);", src))

    def num(v):
        v = v.replace("_", "").rstrip("uU")
        if v.lower().startswith("0x"):
            return float(int(v, 16))
        return float(v.rstrip("fF"))

    both = {**{k: v for k, v in defs.items() if k != "MCRT_ABI_H"}, **enum}
    for k, v in both.items():
        assert k in rust, k
        assert num(rust[k]) == num(v), (k, v, rust[k])
