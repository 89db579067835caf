"""The Rust FFI crate (rust/diomedes-hip-sys/src/lib.rs) against the C header
(include/diomedes_hip.h): every exported function is declared in the crate's
extern block with the same parameters and return type, every header record
has a #[repr(C)] mirror with the same fields, types and array sizes, and the
crate's constants and size assertions carry the header's values.  No Rust
toolchain exists here, so the check is textual: both files are parsed and the
C types mapped to the Rust types the crate must use."""

from __future__ import annotations

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "diomedes_hip.h"
CRATE = ROOT / "rust" / "diomedes-hip-sys"
LIB_RS = CRATE / "src" / "lib.rs"

SCALARS = {"uint32_t": "u32", "uint64_t": "u64", "int": "i32", "float": "f32", "double": "f64",
           "dio_status": "dio_status", "void": "()"}


def _strip_c(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _c_type(decl: str) -> str:
    """A C parameter / return type (qualifiers, pointers) as its Rust form."""
    decl = decl.strip()
    stars = decl.count("*")
    base = decl.replace("*", " ").split()
    const_first = base[0] == "const"
    base = [b for b in base if b != "const"]
    assert len(base) == 1, decl
    name = base[0]
    if name == "char" and stars == 1:
        return "*const c_char"
    if name == "void" and stars > 0:
        name = "c_void"
    rust = SCALARS.get(name, name)
    if stars == 0:
        return rust
    out = rust
    for k in range(stars):
        # the const qualifier of `const T*` applies to the innermost pointee
        out = ("*const " if (k == 0 and const_first) else "*mut ") + out
    return out


def c_functions() -> dict:
    text = _strip_c(HEADER.read_text())
    out = {}
    for m in re.finditer(r"^\s*((?:const\s+)?\w+\s*\**)\s*(dio_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        ret, name, args = m.group(1), m.group(2), " ".join(m.group(3).split())
        params = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                pm = re.match(r"(.*?)(\w+)$", a)
                params.append((pm.group(2), _c_type(pm.group(1))))
        out[name] = (params, _c_type(ret))
    return out


def c_structs() -> dict:
    text = _strip_c(HEADER.read_text())
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} (\w+);", text, flags=re.S):
        assert m.group(1) == m.group(3)
        fields = []
        for line in m.group(2).split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            tm = re.match(r"(\w+) (.*)$", line)
            ctype, names = tm.group(1), tm.group(2)
            for n in names.split(","):
                n = n.strip()
                dims = [int(d) for d in re.findall(r"\[(\d+)\]", n)]
                fname = n.split("[")[0]
                rust = SCALARS.get(ctype, ctype)
                for d in reversed(dims):
                    rust = f"[{rust}; {d}]"
                fields.append((fname, rust))
        out[m.group(1)] = fields
    return out


def c_constants() -> dict:
    text = _strip_c(HEADER.read_text())
    out = {}
    for name, val in re.findall(r"^#define (DIO_\w+) (0x[0-9A-Fa-f]+|\d+)u?$", text, flags=re.M):
        out[name] = int(val, 0)
    for name, val in re.findall(r"\b(DIO_[A-Z_]+) = (\d+)", text):
        out[name] = int(val)
    return out


def _rust_text() -> str:
    text = LIB_RS.read_text()
    return re.sub(r"//[^\n]*", "", text)


def rust_functions() -> dict:
    text = _rust_text()
    block = re.search(r'extern "C" \{(.*?)\n\}', text, flags=re.S).group(1)
    out = {}
    for m in re.finditer(r"pub fn (dio_\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        args = " ".join(m.group(2).split())
        params = []
        if args:
            for a in [x for x in args.split(",") if x.strip()]:
                n, t = a.split(":", 1)
                params.append((n.strip(), " ".join(t.split())))
        out[m.group(1)] = (params, " ".join(m.group(3).split()) if m.group(3) else "()")
    return out


def rust_structs() -> dict:
    text = _rust_text()
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+) \{(.*?)\}", text, flags=re.S):
        fields = []
        for f in m.group(2).split(","):
            f = " ".join(f.split())
            if f:
                fm = re.match(r"pub (\w+): (.*)$", f)
                fields.append((fm.group(1), fm.group(2)))
        out[m.group(1)] = fields
    return out


def rust_constants() -> dict:
    out = {}
    for name, val in re.findall(r"pub const (DIO_\w+): \w+ = (0x[0-9A-Fa-f_]+|\d+);", _rust_text()):
        out[name] = int(val.replace("_", ""), 0)
    return out


def test_header_parses():
    fns, structs = c_functions(), c_structs()
    assert len(fns) >= 35 and "dio_step" in fns and "dio_narrow_pairs_noalloc" in fns
    assert fns["dio_create"] == ([("device", "i32"), ("limits", "*const dio_limits"), ("out", "*mut *mut dio_ctx")],
                                 "dio_status")
    assert fns["dio_last_error"][1] == "*const c_char"
    assert dict(structs["dio_gpu_shape"])["corners"] == "[[f32; 4]; 4]"
    assert [f for f, _ in structs["dio_step_stats"]][-8:-2] == ["ms_dual", "ms_recover", "n_init", "n_primal",
                                                                "n_dual", "n_recover"]


def test_crate_files_exist():
    assert (CRATE / "Cargo.toml").exists() and (CRATE / "build.rs").exists() and LIB_RS.exists()
    cargo = (CRATE / "Cargo.toml").read_text()
    assert 'name = "diomedes-hip-sys"' in cargo and 'links = "diomedes_hip"' in cargo


def test_every_function_matches_the_header():
    c, r = c_functions(), rust_functions()
    assert set(c) == set(r), set(c) ^ set(r)
    for name, (cparams, cret) in c.items():
        rparams, rret = r[name]
        assert [t for _, t in rparams] == [t for _, t in cparams], (name, rparams, cparams)
        assert [n for n, _ in rparams] == [n for n, _ in cparams], (name, "parameter names")
        assert rret == cret, (name, rret, cret)


def test_every_record_matches_the_header():
    c, r = c_structs(), rust_structs()
    assert set(c) <= set(r), set(c) - set(r)
    for name, fields in c.items():
        assert r[name] == fields, (name, r[name], fields)


def test_constants_and_sizes():
    c, r = c_constants(), rust_constants()
    missing = {k for k in c if k not in ("DIO_TUNE_SLOTS",)} - set(r)
    assert not missing, missing
    for k, v in c.items():
        if k in r:
            assert r[k] == v, (k, r[k], v)
    assert r["DIO_TUNE_SLOTS"] == (r["DIO_TUNE_PRIMAL_SHAPES"] + r["DIO_TUNE_BROAD_SHAPES"] + 4
                                   + r["DIO_TUNE_COLOUR_SHAPES"] + 2)
    # the crate asserts each record's size at compile time, with the header's sizes
    from .test_abi import _c_sizeof

    text = _rust_text()
    for name in c_structs():
        m = re.search(rf"size_of::<{name}>\(\) == (\d+)", text)
        assert m, f"no size assertion for {name}"
        assert int(m.group(1)) == _c_sizeof(name), name


@pytest.mark.parametrize("name", ["dio_gpu_body", "dio_avbd_contact", "dio_avbd_constraint", "dio_gpu_contact_out",
                                  "dio_avbd_body_state"])
def test_reference_records_keep_their_reference_names(name):
    """The five records the engine already defines (compute.rs, gpu_physics.rs)
    are re-exported under the engine's names, so a Renderer passes its slices
    unchanged."""
    alias = {"dio_gpu_body": "GpuBody", "dio_avbd_contact": "GpuAvbdContact", "dio_avbd_constraint": "GpuAvbdConstraint",
             "dio_gpu_contact_out": "GpuContactOut", "dio_avbd_body_state": "GpuAvbdBodyState"}[name]
    assert re.search(rf"pub type {alias} = {name};", _rust_text())
