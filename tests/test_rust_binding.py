"""The Rust binding crate (logosq_amd/rust/logosq-hip-sys) against the C
header it binds, with no Rust toolchain: every function declared in
include/logosq_hip.h has one `extern "C"` declaration in src/lib.rs with the
same parameter count and the Rust type that the C type maps to (pointer
constness included), the constants carry the header's values, LqGate lists
the header's fields in order, and the library exports every declared name."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "logosq_hip.h")
CRATE = os.path.join(ROOT, "logosq_amd", "rust", "logosq-hip-sys")

C_TO_RUST = {
    "void": "c_void", "int": "c_int", "char": "c_char", "long long": "c_longlong", "double": "f64",
    "float": "f32", "size_t": "usize", "uint64_t": "u64", "int64_t": "i64", "int32_t": "i32", "uint8_t": "u8",
    "lq_state": "LqState", "lq_mlp": "LqMlp", "lq_gate": "LqGate", "lq_opt_result": "LqOptResult",
    "lq_objective_fn": "LqObjectiveFn", "lq_gradient_fn": "LqGradientFn",
}
RUST_KEYWORDS = {"in", "type", "ref", "fn", "match", "loop", "move"}


def c_type_to_rust(decl):
    """'const uint8_t* const*' -> '*const *const u8'.  A const before the
    first '*' qualifies the pointee; a const after a '*' qualifies that
    pointer, i.e. the pointee of the next level."""
    toks = re.findall(r"\*|\w+", decl)
    stars = [i for i, t in enumerate(toks) if t == "*"]
    head = toks[:stars[0]] if stars else toks
    base_const = "const" in head
    base = " ".join(t for t in head if t != "const")
    out = C_TO_RUST[base]
    pointee_const = base_const
    for k, i in enumerate(stars):
        out =("*const " if pointee_const else "*mut ") + out
        nxt = stars[k + 1] if k + 1 < len(stars) else len(toks)
        pointee_const = "const" in toks[i + 1:nxt]
    return out


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"#[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"^\s*((?:const\s+)?\w+\s*\**)\s*(lq_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        ret, name, params = m.group(1), m.group(2), " ".join(m.group(3).split())
        if name.endswith("_fn"):
            continue
        args = []
        if params and params != "void":
            for p in params.split(","):
                # an array parameter (`const double m[8]`) is a pointer
                pm = re.match(r"(.*?)(\w+)\s*(\[\d*\])?\s*$", p.strip())
                pname, ctype = pm.group(2), pm.group(1) + ("*" if pm.group(3) else "")
                args.append((pname + "_" if pname in RUST_KEYWORDS else pname, c_type_to_rust(ctype)))
        rt = c_type_to_rust(ret)
        out[name] = (None if rt == "c_void" else rt, args)
    return out


def rust_functions():
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    block = re.search(r'extern "C" \{(.*?)\n\}', src, flags=re.S).group(1)
    block = re.sub(r"//[^\n]*", "", block)
    out = {}
    for m in re.finditer(r"pub fn (lq_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", block, flags=re.S):
        name, params, ret = m.group(1), " ".join(m.group(2).split()), m.group(3)
        args = []
        if params:
            for p in params.split(","):
                pname, ptype = p.split(":", 1)
                args.append((pname.strip(), " ".join(ptype.split())))
        assert name not in out, "declared twice: " + name
        out[name] = (ret.strip() if ret else None, args)
    return out


def test_type_mapping_rules():
    assert c_type_to_rust("const uint8_t* const*") == "*const *const u8"
    assert c_type_to_rust("lq_state**") == "*mut *mut LqState"
    assert c_type_to_rust("const lq_state*") == "*const LqState"
    assert c_type_to_rust("void*") == "*mut c_void"
    assert c_type_to_rust("const char*") == "*const c_char"
    assert c_type_to_rust("double") == "f64"


def test_every_header_function_is_bound_with_matching_types():
    c, r = header_functions(), rust_functions()
    assert len(c) >= 60
    assert sorted(c) == sorted(r), (sorted(set(c) - set(r)), sorted(set(r) - set(c)))
    bad = []
    for name, (ret, args) in c.items():
        rret, rargs = r[name]
        if ret != rret or [t for _, t in args] != [t for _, t in rargs]:
            bad.append((name, (ret, args), (rret, rargs)))
    assert not bad, bad[:3]


def test_constants_match_header():
    text = open(HEADER).read()
    want = {}
    for name, val in re.findall(r"#define (LQ_\w+) (\d+)", text):
        want[name] = int(val)
    for name, val in re.findall(r"^\s*(LQ_\w+) = (\d+)", text, flags=re.M):
        want[name] = int(val)
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    got = {n: int(v) for n, v in re.findall(r"pub const (LQ_\w+): \w+ = (\d+);", src)}
    assert want == got, (set(want) ^ set(got), {k: (want[k], got.get(k)) for k in want if want[k] != got.get(k)})


def test_gate_struct_fields_in_header_order():
    text = open(HEADER).read()
    c_body = re.search(r"typedef struct lq_gate \{(.*?)\} lq_gate;", text, flags=re.S).group(1)
    c_body = re.sub(r"/\*.*?\*/", "", c_body, flags=re.S)
    c_fields = []
    for t, name, n in re.findall(r"(\w+) (\w+)(?:\[(\d+)\])?;", c_body):
        rt = C_TO_RUST[t]
        c_fields.append((name, "[%s; %s]" % (rt, n) if n else rt))
    src = open(os.path.join(CRATE, "src", "lib.rs")).read()
    r_body = re.search(r"pub struct LqGate \{(.*?)\}", src, flags=re.S).group(1)
    r_fields = [(n, " ".join(t.split())) for n, t in re.findall(r"pub (\w+): ([^,]+),", r_body)]
    assert c_fields == r_fields


def test_build_script_links_the_library():
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "rustc-link-lib=dylib=logosq_hip" in build
    cargo = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'links = "logosq_hip"' in cargo


def test_library_exports_every_bound_name():
    import logosq_amd
    lib = logosq_amd.load()
    missing = [n for n in rust_functions() if not hasattr(lib, n)]
    assert not missing, missing
