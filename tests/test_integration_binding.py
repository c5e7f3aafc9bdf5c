"""CPU: the Rust binding INTEGRATION.md gives a maintainer (its ```rust
blocks) matches include/cnq.h: every function the header declares is
declared in Rust with the same parameters (count, and per parameter the
same kind: the scalar's width and type, or a pointer to the same element
type with the same constness) and return type, and nothing else; every
struct the header defines has a #[repr(C)] twin whose field names and order
are the header's and whose C layout (offsets and size, computed from the
Rust field types) equals what gcc gives the header's struct; every opaque
handle has its zero-sized twin."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cnq.h"
DOC = ROOT / "INTEGRATION.md"

C_SCALARS = {"float": "f32", "double": "f64", "uint32_t": "u32", "uint64_t": "u64", "int32_t": "i32", "int": "c_int",
             "uint8_t": "u8", "char": "c_char", "void": "c_void"}
SIZES = {"u8": 1, "u32": 4, "i32": 4, "c_int": 4, "f32": 4, "u64": 8, "f64": 8, "ptr": 8}


def camel(name: str) -> str:
    return "".join(p.capitalize() for p in name.split("_"))


def header_text() -> str:
    return re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)


def c_kind(ctype: str, array: bool = False) -> tuple:
    """('ptr', const, element) for pointers and array parameters, else
    ('val', rust scalar or struct name)."""
    t = " ".join(ctype.split())
    const = t.startswith("const ")
    if const:
        t = t[6:]
    stars = t.count("*") + (1 if array else 0)
    base = t.replace("*", "").strip()
    elem = C_SCALARS.get(base) or camel(base)
    if stars == 0:
        return ("val", elem)
    inner = elem if stars == 1 else "*mut " * (stars - 1) + elem
    return ("ptr", const, inner)


def rust_kind(rtype: str) -> tuple:
    t = " ".join(rtype.split())
    if t.startswith("*const ") or t.startswith("*mut "):
        const = t.startswith("*const ")
        inner = t.split(" ", 1)[1]
        return ("ptr", const, inner)
    return ("val", t)


def header_functions() -> dict:
    out = {}
    for m in re.finditer(r"\n((?:const\s+)?\w+\s*\**)\s*(cnq_\w+)\s*\(([^)]*)\)\s*;", header_text()):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        kinds = []
        if params.strip() not in ("", "void"):
            for p in params.split(","):
                a = re.match(r"(.*?)(\w+)\s*(\[\d*\])?$", " ".join(p.split()))
                kinds.append(c_kind(a.group(1), bool(a.group(3))))
        out[name] = (kinds, None if ret.strip() == "void" else c_kind(ret))
    return out


def header_structs() -> dict:
    """name -> [(field, rust type)] for every struct the header defines."""
    out = {}
    for m in re.finditer(r"typedef struct (\w+)?\s*\{(.*?)\}\s*(\w+);", header_text(), flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            mm = re.match(r"((?:const\s+)?\w+\s*\**)\s*(.*)", decl, flags=re.S)
            for nm in mm.group(2).split(","):
                nm = nm.strip()
                stars = len(nm) - len(nm.lstrip("*"))
                a = re.match(r"(\w+)\s*(\[(\d+)\])?", nm.lstrip("*").strip())
                kind = c_kind(mm.group(1) + "*" * stars)
                fields.append((a.group(1), kind, int(a.group(3)) if a.group(3) else None))
        out[m.group(3)] = fields
    return out


def header_opaque() -> set:
    return {b for a, b in re.findall(r"typedef struct (\w+) (\w+);", header_text())}


def rust_blocks() -> str:
    return "\n".join(re.findall(r"```rust\n(.*?)```", DOC.read_text(), flags=re.S))


def rust_functions() -> dict:
    text = rust_blocks()
    out = {}
    for block in re.findall(r'extern "C" \{(.*?)\n\}', text, flags=re.S):
        for m in re.finditer(r"pub fn (\w+)\((.*?)\)\s*(->\s*([^;]+))?;", block, flags=re.S):
            name, params, ret = m.group(1), m.group(2), m.group(4)
            kinds = []
            for p in [x for x in params.split(",") if x.strip()]:
                kinds.append(rust_kind(p.split(":", 1)[1]))
            assert name not in out, f"{name} declared twice"
            out[name] = (kinds, rust_kind(ret) if ret else None)
    return out


def rust_structs() -> tuple:
    text = rust_blocks()
    full, opaque = {}, set()
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+)\s*\{(.*?)\}", text, flags=re.S):
        name, body = m.group(1), m.group(2)
        if body.strip() == "_p: [u8; 0]":
            opaque.add(name)
            continue
        fields = []
        for f in [x.strip() for x in body.split(",") if x.strip()]:
            fname, ftype = (s.strip() for s in f.replace("pub ", "", 1).split(":", 1))
            fields.append((fname, ftype))
        assert name not in full, f"struct {name} defined twice"
        full[name] = fields
    return full, opaque


def rust_layout(fields, structs, cache=None) -> tuple:
    """C layout of a #[repr(C)] struct: (offsets, size, align)."""
    cache = {} if cache is None else cache

    def size_align(t):
        t = " ".join(t.split())
        a = re.match(r"\[(.*);\s*(\d+)\]$", t)
        if a:
            s, al = size_align(a.group(1))
            return s * int(a.group(2)), al
        if t.startswith("*"):
            return 8, 8
        if t in SIZES:
            return SIZES[t], SIZES[t]
        if t not in cache:
            _, sz, al = rust_layout(structs[t], structs, cache)
            cache[t] = (sz, al)
        return cache[t]

    off, align, offsets = 0, 1, []
    for _, t in fields:
        s, al = size_align(t)
        off = (off + al - 1) // al * al
        offsets.append(off)
        off += s
        align = max(align, al)
    return offsets, (off + align - 1) // align * align, align


def test_every_function_is_bound_with_the_headers_parameters():
    h, r = header_functions(), rust_functions()
    assert len(h) >= 70
    assert sorted(r) == sorted(h), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (params, ret) in h.items():
        rparams, rret = r[name]
        assert len(rparams) == len(params), name
        for k, (a, b) in enumerate(zip(params, rparams)):
            assert a == b, f"{name} parameter {k}: header {a}, Rust {b}"
        assert ret == rret, f"{name} return: header {ret}, Rust {rret}"


def test_every_handle_has_an_opaque_twin():
    _, opaque = rust_structs()
    assert {camel(n) for n in header_opaque()} == opaque


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_every_struct_has_the_headers_layout(tmp_path):
    hs = header_structs()
    rs, _ = rust_structs()
    assert sorted(rs) == sorted(camel(n) for n in hs)
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, fields in hs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _, _ in fields:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    gcc = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        n, f, v = line.split()
        gcc[(n, f)] = int(v)
    for cname, fields in hs.items():
        rfields = rs[camel(cname)]
        assert [f for f, _ in rfields] == [f for f, _, _ in fields], cname
        for (f, ckind, dim), (_, rtype) in zip(fields, rfields):
            want = ckind[1] if ckind[0] == "val" else None
            if dim is not None:
                assert rtype.replace(" ", "") == f"[{want};{dim}]", (cname, f, rtype)
            elif want is not None:
                assert rtype == want, (cname, f, rtype)
            else:
                assert rust_kind(rtype) == ckind, (cname, f, rtype)
        offsets, size, _ = rust_layout(rfields, rs)
        assert size == gcc[(cname, "size")], cname
        for (f, _), off in zip(rfields, offsets):
            assert off == gcc[(cname, f)], (cname, f)
