"""The Rust FFI module (rust/gpu.rs) against include/gargantua.h, without a
Rust toolchain: every #[repr(C)] struct has its C typedef's fields in the
same order, with the offsets and size gcc gives the header (computed here
with the C layout rules Rust's repr(C) follows), and the extern "C" block
binds exactly the header's entry points with the header's parameter counts,
all exported by the library."""
import os
import re
import subprocess

import pytest

from gargantua_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUST = os.path.join(ROOT, "rust", "gpu.rs")
HEADER = os.path.join(ROOT, "include", "gargantua.h")

_SCALARS = {"f64": (8, 8), "u64": (8, 8), "i64": (8, 8), "usize": (8, 8), "f32": (4, 4), "i32": (4, 4),
            "u32": (4, 4), "c_int": (4, 4), "u8": (1, 1), "i8": (1, 1)}


def _c_name(rust_name):
    return "gg_" + re.sub(r"(?<!^)(?=[A-Z])", "_", rust_name[2:]).lower()


def _rust_structs():
    text = re.sub(r"//[^\n]*", "", open(RUST).read())
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^\]]*\)\]\s*)?pub struct (\w+)\s*\{(.*?)\}", text, re.S):
        fields = [(n, t.strip()) for n, t in re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", m.group(2).strip())]
        out[m.group(1)] = fields
    return out


def _layout(name, structs, cache):
    """(size, align, {field: offset}) of a repr(C) struct."""
    if name in cache:
        return cache[name]
    off, align, offs = 0, 1, {}
    for f, t in structs[name]:
        size, al = _type_layout(t, structs, cache)
        off = (off + al - 1) // al * al
        offs[f] = off
        off += size
        align = max(align, al)
    size = (off + align - 1) // align * align
    cache[name] = (size, align, offs)
    return cache[name]


def _type_layout(t, structs, cache):
    t = t.strip()
    if t.startswith("*mut") or t.startswith("*const"):
        return 8, 8
    m = re.fullmatch(r"\[(\w+);\s*(\d+)\]", t)
    if m:
        s, a = _type_layout(m.group(1), structs, cache)
        return s * int(m.group(2)), a
    if t in _SCALARS:
        return _SCALARS[t]
    size, al, _ = _layout(t, structs, cache)
    return size, al


def _c_offsets(tmp_path, structs):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gargantua.h"', "int main(void) {"]
    for rs, fields in structs.items():
        c = _c_name(rs)
        lines.append(f'  printf("{c} sizeof %zu\\n", sizeof({c}));')
        for f, _ in fields:
            lines.append(f'  printf("{c} {f} %zu\\n", offsetof({c}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "rust_layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "rust_layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    return {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}


def _c_struct_fields():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"typedef struct (gg_\w+) \{(.*?)\} \1;", text, re.S):
        out[m.group(1)] = re.findall(r"(\w+)(?:\[\d+\])?\s*;", m.group(2))
    return out


def test_rust_structs_mirror_the_header(tmp_path):
    structs = _rust_structs()
    c_fields = _c_struct_fields()
    assert set(_c_name(s) for s in structs) == set(c_fields), "a struct of the header has no Rust mirror"
    for rs, fields in structs.items():
        assert [f for f, _ in fields] == c_fields[_c_name(rs)], rs
    got = _c_offsets(tmp_path, structs)
    cache = {}
    for rs, fields in structs.items():
        size, _, offs = _layout(rs, structs, cache)
        c = _c_name(rs)
        assert got[(c, "sizeof")] == size, (rs, got[(c, "sizeof")], size)
        for f, _ in fields:
            assert got[(c, f)] == offs[f], (rs, f, got[(c, f)], offs[f])


def _c_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\n(?:const )?[a-z_0-9]+\s*\*?\s*(gg_[a-z_0-9]+)\s*\(([^;]*?)\);", text):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return protos


def test_rust_extern_block_binds_every_entry_point():
    text = re.sub(r"//[^\n]*", "", open(RUST).read())
    block = re.search(r'extern "C" \{(.*?)\n\}', text, re.S).group(1)
    fns = {}
    for m in re.finditer(r"pub fn (gg_\w+)\((.*?)\)\s*(?:->[^;]*)?;", block, re.S):
        args = m.group(2).strip()
        fns[m.group(1)] = 0 if not args else args.count(",") + 1 - (1 if args.rstrip().endswith(",") else 0)
    protos = _c_prototypes()
    assert set(fns) == set(protos)
    for name, n in fns.items():
        assert n == protos[name], (name, n, protos[name])
    lib = _abi.load()
    for name in fns:
        assert hasattr(lib, name), name


def test_rust_constants_match_the_header():
    text = open(RUST).read()
    consts = dict(re.findall(r"pub const (GG_\w+): c_int = (\d+);", text))
    header = open(HEADER).read()
    assert int(consts["GG_ABI_VERSION"]) == int(re.search(r"#define GG_ABI_VERSION (\d+)", header).group(1))
    for k in ("GG_PREC_FP64", "GG_PREC_MIXED", "GG_PREC_FP64_FAST"):
        assert int(consts[k]) == int(re.search(rf"\b{k}\s*=\s*(\d+)", header).group(1)), k


@pytest.mark.parametrize("rust_type,c_type", [("GgScene", "gg_scene"), ("GgGradScene", "gg_grad_scene")])
def test_nested_layouts_agree(tmp_path, rust_type, c_type):
    structs = _rust_structs()
    got = _c_offsets(tmp_path, {rust_type: structs[rust_type]})
    assert got[(c_type, "sizeof")] == _layout(rust_type, structs, {})[0]
