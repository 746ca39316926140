"""C ABI checks that need no GPU: the library loads, exports every entry point
include/gargantua.h declares, and its host-side validation rejects invalid
scenes with the documented status codes before anything reaches a kernel."""
import ctypes as C
import math
import os
import re

import pytest

import gargantua_amd as gg
from gargantua_amd import _abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gargantua.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gg_[a-z_0-9]+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert set(_declared()) == {name for name, _, _ in _abi.EXPORTS}


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.gg_abi_version() == _abi.ABI_VERSION


def test_struct_sizes_match_c_layout():
    # sizes of the POD structs as the C compiler lays them out (x86-64 and gfx950 agree)
    assert C.sizeof(_abi.gg_metric) == 24
    assert C.sizeof(_abi.gg_camera) == 96
    assert C.sizeof(_abi.gg_disk) == 72
    assert C.sizeof(_abi.gg_integrator) == 80
    assert C.sizeof(_abi.gg_image) == 24
    assert C.sizeof(_abi.gg_scene) == 24 + 96 + 72 + 80 + 24
    assert C.sizeof(_abi.gg_outputs) == 80
    assert C.sizeof(_abi.gg_hittable) == 128


def test_python_mirror_matches_c_field_offsets(tmp_path):
    """Every struct the ABI mirrors: size and each field's offset as gcc lays
    out include/gargantua.h, against the ctypes mirror."""
    import subprocess
    structs = [v for v in vars(_abi).values()
               if isinstance(v, type) and issubclass(v, C.Structure) and v.__name__.startswith("gg_")]
    assert structs
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gargantua.h"', "int main(void) {"]
    for st in structs:
        lines.append(f'  printf("{st.__name__} sizeof %zu\\n", sizeof({st.__name__}));')
        for f in st._fields_:
            lines.append(f'  printf("{st.__name__} {f[0]} %zu\\n", offsetof({st.__name__}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (l.split() for l in out if l)}
    for st in structs:
        assert got[(st.__name__, "sizeof")] == C.sizeof(st), st.__name__
        for f in st._fields_:
            assert got[(st.__name__, f[0])] == getattr(st, f[0]).offset, (st.__name__, f[0])


def _rays(scene):
    lib = _abi.load()
    n = C.c_uint64()
    sc = scene.to_c()
    rc = lib.gg_scene_rays(C.byref(sc), C.byref(n))
    return rc, n.value, lib.gg_last_error().decode()


def test_scene_rays_counts_subpixels():
    sc = gg.configs.kerr_1024(64, 32)
    sc.samples_per_axis = 3
    assert _rays(sc)[:2] == (0, 64 * 3 * 32 * 3)


@pytest.mark.parametrize("mutate,needle", [
    (lambda s: setattr(s.metric, "mass", 0.0), "mass"),
    (lambda s: setattr(s.integrator, "min_step", 0.0), "min_step"),
    (lambda s: setattr(s.integrator, "max_step", 1e-9), "max_step"),
    (lambda s: setattr(s.integrator, "err_mode", 7), "err_mode"),
    (lambda s: setattr(s.integrator, "rtol", -1.0), "rtol"),
    (lambda s: setattr(s, "width", 0), "width"),
    (lambda s: setattr(s, "samples_per_axis", 0), "samples_per_axis"),
    (lambda s: setattr(s.camera, "fov_y_radians", 4.0), "fov"),
    (lambda s: setattr(s.camera, "look_at", (0.0, 0.0, 1.0)), "look_at"),
    (lambda s: setattr(s.disk, "edge_width", -1.0), "edge_width"),
])
def test_invalid_scenes_are_rejected(mutate, needle):
    sc = gg.configs.kerr_1024(16, 16)
    mutate(sc)
    rc, _, msg = _rays(sc)
    assert rc == -1 and needle in msg, (rc, msg)


def test_kerr_new_panics_on_nonpositive_mass():
    with pytest.raises(ValueError):
        gg.Kerr(0.0, 0.5)


def test_null_arguments_are_rejected():
    lib = _abi.load()
    assert lib.gg_scene_rays(None, None) == -1
    assert lib.gg_context_create(0, None, None) == -1
    assert lib.gg_context_destroy(None) == 0
    assert lib.gg_strerror(-4).decode() == "out of device memory"
