"""CPU: every struct the ctypes bindings mirror has the layout include/cnq.h
gives it: gcc compiles a program that prints each struct's size and each
field's offset, and the ctypes classes must agree field by field."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from continuum_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "cnq.h"

MIRRORS = {
    "cnq_options": _lib.Options, "cnq_sky_options": _lib.SkyOptions, "cnq_scene_desc": _lib.SceneDesc,
    "cnq_placement": _lib.Placement, "cnq_view": _lib.View, "cnq_tuning": _lib.Tuning,
    "cnq_server_config": _lib.ServerConfig, "cnq_mover_tree": _lib.MoverTree,
    "cnq_request_info": _lib.RequestInfo, "cnq_vcache_surfaces": _lib.VcacheSurfaces,
    "cnq_vcache_config": _lib.VcacheConfig, "cnq_vcache_shade_params": _lib.VcacheShadeParams,
    "cnq_moved": _lib.MovedBox,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_the_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in MIRRORS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field in cls._fields_:
            lines.append(f'  printf("{cname} {field[0]} %zu\\n", offsetof({cname}, {field[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, field, value = line.split()
        got[(name, field)] = int(value)
    for cname, cls in MIRRORS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field in cls._fields_:
            assert got[(cname, field[0])] == getattr(cls, field[0]).offset, (cname, field[0])
