"""The C-ABI library loads and exports every symbol include/argent.h
declares (no compute calls: CPU only); argument validation on the host."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "argent.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ag_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(A.LIB_PATH)
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(A.EXPORTED) <= set(declared())


def test_struct_layout_matches_the_c_compiler(tmp_path):
    """ctypes mirrors have the size and field offsets gcc gives the header."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    structs = {"ag_material": A.Material, "ag_panel": A.Panel, "ag_env": A.Env, "ag_floor": A.Floor,
               "ag_scene_desc": A.SceneDesc, "ag_camera": A.Camera, "ag_path_settings": A.PathSettingsC,
               "ag_rays_soa": A.RaysSoA, "ag_hits_soa": A.HitsSoA, "ag_view": A.View, "ag_tuning": A.Tuning,
               "ag_stats": A.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "argent.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_oracle_is_not_linked_by_the_product():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
    assert "or_" not in " ".join(line.split()[-1] for line in out.splitlines() if line.strip() and line.split()[-1].startswith("or_"))
    assert not any(line.split()[-1].startswith("or_") for line in out.splitlines() if line.strip())


def test_descriptor_validation_rejects_bad_meshes():
    p = np.zeros((3, 3), np.float32)
    bad = np.array([[0, 1, 5]], np.uint32)
    d = ag.SceneDesc(p, bad)
    d.add_material(ag.Material.diffuse([0.5, 0.5, 0.5]))
    h = C.c_void_p()
    # no context is needed to reach the descriptor check: a null context fails first
    assert A.lib().ag_scene_create(None, C.byref(d.d), C.byref(h)) == 1
    assert b"null" in A.lib().ag_last_error()
