"""The C-ABI library loads and exports every entry point include/bi_pt.h declares;
the ctypes mirrors match the header's struct layouts (no GPU needed)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import bi_pt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bi_pt.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[a-z_0-9]+\s*\*?\s+)+\**(bi_pt_[a-z_0-9]+)\s*\(", text, re.M)))


def test_header_declarations_are_bound():
    names = declared_functions()
    assert len(names) >= 15
    assert set(names) == set(bi_pt.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(bi_pt.LIB_PATH):
        pytest.skip("libbi_pt.so not built")
    lib = ctypes.CDLL(bi_pt.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", bi_pt.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (bi_pt_[a-z_0-9]+)$", out.stdout, re.M))
    assert set(declared_functions()) <= exported


def test_struct_layouts_match_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "bi_pt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(bi_pt_device_desc), sizeof(bi_pt_mesh),
         sizeof(bi_pt_instance), sizeof(bi_pt_material), sizeof(bi_pt_dir_light), sizeof(bi_pt_camera),
         sizeof(bi_pt_settings), sizeof(bi_pt_stats), offsetof(bi_pt_stats, ms_build), sizeof(bi_pt_texture),
         offsetof(bi_pt_material, base_color_tex));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    import scenes
    mirror = [ctypes.sizeof(bi_pt.DeviceDesc), ctypes.sizeof(bi_pt.Mesh), ctypes.sizeof(bi_pt.Instance),
              4 * scenes.MATERIAL_FLOATS, 64, ctypes.sizeof(bi_pt.CameraC), ctypes.sizeof(bi_pt.Settings),
              ctypes.sizeof(bi_pt.Stats), bi_pt.Stats.ms_build.offset, ctypes.sizeof(bi_pt.TextureC),
              4 * scenes.M_BASE_TEX]
    assert sizes == mirror
    assert sizes[3] == 128 and sizes[4] == 64   # bi_pt_material, bi_pt_dir_light (= DirLightData)


def test_no_gpu_means_loud_failure():
    """Without a visible GPU the product path raises instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        bi_pt.PathTracer()


def test_assets_library_exports_its_header():
    """libbi_assets.so (host-only StaticMesh loader) exports every bi_static_mesh_* in include/bi_assets.h."""
    path = os.path.join(ROOT, "bisemutum-engine_amd", "libbi_assets.so")
    if not os.path.exists(path):
        pytest.skip("libbi_assets.so not built")
    text = open(os.path.join(ROOT, "include", "bi_assets.h")).read()
    names = set(re.findall(r"^\s*(?:[a-z_0-9]+\s*\*?\s+)+\**(bi_static_mesh_[a-z_0-9]+)\s*\(", text, re.M))
    assert names == {"bi_static_mesh_load", "bi_static_mesh_destroy", "bi_static_mesh_counts", "bi_static_mesh_submesh"}
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (bi_static_mesh_[a-z_0-9]+)$", out.stdout, re.M))
    assert names <= exported
