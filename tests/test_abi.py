"""The C ABI library without a GPU: it loads, exports every function
include/shiro_hip.h declares, its struct layouts match the ctypes mirror,
and the defaults are the reference's (Renderer.h:29-102)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from shiro_amd import abi
from shiro_amd.hip import LIB_PATH, load_library
from shiro_amd.scene import PbrMaterial, RenderSettings, Scene, TriangleMesh, marshal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shiro_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(shiro_hip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB_PATH), "libshiro_hip.so not built"
    L = load_library()
    names = declared_functions()
    assert len(names) >= 20
    assert sorted(names) == sorted(abi.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match_ctypes(tmp_path):
    structs = {"shiro_hip_material": abi.Material, "shiro_hip_mesh": abi.Mesh,
               "shiro_hip_directional_light": abi.DirectionalLight, "shiro_hip_dome_light": abi.DomeLight,
               "shiro_hip_texture": abi.Texture, "shiro_hip_scene": abi.Scene, "shiro_hip_camera": abi.Camera,
               "shiro_hip_settings": abi.Settings, "shiro_hip_frame": abi.Frame, "shiro_hip_status": abi.Status,
               "shiro_hip_scene_info": abi.SceneInfo, "shiro_hip_fetch_counts": abi.FetchCounts,
               "shiro_hip_wavefront_counts": abi.WavefrontCounts,
               "shiro_hip_tuning": abi.Tuning, "shiro_hip_stage_times": abi.StageTimes,
               "shiro_hip_batch_item": abi.BatchItem, "shiro_hip_image": abi.Image}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    for cname, cls in structs.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"


def test_default_settings_are_reference_defaults():
    L = load_library()
    s = abi.Settings()
    L.shiro_hip_default_settings(C.byref(s))
    # Renderer.h:29-45.
    ref = dict(width=1280, height=720, samplesPerPixel=32, samplesPerUpdate=1, domeLightSamples=1, maxDepth=3,
               diffuseDepth=2, specularDepth=2, threadLimit=0, maxSubdivLevel=2, backend=0, backgroundVisible=1,
               enableHeadlight=1, enableCaustics=1, enableSss=1, russianRouletteDepth=2, transport=0)
    for k, v in ref.items():
        assert getattr(s, k) == v, k
    py = RenderSettings().to_c()
    for k, _ in abi.Settings._fields_:
        assert getattr(py, k) == getattr(s, k), k


def test_default_material_is_reference_default():
    L = load_library()
    m = abi.Material()
    L.shiro_hip_default_material(C.byref(m))
    assert bytes(m) == bytes(PbrMaterial().to_c())
    assert list(m.baseColor) == pytest.approx([0.8, 0.8, 0.8])
    assert m.roughness == pytest.approx(0.5) and m.coatRoughness == pytest.approx(0.03)
    assert m.ior == pytest.approx(1.5) and m.opacity == 1.0


def _tri_scene(**mat):
    m = PbrMaterial(**mat)
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    return Scene(materials=[m], meshes=[TriangleMesh(p, np.zeros((0, 3), np.float32), np.array([0, 1, 2], np.uint32))])


def test_supports_scene():
    L = load_library()
    assert L.shiro_hip_supports_scene(marshal(_tri_scene()).ptr) == 1
    # Emissive surfaces are sampled as triangle lights on the device.
    assert L.shiro_hip_supports_scene(marshal(_tri_scene(emissionColor=(1, 1, 1), emissionStrength=1.0)).ptr) == 1
    # Lights live in device arrays: scenes with many directional and dome
    # lights are supported (the reference has no limit); only counts beyond
    # the per-path light-sample bound (1024 directional, 256 dome) fall back.
    from shiro_amd.scene import DirectionalLight, DomeLight
    many = _tri_scene()
    many.distantLights = [DirectionalLight() for _ in range(40)]
    many.domeLights = [DomeLight() for _ in range(12)]
    assert L.shiro_hip_supports_scene(marshal(many).ptr) == 1
    many.distantLights = [DirectionalLight() for _ in range(1025)]
    assert L.shiro_hip_supports_scene(marshal(many).ptr) == 0
    assert L.shiro_hip_supports_scene(None) == 0


def test_create_without_device_reports_no_device():
    L = load_library()
    n = C.c_int()
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.shiro_hip_create(0, C.byref(h)) == 1  # SHIRO_HIP_NO_DEVICE
    del n


def test_oracle_library_loads():
    from oracle.bindings import oracle_fns
    f = oracle_fns()
    assert f.qmc(0, 0, 0, 0) == 0.5
