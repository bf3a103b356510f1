"""The C ABI library loads on the CPU and exports every declared entry point.

No compute calls: this runs without a GPU.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))
                 if h.endswith(".h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(osn_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib_path():
    import osn_tsdf_hip as H
    if not os.path.exists(H.LIB_PATH):
        import build
        build.build()
    return H.LIB_PATH


def test_every_declared_symbol_is_exported(lib_path):
    names = declared_functions()
    assert len(names) >= 35
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds(lib_path):
    import osn_tsdf_hip as H
    L = H.lib()
    for n in declared_functions():
        assert getattr(L, n) is not None


def test_reference_mirror_is_complete():
    """Every osn_tsdf_cuda_* entry point of the reference device ABI
    (cpp/include/osn_tsdf/c_api.h:80-126) has an osn_tsdf_hip_* twin."""
    reference = ["create", "destroy", "valid", "device_view", "allocate_blocks", "update_voxels",
                 "extract_mesh", "reset", "allocate_counting_cas", "allocate_no_cas",
                 "allocate_slot_index", "block_count", "drop_count", "synchronize"]
    names = set(declared_functions())
    for r in reference:
        assert f"osn_tsdf_hip_{r}" in names


def test_device_view_layout_matches_reference():
    """OsnTsdfDeviceView is 104 bytes with the field order of c_api.h:61-78."""
    import osn_tsdf_hip as H
    V = H.DeviceView
    assert C.sizeof(V) == 104
    offsets = [getattr(V, f).offset for f, _ in V._fields_]
    assert offsets == [0, 8, 16, 24, 32, 40, 48, 56, 64, 72, 76, 80, 84, 88, 92, 96]


def test_config_and_frame_layout_match_header(tmp_path):
    """The ctypes mirrors of OsnTsdfHipConfig and OsnTsdfHipFrame have the
    header's field count and size (compiled probe of the C header)."""
    import osn_tsdf_hip as H
    src = ('#include "osn_tsdf_hip.h"\n#include <stdio.h>\n#include <stddef.h>\n'
           'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(OsnTsdfHipConfig),'
           ' sizeof(OsnTsdfHipFrame), offsetof(OsnTsdfHipFrame, R_wc64),'
           ' offsetof(OsnTsdfHipFrame, radius_m)); return 0;}\n')
    c = str(tmp_path / "layout_probe.c")
    exe = str(tmp_path / "layout_probe")
    open(c, "w").write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [C.sizeof(H.Config), C.sizeof(H.Frame), H.Frame.R_wc64.offset,
                   H.Frame.radius_m.offset]
