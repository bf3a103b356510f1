"""The timing harness binary on the GPU: records valid cells with the published
block counts, writes the reference CSV schema, exports OSN1 meshes whose vertex
counts are the published ones for both arms (vertex for vertex the same
triangles: pool order depends on the key set alone; positions agree to float
atomics' summation order, as the reference's scatter update documents), and
refuses to run on a shared device."""
import csv
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(os.path.dirname(HERE), "what-irregularity-costs_amd", "osn_harness")


def _free_cached_device_memory():
    try:
        import torch
        torch.cuda.empty_cache()
    except Exception:
        pass


def test_harness_cells_csv_and_export(tmp_path):
    import formats as F
    assert os.path.exists(HARNESS), "build the library (it builds the harness beside it)"
    _free_cached_device_memory()
    out_csv = tmp_path / "h.csv"
    ex = tmp_path / "export"
    ex.mkdir()
    r = subprocess.run([HARNESS, "--cells", "base,plane-320k", "--reps", "4", "--warmup", "1",
                        "--batch", "2", "--csv", str(out_csv), "--export", str(ex),
                        "--max-foreign-gib", "1000"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = list(csv.DictReader(open(out_csv)))
    assert list(rows[0].keys())[:6] == ["run", "device", "device_name", "sm", "sm_count", "cell"]
    assert {(x["cell"], x["arm"], x["stage"]) for x in rows} == {
        (c, a, s) for c in ("base", "plane-320k") for a in ("hip", "hip-noset")
        for s in ("allocate", "update")} | {(c, "hip", "extract-hip-only")
                                            for c in ("base", "plane-320k")}
    blocks = {x["cell"]: int(x["blocks"]) for x in rows}
    assert blocks == {"base": 1160, "plane-320k": 1060}
    for x in rows:
        assert 0 < float(x["min_ms"]) <= float(x["p50_ms"]) <= float(x["p95_ms"]) <= float(x["p99_ms"])
        assert int(x["n"]) == 4
    for cell, verts in (("base", 846288), ("plane-320k", 950424)):
        a = F.read_osn1(str(ex / f"{cell}.hip.raw"))
        b = F.read_osn1(str(ex / f"{cell}.hip-noset.raw"))
        assert len(a) == len(b) == verts
        assert np.abs(a[:, :3] - b[:, :3]).max() < 1e-5


def test_harness_refuses_a_shared_device():
    r = subprocess.run([HARNESS, "--cells", "base", "--max-foreign-gib", "-1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 3 and "REFUSING" in r.stdout
