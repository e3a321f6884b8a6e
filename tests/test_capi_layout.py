"""The ctypes mirrors of the C-ABI structs have the sizes and field offsets
the C compiler gives include/cs_stitch.h (host-only: gcc on a generated probe)."""
import ctypes
import shutil
import subprocess

import pytest

from test_capi_exports import ROOT

STRUCTS = {
    "cs_image": "Image", "cs_scale_info": "ScaleInfo", "cs_keypoints": "Keypoints", "cs_descriptors": "Descriptors",
    "cs_matches": "Matches", "cs_homography_result": "HomographyResult", "cs_top2": "Top2",
    "cs_s2_config": "S2Config", "cs_s4_config": "S4Config", "cs_s5_config": "S5Config",
    "cs_geometry_policy": "GeometryPolicy", "cs_geometry": "Geometry", "cs_render_statistics": "RenderStatistics",
    "cs_harness_config": "HarnessConfig", "cs_run_record": "RunRecord", "cs_pair_spec": "PairSpec",
    "cs_pair_outcome": "PairOutcome", "cs_batch_timings": "BatchTimings", "cs_tuning": "Tuning",
    "cs_server_config": "ServerConfig", "cs_request": "Request", "cs_completion": "Completion",
    "cs_serve_stats": "ServeStats", "cs_kernel_probe": "KernelProbe", "cs_attempt_record": "AttemptRecord",
    "cs_strip_record": "StripRecord",
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_the_header(tmp_path):
    from cs_stitch import _ffi as F
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cs_stitch.h"', "int main(void) {"]
    for c_name, py_name in STRUCTS.items():
        cls = getattr(F, py_name)
        lines.append(f'printf("{c_name} size %zu\\n", sizeof({c_name}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{c_name} {field} %zu\\n", offsetof({c_name}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for c_name, py_name in STRUCTS.items():
        cls = getattr(F, py_name)
        assert got[(c_name, "size")] == ctypes.sizeof(cls), c_name
        for field, _ in cls._fields_:
            assert got[(c_name, field)] == getattr(cls, field).offset, (c_name, field)
