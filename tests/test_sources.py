"""Static checks on the kernel sources that keep graph replay safe: a captured
launch is sized for the capacity it was captured for, so every kernel must
take its job count and tile total from the device (JobSpan::jobs(),
dev::Total) and walk jobs with a grid-stride loop - never index the job
table by blockIdx directly."""
import re
from pathlib import Path

CSRC = Path(__file__).resolve().parents[1] / "crossstage-stitch_amd" / "csrc"


def kernels():
    for f in sorted(CSRC.glob("*.hip")):
        text = f.read_text()
        for m in re.finditer(r"__global__ void\s+(?:__launch_bounds__\([^)]*\)\s+)?(\w+)\s*\(([^)]*)\)", text):
            yield f.name, m.group(1), m.group(2)


def test_no_job_indexed_by_block():
    for f in CSRC.glob("*.hip"):
        assert not re.search(r"job_at\(\s*span\s*,\s*blockIdx", f.read_text()), f.name


def test_span_kernels_take_device_totals():
    n = 0
    for name, kernel, params in kernels():
        if "JobSpan" not in params:
            continue
        n += 1
        assert "long long total" not in params, (name, kernel)
    assert n >= 25
