"""The stress driver (stitch_stress, the reference's CH-05 driver
harness/stress_multires.cu): argument rules on the CPU; on the GPU the S6-only
track (guard bytes, region dumps against the oracle renderer within 1 LSB
with the mask exact) and the end-to-end track (rows equal the plain
runner's)."""
import csv
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

ROOT = Path(__file__).resolve().parents[1]
STRESS = ROOT / "crossstage-stitch_amd" / "bin" / "stitch_stress"
pytestmark = pytest.mark.skipif(not STRESS.exists(), reason="stitch_stress is built by build()")


def run(*args):
    return subprocess.run([str(STRESS), *map(str, args)], capture_output=True, text=True, timeout=300)


def write_ppm(path, img):
    h, w, _ = img.shape
    path.write_bytes(f"P6\n{w} {h}\n255\n".encode() + np.ascontiguousarray(img).tobytes())


def test_argument_rules(tmp_path):
    base = ["--mode", "feather", "--a", "a.ppm", "--b", "b.ppm", "--width", "64", "--height", "48"]
    assert run("--track", "nope", *base, "--output", tmp_path / "o1").returncode == 2
    assert run("--track", "s6_only", *base, "--output", tmp_path / "o2").returncode == 2  # needs --h-file/--rois
    assert run("--track", "end_to_end", *base, "--h-file", "h.txt", "--canvas-width", "9", "--canvas-height", "9",
               "--output", tmp_path / "o3").returncode == 2
    assert run("--track", "end_to_end", *base, "--canvas-width", "9", "--canvas-height", "9", "--repeats", "0",
               "--output", tmp_path / "o4").returncode == 2
    assert run("--track", "end_to_end", *base, "--canvas-width", "9", "--canvas-height", "9", "--bogus", "1",
               "--output", tmp_path / "o5").returncode == 2
    (tmp_path / "exists").mkdir()
    assert run("--track", "end_to_end", *base, "--canvas-width", "9", "--canvas-height", "9",
               "--output", tmp_path / "exists").returncode == 2


def read_metrics(out):
    with open(out / "native_metrics.csv", newline="") as f:
        return list(csv.DictReader(f))


@pytest.mark.gpu
def test_s6_only_track_matches_oracle_renderer(tmp_path):
    from cs_stitch import pipeline as P
    w, h = 640, 480
    a, b, H = P.synth_pair(7300, w, h, overlap=(0.5, 0.7))
    write_ppm(tmp_path / "a.ppm", a)
    write_ppm(tmp_path / "b.ppm", b)
    (tmp_path / "h.txt").write_text(" ".join(repr(float(v)) for v in H.ravel()) + "\n")
    st, geo = O.prepare_geometry(w, h, w, h, H.ravel())
    assert st == 0
    rois = [(0, 0, geo.width, 16), (geo.width // 3, geo.height // 3, 200, 120)]
    (tmp_path / "rois.txt").write_text("".join(f"{x} {y} {bw} {bh}\n" for x, y, bw, bh in rois))
    out = tmp_path / "out"
    r = run("--track", "s6_only", "--mode", "multiband_5", "--a", tmp_path / "a.ppm", "--b", tmp_path / "b.ppm",
            "--h-file", tmp_path / "h.txt", "--rois", tmp_path / "rois.txt", "--width", w, "--height", h,
            "--warmups", 1, "--repeats", 2, "--output", out)
    assert r.returncode == 0, r.stderr
    receipt = json.loads((out / "native_receipt.json").read_text())
    assert receipt["status_code"] == "COMPLETED", receipt
    assert (receipt["canvas_width"], receipt["canvas_height"]) == (geo.width, geo.height)
    rows = read_metrics(out)
    assert len(rows) == 3 and all(row["status_code"] == "SUCCESS" for row in rows)
    pano, mask, _ = O.s6_render(a, b, geo, O.MODES["multiband_5"])
    for i, (x, y, bw, bh) in enumerate(rois):
        got = np.frombuffer((out / f"roi{i}.rgb").read_bytes(), np.uint8).reshape(bh, bw, 3)
        gm = np.frombuffer((out / f"roi{i}.mask").read_bytes(), np.uint8).reshape(bh, bw)
        assert np.array_equal(gm, mask[y:y + bh, x:x + bw])
        assert np.abs(got.astype(int) - pano[y:y + bh, x:x + bw].astype(int)).max() <= 1


@pytest.mark.gpu
def test_end_to_end_track_rows_match_runner(tmp_path):
    from cs_stitch import pipeline as P
    w, h = 640, 480
    a, b, _ = P.synth_pair(7301, w, h, overlap=(0.5, 0.7))
    write_ppm(tmp_path / "a.ppm", a)
    write_ppm(tmp_path / "b.ppm", b)
    out = tmp_path / "out"
    r = run("--track", "end_to_end", "--mode", "feather", "--a", tmp_path / "a.ppm", "--b", tmp_path / "b.ppm",
            "--width", w, "--height", h, "--canvas-width", 2048, "--canvas-height", 1024, "--scale", "0.5",
            "--keypoints", 1000, "--ransac", 256, "--seed", 3, "--warmups", 1, "--repeats", 2, "--output", out)
    assert r.returncode == 0, r.stderr
    receipt = json.loads((out / "native_receipt.json").read_text())
    assert receipt["status_code"] == "COMPLETED", receipt
    rows = read_metrics(out)
    runner = P.PipelineRunner(w, h, max_k=1000, max_r=256, canvas_width=2048, canvas_height=1024)
    runner.upload_pair(a, b)
    rec = runner.execute_run(P.harness_config(S=0.5, K=1000, R=256, seed=3))
    assert len(rows) == 3
    for row in rows:
        assert row["status_code"] == ("SUCCESS" if rec.status == b"OK" else rec.status.decode())
        assert int(row["inlier_count"]) == rec.num_inliers
        if rec.status == b"OK":
            assert [float(row[k]) for k in ("H00", "H01", "H02", "H10", "H11", "H12", "H20", "H21", "H22")] == \
                list(rec.H)
    runner.close()
