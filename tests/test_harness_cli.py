"""stitch_harness, the command-line campaign driver (harness/main.cpp:120-483):
argument handling on the CPU; on the GPU, a JSON manifest of PPM pairs run in
fixed mode against the plain runner's own results, and an adaptive campaign
with its per-attempt trace."""
import csv
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HARNESS = ROOT / "crossstage-stitch_amd" / "bin" / "stitch_harness"

pytestmark = pytest.mark.skipif(not HARNESS.exists(), reason="stitch_harness is built by build()")


def run(*args, cwd=None):
    return subprocess.run([str(HARNESS), *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=300)


def write_ppm(path, rgb):
    h, w, _ = rgb.shape
    path.write_bytes(f"P6\n# synthetic\n{w} {h}\n255\n".encode() + np.ascontiguousarray(rgb, np.uint8).tobytes())


def test_help_and_argument_errors(tmp_path):
    r = run("--help")
    assert r.returncode == 0 and "--manifest" in r.stdout and "--controller-policy" in r.stdout
    assert run("--bogus").returncode == 1
    assert run("--repeats", "0", "--img-a", "a.ppm", "--img-b", "b.ppm").returncode == 1
    assert run("--seed", "12x", "--img-a", "a.ppm", "--img-b", "b.ppm").returncode == 1
    # A policy needs adaptive mode; adaptive mode needs a supported policy.
    assert run("--controller-policy", "legacy-v1", "--img-a", "a.ppm", "--img-b", "b.ppm").returncode == 1
    r = run("--adaptive", "--img-a", "a.ppm", "--img-b", "b.ppm", "--out", tmp_path / "r.csv")
    assert r.returncode == 1 and "requires --controller-policy" in r.stderr
    policy = tmp_path / "policy.json"
    policy.write_text("{}")
    r = run("--adaptive", "--controller-policy", policy, "--img-a", "a.ppm", "--img-b", "b.ppm",
            "--out", tmp_path / "r.csv")
    assert r.returncode == 1 and "schema_version" in r.stderr
    # A policy file's provenance seed must agree with --seed.
    import json
    from test_policies import V2
    doc = dict(V2, provenance={"seed": 42})
    policy.write_text(json.dumps(doc))
    r = run("--adaptive", "--controller-policy", policy, "--seed", "7", "--img-a", "a.ppm", "--img-b", "b.ppm",
            "--out", tmp_path / "r.csv")
    assert r.returncode == 1 and "provenance.seed" in r.stderr
    # Neither a manifest nor a pair; a manifest that is not an object.
    assert run("--out", tmp_path / "r.csv").returncode == 1
    bad = tmp_path / "m.json"
    bad.write_text("[1, 2]")
    assert run("--manifest", bad, "--out", tmp_path / "r.csv").returncode == 1
    # The output may not overwrite an input.
    assert run("--manifest", bad, "--out", bad).returncode == 1


def pair_files(tmp_path, n):
    from cs_stitch import pipeline as P
    entries = {}
    for i in range(n):
        a, b, _ = P.synth_pair(31000 + i, 640, 480, overlap=(0.5, 0.7))
        write_ppm(tmp_path / f"p{i}_a.ppm", a)
        write_ppm(tmp_path / f"p{i}_b.ppm", b)
        entries[f"pair_{i}"] = {"a": f"p{i}_a.ppm", "b": f"p{i}_b.ppm"}
    (tmp_path / "manifest.json").write_text(json.dumps(entries))
    return entries


def read_csv(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    header, body = rows[0], rows[1:]
    assert all(len(r) == len(header) for r in body)
    return [dict(zip(header, r)) for r in body]


@pytest.mark.gpu
def test_fixed_mode_rows_match_the_runner(tmp_path):
    from cs_stitch import pipeline as P
    entries = pair_files(tmp_path, 2)
    cfgs = [{"S": 1.0, "K": 800, "R": 512, "blend_mode": "multiband5"},
            {"config_id": "small", "S": 0.5, "K": 500, "R": 256, "blend_mode": "feather"}]
    (tmp_path / "configs.json").write_text(json.dumps(cfgs))
    out = tmp_path / "runs.csv"
    feats = tmp_path / "features.csv"
    r = run("--manifest", tmp_path / "manifest.json", "--configs", tmp_path / "configs.json", "--out", out,
            "--warmup", 1, "--repeats", 2, "--run-id", "t1", "--seed", 5, "--features-out", feats)
    assert r.returncode == 0, r.stderr
    rows = read_csv(out)
    frows = read_csv(feats)
    assert len(frows) == len(rows) and len(frows[0]) == 7 + 32 + 38
    assert len(rows) == 2 * 2 * 3
    assert [row["pair_id"] for row in rows[::6]] == ["pair_0", "pair_1"]
    assert {row["config_id"] for row in rows} == {"S1.00_K800_R512_Bmultiband5", "small"}
    assert [row["is_warmup"] for row in rows[:3]] == ["1", "0", "0"]
    runner = P.PipelineRunner(640, 480, max_k=4000, max_r=2048)
    runner.set_collect_features(True)
    for i, pid in enumerate(entries):
        a, b, _ = P.synth_pair(31000 + i, 640, 480, overlap=(0.5, 0.7))
        runner.upload_pair(a, b)
        for c in cfgs:
            rec = runner.execute_run(P.harness_config(S=c["S"], K=c["K"], R=c["R"], blend=c["blend_mode"]
                                                      .replace("multiband5", "multiband_5"), seed=5))
            got = [row for row in rows if row["pair_id"] == pid and row["S"].startswith(str(c["S"]))
                   and int(row["K"]) == c["K"]]
            assert len(got) == 3
            for row in got:
                assert row["status"] == rec.status.decode()
                assert (int(row["num_kp_a"]), int(row["num_kp_b"]), int(row["num_matches"]),
                        int(row["num_inliers"])) == (rec.num_kp_a, rec.num_kp_b, rec.num_matches, rec.num_inliers)
                if rec.status == b"OK":
                    assert [float(row[f"h{k}"]) for k in range(9)] == list(rec.H)
                    assert int(row["canvas_pixels"]) == rec.canvas_pixels
            # The observations CSV carries the runner's features to the last bit.
            scene, gate, present, _ = runner.features()
            fgot = [row for row in frows if row["pair_id"] == pid and row["config_id"] == got[0]["config_id"]]
            assert len(fgot) == 3
            for row in fgot:
                assert row["features_present"] == str(int(present))
                assert [float(row[f"scene_{k}"]) for k in range(32)] == list(scene)
                assert [float(row[f"gate_{k}"]) for k in range(38)] == list(gate)
    runner.close()


@pytest.mark.gpu
def test_adaptive_campaign_with_attempt_trace(tmp_path):
    pair_files(tmp_path, 1)
    out, trace = tmp_path / "runs.csv", tmp_path / "attempts.csv"
    r = run("--manifest", tmp_path / "manifest.json", "--adaptive", "--controller-policy", "legacy-v1",
            "--out", out, "--trace-out", trace, "--warmup", 0, "--repeats", 2)
    assert r.returncode == 0, r.stderr
    rows = read_csv(out)
    attempts = read_csv(trace)
    assert len(rows) == 2 and all(row["policy_id"] for row in rows)
    assert rows[0]["status"] == rows[1]["status"] and rows[0]["attempt_count"] == rows[1]["attempt_count"]
    n = int(rows[0]["attempt_count"])
    assert len(attempts) == 2 * n
    last = attempts[n - 1]
    assert last["terminal_status"] == rows[0]["status"]
    assert last["action"] == ("ACCEPT" if rows[0]["status"] == "OK" else "REJECT")
    assert all(a["action"] == "RETRY" for a in attempts[: n - 1])
    # The trace's last knobs name the run's config.
    assert rows[0]["config_id"].startswith(f"S{float(last['S']):.2f}_K{last['K']}_R{last['R']}")
    # A trace path must be fresh.
    assert run("--manifest", tmp_path / "manifest.json", "--adaptive", "--controller-policy", "legacy-v1",
               "--out", tmp_path / "r2.csv", "--trace-out", trace).returncode == 1


@pytest.mark.gpu
def test_adaptive_campaign_with_learned_policy_file(tmp_path):
    """A controller-policy-v3 file drives the adaptive campaign; the features
    CSV carries the scene and last-attempt gate observations the runner
    collected, equal to the Python runner's under the same policy."""
    import json
    from cs_stitch import pipeline as P
    from test_policies import gated_v3_policy
    pair_files(tmp_path, 1)
    policy = tmp_path / "policy.json"
    policy.write_text(json.dumps(gated_v3_policy()))
    out, trace, feats = tmp_path / "runs.csv", tmp_path / "attempts.csv", tmp_path / "features.csv"
    r = run("--manifest", tmp_path / "manifest.json", "--adaptive", "--controller-policy", policy, "--out", out,
            "--trace-out", trace, "--features-out", feats, "--warmup", 0, "--repeats", 1, "--seed", 5)
    assert r.returncode == 0, r.stderr
    rows, frows = read_csv(out), read_csv(feats)
    assert len(rows) == 1 and rows[0]["policy_id"] == "gated-v3"
    a, b, _ = P.synth_pair(31000, 640, 480, overlap=(0.5, 0.7))
    runner = P.PipelineRunner(640, 480, max_k=4000, max_r=2048)
    runner.upload_pair(a, b)
    rec = runner.execute_adaptive_run(P.Controller(policy_json=policy.read_text()), seed=5)
    scene, gate, present, _ = runner.features()
    assert rows[0]["status"] == rec.status.decode() and int(rows[0]["attempt_count"]) == rec.attempt_count
    row = frows[0]
    assert row["features_present"] == str(int(present))
    assert [float(row[f"scene_{k}"]) for k in range(32)] == list(scene)
    assert [float(row[f"gate_{k}"]) for k in range(38)] == list(gate)
    runner.close()
