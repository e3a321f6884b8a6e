"""Controller policy files (controller-policy-v2/v3/v4) and the learned
controller branches: selector and gate trees, retry targets, pooled boosted
models (src/controller/controller.cpp:40-152, policy.cpp:13-50, tree.hpp,
ensemble.hpp; harness/controller_policy.hpp:60-206).

Pinned three ways: the loader cases of the reference's
tests/controller/test_policy_loader.cpp (its JSON fixtures are data and are
restated here); closed-form walks worked out by hand; and, when the reference's
controller and loader compiled from their sources are present (oracle/_ref),
decision-for-decision agreement on random synthetic policies, scenes and
feedback.
"""
import copy
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref" / "libref.so"
needs_ref = pytest.mark.skipif(not REF.exists(), reason="reference controller build (oracle/_ref) not present")

V2 = {"schema_version": "controller-policy-v2", "policy_id": "synthetic-test-only", "probe_scale": 0.25,
      "ladder": [{"S": 0.75, "K": 500, "R": 128, "blend_mode": "feather"},
                 {"S": 1.0, "K": 2000, "R": 1024, "blend_mode": "feather"}],
      "self_checks": [{"min_inliers": 4, "min_inlier_ratio": 0, "min_inlier_coverage": 0, "max_residual_s1": 2},
                      {"min_inliers": 4, "min_inlier_ratio": 0, "min_inlier_coverage": 0, "max_residual_s1": 1e300}],
      "scene_rules": [{"min_pair_gradient": 0.05, "first_index": 0}, {"min_pair_gradient": 0, "first_index": 1}]}


def v3_fixture():
    p = copy.deepcopy(V2)
    p["schema_version"] = "controller-policy-v3"
    p["feature_schema"] = "controller-features-v1"
    p["selector"] = [{"feature": 0, "threshold": 0.1, "left": 1, "right": 2}, {"value": 0}, {"value": 1}]
    p["gates"] = [[{"value": 0}], [{"value": 1}]]
    return p


def v4_fixture():
    p = copy.deepcopy(V2)
    p["schema_version"] = "controller-policy-v4"
    p["feature_schema"] = "controller-features-v2"
    p.update(reference_index=1, pre_threshold=0.8, gate_threshold=0.95, extra_ms=0.5, pair_relations=False,
             relation_mask=0)
    sig = {"bias": 0, "transform": "sigmoid", "calibration": {"slope": 1, "intercept": 0}}
    p["models"] = {
        "pre_quality": dict(sig, features=39, trees=[[{"leaf": 5}]]),
        "post_quality": dict(sig, features=77, trees=[[{"leaf": 2}]]),
        "latency": {"features": 39, "bias": 0, "transform": "exp", "calibration": {"slope": 1, "intercept": 0},
                    "trees": [[{"feature": 31, "threshold": 0.5, "left": 1, "right": 2}, {"leaf": 0.69314718},
                               {"leaf": 2.30258509}]]}}
    return p


def accepted(doc) -> bool:
    from cs_stitch import pipeline as P
    try:
        P.Controller(policy_json=doc if isinstance(doc, str) else json.dumps(doc))
        return True
    except ValueError:
        return False


def ref_lib():
    ref = C.CDLL(str(REF))
    ref.ref_policy_check.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    ref.ref_policy_walk.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    return ref


def mutations():
    """(document, accepted?) per case of test_policy_loader.cpp:24-95."""
    out = [(V2, True)]
    p = copy.deepcopy(V2); del p["scene_rules"]; out.append((p, False))
    p = copy.deepcopy(V2); p["ladder"][0]["K"] = 500.5; out.append((p, False))
    p = copy.deepcopy(V2); p["ladder"][0]["K"] = 4001; out.append((p, False))
    p = copy.deepcopy(V2); p["scene_rules"][0]["first_index"] = 2; out.append((p, False))
    p = copy.deepcopy(V2); p["scene_rules"].pop(); out.append((p, False))
    p = copy.deepcopy(V2); p["self_checks"].pop(); out.append((p, False))
    p = copy.deepcopy(V2); p["ladder"][0]["blend_mode"] = "unknown"; out.append((p, False))
    p = copy.deepcopy(V2); p["ladder"] = (p["ladder"] * 9)[:17]; out.append((p, False))
    p = copy.deepcopy(V2); p["unexpected"] = None; out.append((p, False))
    p = copy.deepcopy(V2); p["residual_unit"] = "original_pixels"; out.append((p, False))
    p = copy.deepcopy(V2); p["provenance"] = {"seed": 42}; out.append((p, True))
    p = copy.deepcopy(V2); p["provenance"] = {"seed": 4.5}; out.append((p, False))
    out.append((v3_fixture(), True))
    p = v3_fixture(); p["selector"][0]["left"] = 0; out.append((p, False))
    p = v3_fixture(); p["gates"][0][0]["value"] = 2; out.append((p, False))
    p = v3_fixture(); p["retry_targets"] = [1, 2]; out.append((p, True))
    p = v3_fixture(); p["retry_targets"] = [0, 2]; out.append((p, False))  # must move forward
    p = v3_fixture(); p["feature_schema"] = "controller-features-v2"; out.append((p, False))
    out.append((v4_fixture(), True))
    p = v4_fixture(); p["models"]["latency"]["trees"][0][0]["left"] = 0; out.append((p, False))
    p = v4_fixture(); p["reference_index"] = 0; out.append((p, False))  # the reference rung is the last
    p = v4_fixture(); p["ladder"][1]["K"] = 1000; out.append((p, False))  # and the QoR baseline
    p = v4_fixture(); p["models"]["latency"]["transform"] = "sigmoid"; out.append((p, False))
    p = v4_fixture(); p["relation_mask"] = 8; out.append((p, False))
    return out


def test_loader_cases_of_the_reference_test():
    for doc, ok in mutations():
        assert accepted(doc) == ok, json.dumps(doc)[:200]
    for malformed in ('{"S":0.25,"S":0.75}', "{} {}"):
        assert not accepted(malformed)


@needs_ref
def test_loader_agrees_with_reference_loader():
    ref = ref_lib()
    err = C.create_string_buffer(256)
    for doc, ok in mutations():
        assert (ref.ref_policy_check(json.dumps(doc).encode(), err, 256) == 0) == ok


def healthy(matches=10, inliers=10, coverage=0.5, residual=0.1):
    from cs_stitch import _ffi as F
    h = F.HomographyResult()
    for i in range(9):
        h.H[i] = h.H_s[i] = 1.0 if i % 4 == 0 else 0.0
    h.status, h.num_matches, h.num_inliers = 0, matches, inliers
    h.inlier_ratio, h.inlier_coverage, h.reproj_residual = inliers / matches, coverage, residual
    return h


def test_pooled_walk_closed_form():
    """Latency exp(ln 2) = 2 ms at K/2000 < 0.5 and exp(ln 10) = 10 ms at the
    reference rung; pre-quality sigmoid(5) = 0.9933 >= 0.8, so rung 0 costs
    2 + 0.0067 * 10 + 0.5 < 10 and is chosen. Post-quality sigmoid(2) = 0.88
    is below the 0.95 gate: retry at the reference rung, which needs no gate
    features and accepts."""
    from cs_stitch import pipeline as P
    ctl = P.Controller(policy_json=json.dumps(v4_fixture()))
    assert ctl.observations() == (True, False, 0, True)  # before begin_pair the rung is 0
    scene = np.full(32, 0.1)
    rows = ctl.walk(.1, .1, 100, 100, 100, 100, [healthy(), healthy()], [0, 0], scene=scene,
                    gates=[np.zeros(38), None])
    assert rows == [(-1, -1, 0.75, 500, 128), (1, 9, 1.0, 2000, 1024), (0, 0, 1.0, 2000, 1024)]
    # Without scene features a pooled policy cannot start.
    assert ctl.walk(.1, .1, 100, 100, 100, 100, [healthy()], [0], scene=None) == []


def test_selector_and_gate_walk_closed_form():
    """The selector sends scene feature 0 <= 0.1 to rung 0; rung 0's gate
    always says 0 (reject), rung 1's says 1 (accept)."""
    from cs_stitch import pipeline as P
    ctl = P.Controller(policy_json=json.dumps(v3_fixture()))
    assert ctl.observations()[:3] == (True, False, 0)
    low, high = np.full(32, 0.05), np.full(32, 0.5)
    g = np.zeros(38)
    assert ctl.walk(.1, .1, 100, 100, 100, 100, [healthy(), healthy()], [0, 0], scene=low, gates=[g, g]) == \
        [(-1, -1, 0.75, 500, 128), (1, 9, 1.0, 2000, 1024), (0, 0, 1.0, 2000, 1024)]
    assert ctl.walk(.1, .1, 100, 100, 100, 100, [healthy()], [0], scene=high, gates=[g]) == \
        [(-1, -1, 1.0, 2000, 1024), (0, 0, 1.0, 2000, 1024)]
    # A passing self-check without gate features is rejected by the gate.
    assert ctl.walk(.1, .1, 100, 100, 100, 100, [healthy()], [0], scene=high, gates=[None]) == \
        [(-1, -1, 1.0, 2000, 1024), (2, 8, 1.0, 2000, 1024)]


# ------------------------------------------------------- random synthetic policies

def random_tree(rng, features, outputs, max_nodes=15):
    """Bounded tree in breadth-first order (children after their parent)."""
    nodes = [None]
    queue = [(0, 0)]
    while queue:
        i, depth = queue.pop(0)
        if depth < 3 and len(nodes) + 2 <= max_nodes and rng.random() < 0.7:
            left, right = len(nodes), len(nodes) + 1
            nodes += [None, None]
            nodes[i] = {"feature": int(rng.integers(0, features)), "threshold": float(rng.uniform(0, 1)),
                        "left": left, "right": right}
            queue += [(left, depth + 1), (right, depth + 1)]
        else:
            nodes[i] = {"value": int(rng.integers(0, outputs))}
    return nodes


def random_ensemble(rng, features, exponential):
    trees = []
    for _ in range(int(rng.integers(1, 6))):
        t = random_tree(rng, features, 1)
        for n in t:
            if "value" in n:
                del n["value"]
                n["leaf"] = float(rng.normal(0, 1))
        trees.append(t)
    return {"features": features, "bias": float(rng.normal(0, 0.5)), "transform": "exp" if exponential else "sigmoid",
            "calibration": {"slope": float(rng.uniform(0.5, 2)), "intercept": float(rng.normal(0, 0.5))},
            "trees": trees}


def random_ladder(rng, n, pooled):
    rungs = [(.25, 500, 128), (.5, 1000, 128), (.5, 1000, 256), (.75, 2000, 512), (.5, 2000, 256), (.75, 1000, 128)]
    pick = [rungs[i] for i in rng.permutation(len(rungs))[: n - 1 if pooled else n]]
    if pooled:
        pick.append((1.0, 2000, 1024))
    ladder = [{"S": s, "K": k, "R": r, "blend_mode": "feather"} for s, k, r in pick]
    checks = [{"min_inliers": int(rng.integers(4, 40)), "min_inlier_ratio": float(rng.uniform(0, .3)),
               "min_inlier_coverage": float(rng.uniform(0, .2)), "max_residual_s1": float(rng.uniform(1, 5))}
              for _ in range(n)]
    return ladder, checks


def random_policy(rng, kind):
    n = int(rng.integers(2, 6))
    ladder, checks = random_ladder(rng, n, kind == "v4")
    p = {"schema_version": f"controller-policy-{kind}", "policy_id": f"random-{kind}", "probe_scale": 0.25,
         "ladder": ladder, "self_checks": checks, "scene_rules": [{"min_pair_gradient": 0, "first_index": 0}]}
    if kind == "v3":
        p["feature_schema"] = "controller-features-v1"
        p["selector"] = random_tree(rng, 30, n)
        p["gates"] = [random_tree(rng, 24, 2) for _ in range(n)]
        if rng.random() < 0.5:
            p["retry_targets"] = [int(rng.integers(i + 1, n + 1)) for i in range(n)]
    else:
        p["feature_schema"] = "controller-features-v2"
        p.update(reference_index=n - 1, pre_threshold=float(rng.uniform(0.2, 0.9)),
                 gate_threshold=float(rng.uniform(0.2, 0.9)), extra_ms=float(rng.uniform(0, 2)),
                 pair_relations=bool(rng.random() < 0.5), relation_mask=int(rng.integers(0, 8)))
        p["models"] = {"pre_quality": random_ensemble(rng, 39, False), "post_quality": random_ensemble(rng, 77, False),
                       "latency": random_ensemble(rng, 39, True)}
    return p


def random_feedback(rng):
    kind = rng.integers(0, 8)
    if kind == 0:
        h = healthy(int(rng.integers(4, 9)), 4)
        h.status, h.num_matches = 4, int(rng.integers(0, 4))
        return h
    if kind == 1:
        h = healthy(int(rng.integers(4, 50)), 4)
        h.status = 5
        return h
    m = int(rng.integers(4, 400))
    return healthy(m, int(rng.integers(4, m + 1)), float(rng.uniform(0, .5)), float(rng.uniform(0, 5)))


@needs_ref
@pytest.mark.parametrize("kind", ["v3", "v4"])
def test_learned_walks_match_reference_controller(kind):
    from cs_stitch import pipeline as P
    ref = ref_lib()
    rng = np.random.default_rng(17 if kind == "v3" else 18)
    walks = 0
    for _ in range(60):
        doc = random_policy(rng, kind)
        text = json.dumps(doc)
        ctl = P.Controller(policy_json=text)
        err = C.create_string_buffer(256)
        assert ref.ref_policy_check(text.encode(), err, 256) == 0, err.value
        for _ in range(10):
            n = int(rng.integers(1, 6))
            fbs = [random_feedback(rng) for _ in range(n)]
            st = [0] * n
            scene = rng.uniform(0, 1, 32) if rng.random() < 0.9 else None
            gates = [rng.uniform(0, 1, 38) if rng.random() < 0.8 else None for _ in range(n)]
            mine = ctl.walk(.1, .1, 640, 480, 600, 400, fbs, st, scene=scene, gates=gates)
            arr = (P.F.HomographyResult * n)(*fbs)
            sts = (C.c_int * n)(*st)
            g = np.zeros((n, 38))
            present = (C.c_int * n)()
            for i, row in enumerate(gates):
                if row is not None:
                    g[i], present[i] = row, 1
            sc = np.ascontiguousarray(scene) if scene is not None else None
            trace = (C.c_double * (5 * (n + 2)))()
            rows = ref.ref_policy_walk(text.encode(), .1, .1, 640, 480, 600, 400,
                                       sc.ctypes.data if sc is not None else None, arr, sts, g.ctypes.data, present,
                                       n, trace)
            assert mine == [tuple(trace[5 * i:5 * i + 5]) for i in range(rows)], (text, mine)
            walks += 1
    assert walks == 600


def _ladder(rungs):
    return [{"S": s, "K": k, "R": r, "blend_mode": "feather"} for s, k, r in rungs]


def _checks(n):
    return [{"min_inliers": 8, "min_inlier_ratio": 0.05, "min_inlier_coverage": 0.0, "max_residual_s1": 6.0}] * n


def gated_v3_policy():
    """Selector always picks rung 0; rung 0's gate rejects whenever the
    photometric overlap share (gate feature 10) is finite, rung 1's accepts
    unless the photometric NCC (feature 13) is below -2. Exercises gate
    features on a retry."""
    rungs = [(.5, 1000, 128), (.75, 2000, 512), (1.0, 2000, 1024)]
    return {"schema_version": "controller-policy-v3", "policy_id": "gated-v3", "probe_scale": 0.25,
            "ladder": _ladder(rungs), "self_checks": _checks(3),
            "scene_rules": [{"min_pair_gradient": 0, "first_index": 0}], "feature_schema": "controller-features-v1",
            "selector": [{"value": 0}],
            "gates": [[{"feature": 10, "threshold": 2.0, "left": 1, "right": 2}, {"value": 0}, {"value": 1}],
                      [{"feature": 13, "threshold": -2.0, "left": 1, "right": 2}, {"value": 0}, {"value": 1}],
                      [{"value": 1}]]}


def gated_v4_policy():
    """Latency 1 ms for K/2000 <= 0.6 and 10 ms otherwise, pre-quality 0.993:
    the cheapest non-reference rung starts. The post-quality model accepts
    while the photometric overlap share (input 43 = gate feature 10) is
    finite. Pair and all alignment relations are collected."""
    rungs = [(.5, 1000, 128), (.75, 2000, 512), (1.0, 2000, 1024)]
    sig = {"bias": 0, "transform": "sigmoid", "calibration": {"slope": 1, "intercept": 0}}
    return {"schema_version": "controller-policy-v4", "policy_id": "gated-v4", "probe_scale": 0.25,
            "ladder": _ladder(rungs), "self_checks": _checks(3),
            "scene_rules": [{"min_pair_gradient": 0, "first_index": 0}], "feature_schema": "controller-features-v2",
            "reference_index": 2, "pre_threshold": 0.0, "gate_threshold": 0.5, "extra_ms": 0.5,
            "pair_relations": True, "relation_mask": 7,
            "models": {"pre_quality": dict(sig, features=39, trees=[[{"leaf": 5}]]),
                       "post_quality": dict(sig, features=77, trees=[[{"feature": 43, "threshold": 2.0, "left": 1,
                                                                       "right": 2}, {"leaf": 5}, {"leaf": -5}]]),
                       "latency": {"features": 39, "bias": 0, "transform": "exp",
                                   "calibration": {"slope": 1, "intercept": 0},
                                   "trees": [[{"feature": 31, "threshold": 0.6, "left": 1, "right": 2},
                                              {"leaf": 0.0}, {"leaf": 2.30258509}]]}}}


def test_gated_policies_load():
    assert accepted(gated_v3_policy()) and accepted(gated_v4_policy())
