"""The wire contract's host side without a GPU: SHA-256, the strict mesh-artifact parser and
the motor single-sample and batch protocols, all through the library's C ABI
(include/femm_wire.h).

Pins: SHA-256 against the FIPS 180-2 example digests; the canonical identity against the
identities the reference's own preparer wrote into the golden artifacts (every one of them
must parse); the parsed model against the reference's parser (tests/test_oracle.py); each
rejection rule against a document that breaks only that rule (gpu_solver/tests/native/
protocol_contract_tests.cu); the batch parse rules of app/batch_prepare.cpp:74-131 and the
response of app/batch_response.cpp:140-371 (PARTIAL_FAILURE, items in request order)."""

import copy
import ctypes
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import femm_fixtures as ff
from femm_gpu_amd import sample, wire
from femm_gpu_amd.solver import library

GENERIC = ff.GOLDEN / "periodic_strip_v1" / "periodic_strip.gpu.json"

MOTOR_KEYS = ["protocol", "mesh_artifact_sha256", "base_motor_fem_sha256", "source_fem_sha256",
              "status", "solve_status", "error_identifier", "error_message", "Fx_N", "Fy_N",
              "torque_Nm", "actual_circuit_currents_A", "circuit_flux_linkage_Wb",
              "airgap_sample_angles_deg", "airgap_radial_flux_density_T", "mesh_element_count",
              "convergence"]
BATCH_KEYS = ["protocol", "status", "solve_status", "error_identifier", "error_message", "cache",
              "items"]


def _gpu_present():
    text = ctypes.create_string_buffer(256)
    return library().femm_device_info(text, 256) == 0


# ---- SHA-256 -------------------------------------------------------------------------------

@pytest.mark.parametrize("message,digest", [
    (b"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
    (b"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
    (b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
     "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"),
])
def test_sha256_fips_examples(message, digest):
    assert wire.sha256_hex(message) == digest


def test_sha256_block_boundaries_and_a_million_a():
    for n in (55, 56, 63, 64, 65, 119, 120, 128, 1000):
        data = bytes((7 * k + 3) % 256 for k in range(n))
        assert wire.sha256_hex(data) == hashlib.sha256(data).hexdigest()
    assert wire.sha256_hex(b"a" * 1000000) == \
        "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0"


_SHA_CHECK = '''
import hashlib, random, sys
sys.path.insert(0, sys.argv[1])
from femm_gpu_amd import wire
rng = random.Random(5)
for n in list(range(0, 260)) + [4095, 4096, 4097, 100003]:
    data = bytes(rng.randrange(256) for _ in range(n))
    assert wire.sha256_hex(data) == hashlib.sha256(data).hexdigest(), n
print("ok")
'''


@pytest.mark.parametrize("portable", ["0", "1"])
def test_sha256_both_compressions_match_hashlib(portable):
    """Messages of 0-259 bytes and around block multiples hash as hashlib does, with the x86
    SHA-extension compression (where the CPU has it) and with the portable one
    (FEMM_GPU_AMD_PORTABLE_SHA=1); each in a fresh process, since the choice is made once."""
    import os
    import subprocess
    import sys
    env = dict(os.environ, FEMM_GPU_AMD_PORTABLE_SHA=portable)
    pkg = str(Path(wire.__file__).resolve().parents[1])
    out = subprocess.run([sys.executable, "-c", _SHA_CHECK, pkg], env=env, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr


# ---- artifacts -----------------------------------------------------------------------------

@pytest.mark.parametrize("path", sorted(ff.GOLDEN.glob("*/*.gpu.json")), ids=lambda p: p.stem)
def test_golden_artifacts_parse_with_their_canonical_identity(path):
    """Each golden artifact carries the canonical identity the reference's preparer computed;
    the strict parser recomputes and checks it."""
    model, currents = wire.load_artifact(path)
    assert model.node_count > 0 and len(currents) == model.circuit_count


def _doc(path=GENERIC):
    return json.loads(Path(path).read_text(encoding="utf-8"))


def _with_identity(doc):
    canonical = json.dumps(doc["resolved"], sort_keys=True, separators=(",", ":"))
    doc["canonical_identity_sha256"] = hashlib.sha256(canonical.encode()).hexdigest()
    return json.dumps(doc)


def _reject(text, reason):
    with pytest.raises(wire.ArtifactError) as err:
        wire.parse_artifact(text)
    assert str(err.value) == reason


def test_identity_binds_the_number_spelling():
    """The identity hashes numbers as written: 0.2 and 0.20 parse to the same model, but only
    the spelling the identity was computed on is accepted."""
    text = GENERIC.read_text(encoding="utf-8")
    doc = _doc()
    first = json.dumps(doc["resolved"]["model"]["depth_mm"])
    assert first in text
    respelled = text.replace('"depth_mm": %s' % first, '"depth_mm": %s0' % first, 1)
    if respelled == text:
        respelled = text.replace('"depth_mm":%s' % first, '"depth_mm":%s0' % first, 1)
    assert respelled != text
    _reject(respelled, "canonical artifact identity mismatch")


def _mutations():
    """(name, mutate(doc) -> None or text, expected reason) on the generic golden artifact."""
    def res(f):
        def apply(doc):
            f(doc["resolved"])
            return _with_identity(doc)
        return apply

    def mat(key, value):
        return res(lambda r: r["materials"][0].__setitem__(key, value))

    def reg(key, value):
        return res(lambda r: r["regions"][0].__setitem__(key, value))

    yield "unknown_root_member", lambda d: json.dumps(dict(d, extra=1)), \
        "unexpected JSON object fields"
    yield "unknown_resolved_member", res(lambda r: r.__setitem__("extra", 1)), \
        "unexpected JSON object fields"
    yield "missing_member", res(lambda r: r.pop("node_constraints")), \
        "unexpected JSON object fields"
    yield "identity_upper_case", lambda d: json.dumps(
        dict(d, canonical_identity_sha256=d["canonical_identity_sha256"].upper())), \
        "invalid GPU FEMM mesh artifact header"
    yield "identity_mismatch", lambda d: json.dumps(dict(d, canonical_identity_sha256="0" * 64)), \
        "canonical artifact identity mismatch"
    yield "source_mismatch", res(lambda r: r.__setitem__("source_fem_sha256", "c" * 64)), \
        "source FEM SHA mismatch"
    yield "frequency", res(lambda r: r["model"].__setitem__("frequency_hz", 50)), \
        "unsupported model physics"
    yield "zero_depth", res(lambda r: r["model"].__setitem__("depth_mm", 0)), \
        "unsupported model physics"
    yield "nonzero_dirichlet", res(lambda r: r["outer_dirichlet"]["A_Wb_per_m"].__setitem__(0, 1e-6)), \
        "outer Dirichlet must be zero and in range"
    yield "dirichlet_range", res(lambda r: r["outer_dirichlet"]["node_indices"].__setitem__(0, 10**6)), \
        "outer Dirichlet must be zero and in range"
    yield "anisotropic", mat("mu_y", 2.0), "invalid material definition"
    yield "laminated", mat("lam_type", 1), "invalid material definition"
    yield "lam_fill", mat("lam_fill", 0.95), "invalid material definition"
    yield "negative_hc", mat("H_c_A_per_m", -1.0), "invalid material definition"
    yield "material_name", mat("name", ""), "invalid material definition"
    yield "bh_lengths", mat("B_T", [0.0, 1.0]), "invalid material definition"
    yield "bh_decreasing", res(lambda r: r["materials"][0].update(B_T=[0.0, 1.0, 0.5],
                                                                  H_A_per_m=[0.0, 10.0, 20.0])), \
        "invalid B-H point"
    yield "bh_spline", res(lambda r: r["materials"][0].update(B_T=[0.1, 1.0],
                                                              H_A_per_m=[1.0, 10.0])), \
        "invalid B-H spline"
    yield "circuit_order", res(lambda r: r["circuits"][0].__setitem__("index", 1)), \
        "circuits must be ordered contiguous current-driven series circuits"
    yield "circuit_type", res(lambda r: r["circuits"][0].__setitem__("type", "parallel")), \
        "circuits must be ordered contiguous current-driven series circuits"
    yield "turns_without_circuit", reg("circuit_index", -1), "invalid region definition"
    yield "circuit_without_turns", res(lambda r: [g.__setitem__("turns", 0.0) for g in r["regions"]
                                                  if g["circuit_index"] >= 0]), \
        "invalid region definition"
    yield "unknown_material", reg("material_id", 999), "invalid region definition"
    yield "negative_group", reg("group_number", -1), "invalid region definition"
    yield "clockwise", res(lambda r: r["triangles"]["node_indices"][0].reverse()), \
        "triangle is not CCW"
    yield "unknown_region", res(lambda r: r["triangles"]["region_ids"].__setitem__(0, 999)), \
        "triangle references unknown region"
    yield "node_range", res(lambda r: r["triangles"]["node_indices"][0].__setitem__(0, 10**6)), \
        "triangle node out of range"
    yield "constraint_relation", res(lambda r: r["node_constraints"][0].__setitem__("relation", "mirror")), \
        "invalid periodic node constraint"
    yield "constraint_self", res(lambda r: r["node_constraints"][0].__setitem__(
        "node_b", r["node_constraints"][0]["node_a"])), "invalid periodic node constraint"
    # a planar-DC document relabelled as a sliding-band motor mesh lacks the motor header
    yield "sliding_band_schema", lambda d: json.dumps(dict(d, schema_version="gpu_femm_mesh_v2")), \
        "unexpected JSON object fields"
    yield "sector_schema", lambda d: json.dumps(dict(d, schema_version="gpu_femm_mesh_v4")), \
        "unsupported artifact schema: sector motor meshes"
    yield "duplicate_key", lambda d: json.dumps(d)[:-1] + ', "schema_version": "x"}', \
        "duplicate JSON object key"
    yield "trailing", lambda d: json.dumps(d) + " 0", "trailing JSON data"
    yield "nonfinite", lambda d: json.dumps(d).replace('"frequency_hz": 0', '"frequency_hz": 1e999'), \
        "nonfinite JSON number"


@pytest.mark.parametrize("name,mutate,reason", list(_mutations()), ids=[m[0] for m in _mutations()])
def test_artifact_rules(name, mutate, reason):
    doc = copy.deepcopy(_doc())
    text = mutate(doc)
    _reject(text, reason)


def test_region_without_area_is_rejected():
    doc = _doc()
    r = doc["resolved"]
    r["regions"].append(dict(id=99, group_number=7, material_id=r["materials"][0]["id"],
                             circuit_index=-1, turns=0.0, pm_magnetization_deg=0.0))
    _reject(_with_identity(doc), "region has no triangle area")


def test_planar_dc_schema_is_planar_only_and_has_no_constraints():
    doc = _doc(GENERIC)
    doc["schema_version"] = "gpu_femm_planar_dc_mesh_v1"
    _reject(_with_identity(doc), "unexpected JSON object fields")     # node_constraints
    del doc["resolved"]["node_constraints"]
    doc["resolved"]["model"]["problem_type"] = "axisymmetric"
    _reject(_with_identity(doc), "unsupported model physics")


def test_motor_artifact_maps_labels_identities_and_pose(tmp_path):
    model, _, current, _ = ff.nonlinear_pm_coil()
    art = tmp_path / "motor.gpu.json"
    ff.write_artifact(art, model, [current], schema="gpu_femm_mesh_v1", turns={2: 50.0},
                      rotor_angle_deg=3.0, displacement_mm=(0.1, 0.0))
    parsed, currents = wire.load_artifact(art)
    assert currents.tolist() == [current]
    assert parsed.node_count == model.node_count and parsed.triangle_count == model.triangle_count
    # base materials first, then one label per region (group = label of the writer)
    labels = parsed.materials[len(model.materials):]
    assert [m.group for m in labels] == list(range(len(model.materials)))
    assert [m.circuit for m in labels] == [m.circuit for m in model.materials]
    coil = labels[2]
    area = float(model.element_areas()[model.triangle_material == 2].sum())
    assert abs(coil.source_j_per_a - 50.0 / area) <= 1e-9 * coil.source_j_per_a
    doc = json.loads(art.read_text())
    doc["resolved"]["base_motor_fem_sha256"] = "c" * 64
    _reject(json.dumps(doc), "source FEM SHA mismatch")
    doc = json.loads(art.read_text())
    doc["resolved"]["pose"]["displacement_mm"] = [0.1]
    _reject(json.dumps(doc), "invalid pose")


# ---- motor single sample and batch ---------------------------------------------------------

def _motor_setup(tmp_path):
    model, _, current, _ = ff.nonlinear_pm_coil()
    post = ff.nonlinear_pm_coil_postprocess()
    art = tmp_path / "pm_coil.motor.gpu.json"
    sha = ff.write_artifact(art, model, [current], schema="gpu_femm_mesh_v1", turns={2: 50.0})
    return model, current, post, art, sha


def _motor_request(art, sha, current, post, **kw):
    r = dict(protocol=sample.MOTOR, mesh_artifact_path=str(art), mesh_artifact_sha256=sha,
             base_motor_fem_sha256="b" * 64, source_fem_sha256="a" * 64,
             circuit_currents_A=[current], selected_group_number=post["selected_label"],
             air_group_number=3, airgap_radius_mm=post["radius_m"] * 1e3,
             airgap_angles_deg=[float(a) * 180.0 / np.pi for a in post["angles"]],
             rotor_angle_deg=0.0, displacement_mm=[0.0, 0.0])
    r.update(kw)
    return r


def test_motor_single_sample_checks_and_failure_layout(tmp_path):
    model, current, post, art, sha = _motor_setup(tmp_path)
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    for kw, status in ((dict(rotor_angle_deg=1.0), "INVALID_ARGUMENT"),       # pose mismatch
                       (dict(source_fem_sha256="c" * 64), "INVALID_ARGUMENT"),
                       (dict(circuit_currents_A=[1.0, 2.0]), "INVALID_ARGUMENT"),
                       (dict(mesh_artifact_sha256="0" * 64), "INVALID_ARGUMENT"),
                       (dict(mesh_artifact_path=str(tmp_path / "absent")), "INPUT_IO")):
        req.write_text(json.dumps(_motor_request(art, sha, current, post, **kw)))
        assert sample.run_motor_sample(str(req), str(resp)) == 1
        doc = json.loads(resp.read_text())
        assert list(doc) == MOTOR_KEYS
        assert (doc["status"], doc["solve_status"]) == ("FAIL", status)
        assert doc["Fx_N"] is None and doc["convergence"] == {"iterations": 0, "residual_l2": None}
    # an invalid request: the anonymous failure response
    req.write_text(json.dumps(_motor_request(art, sha, current, post, selected_group_number=3)))
    assert sample.run_motor_sample(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text())
    assert doc["protocol"] == sample.MOTOR and doc["mesh_artifact_sha256"] == ""
    # the upper-case artifact SHA is accepted by the motor protocol (any-case hex)
    req.write_text(json.dumps(_motor_request(art, sha.upper(), current, post)))
    sample.run_motor_sample(str(req), str(resp))
    doc = json.loads(resp.read_text())
    assert doc["solve_status"] != "INVALID_ARGUMENT" or doc["error_message"] == "INVALID_ARGUMENT"


def _batch(items, chunk=4, protocol=sample.MOTOR_BATCH):
    return dict(protocol=protocol, max_items_per_chunk=chunk,
                items=[dict(task_id=t, request=r) for t, r in items])


@pytest.mark.parametrize("mutate", [
    lambda b: b.update(protocol="gpu_femm_motor_batch_v9"),
    lambda b: b.update(max_items_per_chunk=0),
    lambda b: b.update(max_items_per_chunk=4097),
    lambda b: b.update(max_items_per_chunk=1.5),
    lambda b: b.update(items=[]),
    lambda b: b.update(extra=1),
    lambda b: b["items"][1].update(task_id=b["items"][0]["task_id"]),       # duplicate id
    lambda b: b["items"][0].update(task_id="bad id"),
    lambda b: b["items"][0].update(task_id="x" * 257),
    lambda b: b["items"][0].update(extra=1),
    lambda b: b["items"][0]["request"].update(protocol="gpu_femm_motor_sample_v2"),
    lambda b: b["items"][0]["request"].update(selected_group_number=-1),
    lambda b: b.update(protocol="gpu_femm_motor_batch_v2"),
], ids=["protocol", "chunk0", "chunk4097", "chunk_fraction", "empty", "member", "duplicate_id",
        "id_chars", "id_length", "item_member", "item_protocol", "item_group", "batch_v2"])
def test_batch_parse_failures_write_no_response(tmp_path, mutate):
    model, current, post, art, sha = _motor_setup(tmp_path)
    r = _motor_request(art, sha, current, post)
    batch = _batch([("a", dict(r)), ("b", dict(r))])
    mutate(batch)
    req, resp = tmp_path / "batch.json", tmp_path / "batch.response.json"
    req.write_text(json.dumps(batch))
    assert sample.run_motor_batch(str(req), str(resp)) == 1
    assert not resp.exists()


def test_batch_item_cap(tmp_path):
    model, current, post, art, sha = _motor_setup(tmp_path)
    r = _motor_request(art, sha, current, post)
    req, resp = tmp_path / "batch.json", tmp_path / "batch.response.json"
    req.write_text(json.dumps(_batch([(f"t{k}", r) for k in range(4097)])))
    assert sample.run_motor_batch(str(req), str(resp)) == 1 and not resp.exists()


def test_batch_response_order_statuses_and_partial_failure(tmp_path):
    """Six items, chunks of four: three valid (one artifact, one plan), a SHA mismatch, a pose
    mismatch and a missing artifact.  Items keep the request order; the invalid ones fail
    before any solve with their reasons; the batch is a PARTIAL_FAILURE.  Without a GPU the
    valid items report GPU_UNAVAILABLE; with one they pass."""
    model, current, post, art, sha = _motor_setup(tmp_path)
    r = _motor_request(art, sha, current, post)
    items = [("ok-1", r), ("sha", dict(r, mesh_artifact_sha256="0" * 64)),
             ("ok.2", dict(r, circuit_currents_A=[-current])),
             ("pose", dict(r, displacement_mm=[0.5, 0.0])),
             ("missing", dict(r, mesh_artifact_path=str(tmp_path / "absent.json"))),
             ("ok_3", dict(r, circuit_currents_A=[0.5 * current]))]
    req, resp = tmp_path / "batch.json", tmp_path / "batch.response.json"
    req.write_text(json.dumps(_batch(items, chunk=4)))
    assert sample.run_motor_batch(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text())
    assert list(doc) == BATCH_KEYS
    assert (doc["status"], doc["solve_status"], doc["error_identifier"], doc["error_message"]) == \
        ("FAIL", "PARTIAL_FAILURE", "GPU_FEMM_BATCH_PARTIAL_FAILURE",
         "one or more batch items failed")
    assert [i["task_id"] for i in doc["items"]] == [t for t, _ in items]
    for item in doc["items"]:
        assert list(item) == ["task_id", "response"]
        assert list(item["response"]) == MOTOR_KEYS
    status = {i["task_id"]: i["response"]["solve_status"] for i in doc["items"]}
    assert (status["sha"], status["pose"], status["missing"]) == \
        ("INVALID_ARGUMENT", "INVALID_ARGUMENT", "INPUT_IO")
    valid = "OK" if _gpu_present() else "GPU_UNAVAILABLE"
    assert status["ok-1"] == status["ok.2"] == status["ok_3"] == valid
    cache = doc["cache"]
    assert (cache["chunk_count"], cache["requested_chunk_size"]) == (2, 4)
    assert cache["effective_chunk_size"] == 4
    # one plan for the artifact: built in the first chunk, reused in the second
    assert cache["mesh_cache_hits"] == 2 and cache["csr_symbolic_cache_hits"] == 1
    assert cache["batched_pcg_launches"] == 2 and cache["actual_parallel_width"] == 2
