"""Single-sample wire contract on the host (no GPU), through the library's command line
(femm_cli_main): request validation (io/sample_request.cpp:152-210), the artifact checks that
run before any solve (app/sample_adapter.cpp:364-392), the failure response layout
(:226-327) and the publish rules (:442-487).  The expected failure response is the one the
reference's CLI tests build (gpu_solver/tests/test_femm_gpu_cli.py:60-76)."""

import ctypes
import hashlib
import json
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

import femm_fixtures as ff
from femm_gpu_amd import sample
from femm_gpu_amd.solver import library

ROOT = Path(__file__).resolve().parents[1]
ARTIFACT = ff.GOLDEN / "periodic_strip_v1" / "periodic_strip.gpu.json"
CLI = ROOT / "femm-gpu_amd" / "bin" / "femm_gpu_amd"

# Keys and order of a planar-DC response (test_femm_gpu_cli.py:60-76).
PLANAR_KEYS = ["protocol", "mesh_artifact_sha256", "status", "solve_status", "postprocess_status",
               "error_identifier", "error_message", "Fx_N", "Fy_N", "torque_Nm",
               "actual_circuit_currents_A", "circuit_flux_linkage_Wb", "airgap_sample_angles_deg",
               "airgap_radial_flux_density_T", "mesh_node_count", "mesh_element_count",
               "field_solution_included", "node_A_Wb_per_m", "element_Bx_T", "element_By_T",
               "convergence"]
GENERIC_KEYS = PLANAR_KEYS[:2] + ["problem_type", "field_components", "node_potential_quantity",
                                  "node_potential_unit"] + PLANAR_KEYS[2:17] + \
    ["node_potential", "element_B_component_1_T", "element_B_component_2_T", "convergence"]


def _request(path, sha, **kw):
    r = dict(protocol=sample.GENERIC, mesh_artifact_path=str(path), mesh_artifact_sha256=sha,
             circuit_currents_A=[1.0], force_group_number=-1, stress_air_group_number=-1,
             airgap_radius_mm=0.0, airgap_angles_deg=[], sliding_band_angle_deg=0.0,
             compute_force_torque=False, include_field_solution=False)
    r.update(kw)
    return r


def _sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def _run(tmp_path, text):
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(text, encoding="utf-8")
    code = sample.run_sample(str(req), str(resp))
    return code, json.loads(resp.read_text(encoding="utf-8"))


def _gpu_present():
    text = ctypes.create_string_buffer(256)
    return library().femm_device_info(text, 256) == 0


def _parse_failure(doc):
    """The response of a request that failed to parse (no trusted request identity)."""
    assert list(doc) == PLANAR_KEYS
    assert doc["protocol"] == sample.PLANAR_DC and doc["mesh_artifact_sha256"] == ""
    assert (doc["status"], doc["solve_status"], doc["postprocess_status"]) == \
        ("FAIL", "INVALID_ARGUMENT", "NOT_REQUESTED")
    assert doc["error_identifier"] == "GPU_FEMM_INVALID_ARGUMENT"
    assert doc["Fx_N"] is None and doc["node_A_Wb_per_m"] == [] and doc["mesh_node_count"] == 0
    assert doc["convergence"] == {"iterations": 0, "residual_l2": None}


def test_a_valid_request_reaches_the_artifact_checks(tmp_path):
    code, doc = _run(tmp_path, json.dumps(_request(
        tmp_path / "absent.json", "a" * 64, force_group_number=2, stress_air_group_number=0,
        protocol=sample.PLANAR_DC)))
    assert code == 1
    assert (doc["solve_status"], doc["error_message"]) == ("INPUT_IO", "cannot open mesh artifact")
    assert doc["mesh_artifact_sha256"] == "a" * 64


@pytest.mark.parametrize("change", [
    dict(protocol="gpu_femm_single_sample_v1"),
    dict(mesh_artifact_path=""),
    dict(mesh_artifact_sha256="A" * 64),                  # upper case
    dict(mesh_artifact_sha256="a" * 63),
    dict(force_group_number=1),                           # groups come in pairs
    dict(force_group_number=1, stress_air_group_number=1),
    dict(force_group_number=-2, stress_air_group_number=-2),
    dict(force_group_number=1.5, stress_air_group_number=0),
    dict(force_group_number=True, stress_air_group_number=0),
    dict(airgap_radius_mm=-1.0),
    dict(airgap_angles_deg=[0.0, 90.0]),                  # angles need a radius > 0
    dict(circuit_currents_A=[1.0, "2"]),
    dict(compute_force_torque=1),
    dict(include_field_solution="yes"),
])
def test_invalid_fields_get_the_parse_failure_response(tmp_path, change):
    code, doc = _run(tmp_path, json.dumps(_request("m.json", "a" * 64, **change)))
    assert code == 1
    _parse_failure(doc)
    # The first reason found: a member of the wrong JSON type, or the request rules.
    assert doc["error_message"].startswith(("wrong JSON type: ",
                                            "invalid gpu_femm_planar_dc_sample_v1 request"))


@pytest.mark.parametrize("text,reason", [
    ("{not json", "JSON object key expected"),
    ("[1, 2]", "unexpected JSON object fields"),
    ("{}  x", "trailing JSON data"),
])
def test_malformed_documents(tmp_path, text, reason):
    code, doc = _run(tmp_path, text)
    assert code == 1
    _parse_failure(doc)
    assert doc["error_message"] == reason


def test_unknown_missing_duplicate_and_nonstandard_members(tmp_path):
    good = _request("m.json", "a" * 64)
    cases = [json.dumps(dict(good, extra=1)),
             json.dumps({k: v for k, v in good.items() if k != "sliding_band_angle_deg"}),
             json.dumps(good)[:-1] + ', "protocol": "%s"}' % sample.GENERIC,
             json.dumps(good).replace("0.0", "NaN", 1),
             json.dumps(good).replace('"m.json"', '"m\\u002ejson"')]
    reasons = ["unexpected JSON object fields", "unexpected JSON object fields",
               "duplicate JSON object key", "invalid JSON value", "unsupported JSON string escape"]
    for text, reason in zip(cases, reasons):
        code, doc = _run(tmp_path, text)
        assert code == 1
        _parse_failure(doc)
        assert doc["error_message"] == reason


@pytest.mark.parametrize("case", ["sha", "missing", "physics", "circuits", "postprocess",
                                  "sliding"])
def test_artifact_checks_fail_before_any_solve(tmp_path, case):
    art = tmp_path / "mesh.gpu.json"
    shutil.copyfile(ARTIFACT, art)
    sha = _sha(art)
    kw = {}
    if case == "sha":
        sha = "0" * 64
    elif case == "missing":
        art = tmp_path / "absent.gpu.json"
    elif case == "physics":
        kw = dict(protocol=sample.PLANAR_DC)        # planar-DC request on a generic artifact
    elif case == "circuits":
        kw = dict(circuit_currents_A=[1.0, 2.0])
    elif case == "postprocess":
        kw = dict(compute_force_torque=True)        # the generic protocol solves fields only
    else:
        kw = dict(sliding_band_angle_deg=7.5)
    code, doc = _run(tmp_path, json.dumps(_request(art, sha, **kw)))
    assert code == 1
    mismatch = "artifact physics, circuit count, group, or sliding-band request mismatch"
    expected = {"sha": ("INVALID_ARGUMENT", "mesh artifact SHA mismatch"),
                "missing": ("INPUT_IO", "cannot open mesh artifact")}.get(
                    case, ("INVALID_ARGUMENT", mismatch))
    assert (doc["solve_status"], doc["error_message"]) == expected
    assert doc["status"] == "FAIL" and doc["error_identifier"] == "GPU_FEMM_" + expected[0]
    assert doc["convergence"]["iterations"] == 0
    assert list(doc) == (PLANAR_KEYS if case == "physics" else GENERIC_KEYS)


def test_a_rejected_artifact_reports_the_parser_reason(tmp_path):
    art = tmp_path / "mesh.gpu.json"
    doc = json.loads(ARTIFACT.read_text(encoding="utf-8"))
    doc["resolved"]["outer_dirichlet"]["A_Wb_per_m"][0] = 1e-3
    art.write_text(json.dumps(doc), encoding="utf-8")
    code, resp = _run(tmp_path, json.dumps(_request(art, _sha(art))))
    assert code == 1 and resp["solve_status"] == "INVALID_ARGUMENT"
    assert resp["error_message"] == "canonical artifact identity mismatch"


def test_response_may_not_overwrite_the_request_or_the_artifact(tmp_path):
    art = tmp_path / "mesh.gpu.json"
    shutil.copyfile(ARTIFACT, art)
    req = tmp_path / "request.json"
    text = json.dumps(_request(art, _sha(art)))
    req.write_text(text, encoding="utf-8")
    assert sample.run_sample(str(req), str(req)) == 1
    assert req.read_text(encoding="utf-8") == text
    before = art.read_bytes()
    assert sample.run_sample(str(req), str(art)) == 1
    assert art.read_bytes() == before
    assert sample.main(["only-one-argument"]) == 2
    assert list(tmp_path.glob("*.tmp.*")) == []


def test_solver_failure_is_reported_in_the_response(tmp_path):
    """A valid request on a host without a GPU: the library's GPU_UNAVAILABLE becomes the
    response's solve status (no solve result, as when the reference's Initialize fails)."""
    if _gpu_present():
        pytest.skip("a GPU is present")
    art = tmp_path / "mesh.gpu.json"
    shutil.copyfile(ARTIFACT, art)
    code, doc = _run(tmp_path, json.dumps(_request(art, _sha(art))))
    assert code == 1
    assert (doc["status"], doc["solve_status"], doc["error_identifier"], doc["error_message"]) == \
        ("FAIL", "GPU_UNAVAILABLE", "GPU_FEMM_GPU_UNAVAILABLE", "GPU_UNAVAILABLE")
    assert doc["problem_type"] == "planar" and doc["convergence"]["iterations"] == 0
    assert list(tmp_path.glob("*.tmp.*")) == []


@pytest.mark.parametrize("front", ["python", "executable"])
def test_command_line_exit_codes(tmp_path, front):
    """2 for a wrong command line, 1 with a published FAIL response for an invalid request
    (the reference CLI's exit contract, app/cli.cpp:20-181); the executable and
    `python -m femm_gpu_amd.sample` are the same command line."""
    if front == "python":
        env = dict(os.environ, PYTHONPATH=str(ROOT / "femm-gpu_amd"))
        cmd = [sys.executable, "-m", "femm_gpu_amd.sample"]
    else:
        env, cmd = dict(os.environ), [str(CLI)]
        if not CLI.exists():
            pytest.skip("command-line executable not built")
    assert subprocess.run(cmd, env=env, capture_output=True).returncode == 2
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(_request("m.json", "a" * 64, extra=1)), encoding="utf-8")
    args = [str(req), str(resp)] if front == "python" else ["--solve", str(req), str(resp)]
    done = subprocess.run(cmd + args, env=env, capture_output=True, text=True)
    assert done.returncode == 1 and "FAIL planar-DC request" in done.stderr
    assert json.loads(resp.read_text(encoding="utf-8"))["status"] == "FAIL"
    caps = subprocess.run(cmd + ["--capabilities"], env=env, capture_output=True, text=True)
    assert caps.returncode == 0
    assert "gpu_femm_motor_batch_v1" in json.loads(caps.stdout)["motor_batch_protocols"]


def test_unwritable_response_location_fails_cleanly(tmp_path):
    req = tmp_path / "request.json"
    req.write_text("{}", encoding="utf-8")
    assert sample.run_sample(str(req), str(tmp_path / "missing_dir" / "response.json")) == 1
