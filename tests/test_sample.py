"""Single-sample wire contract on the host (no GPU): request validation
(io/sample_request.cpp:152-210), the artifact checks that run before any solve
(app/sample_adapter.cpp:364-392), the response layout (:226-327) and the publish rules
(:442-487).  The expected failure response is the one the reference's CLI tests build
(gpu_solver/tests/test_femm_gpu_cli.py:60-76)."""

import hashlib
import json
import shutil
from pathlib import Path

import numpy as np
import pytest

import femm_fixtures as ff
from femm_gpu_amd import sample
from femm_gpu_amd.model import load_artifact

ARTIFACT = ff.GOLDEN / "periodic_strip_v1" / "periodic_strip.gpu.json"

# Keys and order of a planar-DC response (test_femm_gpu_cli.py:60-76).
PLANAR_KEYS = ["protocol", "mesh_artifact_sha256", "status", "solve_status", "postprocess_status",
               "error_identifier", "error_message", "Fx_N", "Fy_N", "torque_Nm",
               "actual_circuit_currents_A", "circuit_flux_linkage_Wb", "airgap_sample_angles_deg",
               "airgap_radial_flux_density_T", "mesh_node_count", "mesh_element_count",
               "field_solution_included", "node_A_Wb_per_m", "element_Bx_T", "element_By_T",
               "convergence"]


def _request(path, sha, **kw):
    r = dict(protocol=sample.GENERIC, mesh_artifact_path=str(path), mesh_artifact_sha256=sha,
             circuit_currents_A=[1.0], force_group_number=-1, stress_air_group_number=-1,
             airgap_radius_mm=0.0, airgap_angles_deg=[], sliding_band_angle_deg=0.0,
             compute_force_torque=False, include_field_solution=False)
    r.update(kw)
    return r


def _sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def test_parse_accepts_a_valid_request():
    r = sample.parse_request(json.dumps(_request("m.json", "a" * 64, force_group_number=2,
                                                 stress_air_group_number=0,
                                                 protocol=sample.PLANAR_DC)))
    assert r["force_group_number"] == 2 and r["circuit_currents_A"] == [1.0]


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
def test_parse_rejects_invalid_fields(change):
    with pytest.raises(sample.RequestError):
        sample.parse_request(json.dumps(_request("m.json", "a" * 64, **change)))


def test_parse_rejects_unknown_missing_and_duplicate_keys():
    good = _request("m.json", "a" * 64)
    with pytest.raises(sample.RequestError):
        sample.parse_request(json.dumps(dict(good, extra=1)))
    missing = dict(good)
    del missing["sliding_band_angle_deg"]
    with pytest.raises(sample.RequestError):
        sample.parse_request(json.dumps(missing))
    text = json.dumps(good)[:-1] + ', "protocol": "%s"}' % sample.GENERIC
    with pytest.raises(sample.RequestError):
        sample.parse_request(text)
    with pytest.raises(sample.RequestError):
        sample.parse_request(json.dumps(good).replace("0.0", "NaN", 1))
    with pytest.raises(sample.RequestError):
        sample.parse_request("[1, 2]")


def test_unparsable_request_gets_the_reference_failure_response(tmp_path):
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text("{not json", encoding="utf-8")
    resp.write_text("old response\n", encoding="utf-8")
    assert sample.run_sample(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text(encoding="utf-8"))
    assert list(doc) == PLANAR_KEYS
    assert doc["protocol"] == sample.PLANAR_DC and doc["mesh_artifact_sha256"] == ""
    assert (doc["status"], doc["solve_status"], doc["postprocess_status"]) == \
        ("FAIL", "INVALID_ARGUMENT", "NOT_REQUESTED")
    assert doc["error_identifier"] == "GPU_FEMM_INVALID_ARGUMENT"
    assert doc["Fx_N"] is None and doc["node_A_Wb_per_m"] == [] and doc["mesh_node_count"] == 0
    assert doc["convergence"] == {"iterations": 0, "residual_l2": None}
    assert list(tmp_path.glob("*.tmp.*")) == []


@pytest.mark.parametrize("case", ["sha", "missing", "physics", "circuits"])
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
    else:
        kw = dict(circuit_currents_A=[1.0, 2.0])
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(_request(art, sha, **kw)), encoding="utf-8")
    assert sample.run_sample(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text(encoding="utf-8"))
    expected = {"sha": ("INVALID_ARGUMENT", "mesh artifact SHA mismatch"),
                "missing": ("INPUT_IO", "cannot open mesh artifact"),
                "physics": ("INVALID_ARGUMENT", "artifact physics, circuit count, group, or "
                                                "sliding-band request mismatch"),
                "circuits": ("INVALID_ARGUMENT", "artifact physics, circuit count, group, or "
                                                 "sliding-band request mismatch")}[case]
    assert (doc["solve_status"], doc["error_message"]) == expected
    assert doc["status"] == "FAIL" and doc["error_identifier"] == "GPU_FEMM_" + expected[0]
    assert doc["convergence"]["iterations"] == 0


def test_generic_requests_may_not_ask_for_postprocessing():
    model, _ = load_artifact(ARTIFACT)
    base = sample.parse_request(json.dumps(_request("m.json", "a" * 64)))
    assert sample.request_matches_artifact(base, sample.GENERIC_MESH, model)
    assert not sample.request_matches_artifact(dict(base, compute_force_torque=True),
                                               sample.GENERIC_MESH, model)
    assert not sample.request_matches_artifact(dict(base, sliding_band_angle_deg=7.5),
                                               sample.GENERIC_MESH, model)


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


def test_success_layout_and_number_format():
    """A solved planar-DC response with postprocess: keys in the reference's order, numbers
    with 17 significant digits (setprecision(17)), fields only when requested."""
    model, _ = load_artifact(ARTIFACT)
    req = sample.parse_request(json.dumps(_request(
        "m.json", "b" * 64, protocol=sample.PLANAR_DC, force_group_number=1,
        stress_air_group_number=0, airgap_radius_mm=1.0, airgap_angles_deg=[0.0, 90.0],
        compute_force_torque=True, include_field_solution=True)))
    sol = dict(a=np.array([0.1, 1.0 / 3.0]), bx=np.array([1e-5]), by=np.array([-0.0]),
               flux=np.array([4.6163396993973211e-07]), currents=[1.0], iterations=4,
               residual_l2=2.5e-9)
    post = dict(loads=np.array([-7.6799913871662699, -1.39915483064456, 0.0054151955763548599]),
                radial_b=np.array([0.5, -0.25]))
    text = sample.format_response(0, 0, True, 0, req, sol, post, "", model)
    doc = json.loads(text)
    assert list(doc) == PLANAR_KEYS
    assert (doc["status"], doc["postprocess_status"], doc["error_message"]) == ("PASS", "OK", "")
    assert doc["Fx_N"] == -7.6799913871662699 and doc["torque_Nm"] == 0.0054151955763548599
    assert doc["node_A_Wb_per_m"] == [0.1, 1.0 / 3.0] and doc["mesh_element_count"] == 1
    assert doc["airgap_sample_angles_deg"] == [0.0, 90.0]
    assert doc["convergence"] == {"iterations": 4, "residual_l2": 2.5e-9}
    assert '"node_A_Wb_per_m": [0.10000000000000001, 0.33333333333333331]' in text
    assert '"element_By_T": [-0]' in text and '"actual_circuit_currents_A": [1]' in text
    assert '"element_Bx_T": [1.0000000000000001e-05]' in text
    # Not-converged solve: counts and arrays stay empty, the convergence record is kept.
    text = sample.format_response(14, 14, True, 0, req, sol, None, "", model)
    doc = json.loads(text)
    assert (doc["status"], doc["postprocess_status"]) == ("FAIL", "NOT_RUN")
    assert doc["mesh_node_count"] == 0 and doc["convergence"]["iterations"] == 4


def test_solver_failure_is_reported_in_the_response(tmp_path):
    """A valid request on a host without a GPU: the library's GPU_UNAVAILABLE becomes the
    response's solve status (no solve attempted, as when the reference's Initialize fails)."""
    import ctypes

    from femm_gpu_amd.solver import library

    text = ctypes.create_string_buffer(256)
    if library().femm_device_info(text, 256) == 0:   # the library's own device query
        pytest.skip("a GPU is present")
    art = tmp_path / "mesh.gpu.json"
    shutil.copyfile(ARTIFACT, art)
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(_request(art, _sha(art))), encoding="utf-8")
    assert sample.run_sample(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text(encoding="utf-8"))
    assert (doc["status"], doc["solve_status"], doc["error_identifier"], doc["error_message"]) == \
        ("FAIL", "GPU_UNAVAILABLE", "GPU_FEMM_GPU_UNAVAILABLE", "GPU_UNAVAILABLE")
    assert doc["problem_type"] == "planar" and doc["convergence"]["iterations"] == 0
    assert list(tmp_path.glob("*.tmp.*")) == []


def test_command_line_exit_codes(tmp_path):
    """`python -m femm_gpu_amd.sample`: 2 for a wrong command line, 1 with a published FAIL
    response for an invalid request (the reference CLI's exit contract, app/cli.cpp:20-181)."""
    import os
    import subprocess
    import sys

    env = dict(os.environ, PYTHONPATH=str(Path(sample.__file__).resolve().parents[1]))
    cmd = [sys.executable, "-m", "femm_gpu_amd.sample"]
    assert subprocess.run(cmd, env=env, capture_output=True).returncode == 2
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(_request("m.json", "a" * 64, extra=1)), encoding="utf-8")
    done = subprocess.run(cmd + [str(req), str(resp)], env=env, capture_output=True, text=True)
    assert done.returncode == 1 and "FAIL planar-DC request" in done.stderr
    assert json.loads(resp.read_text(encoding="utf-8"))["status"] == "FAIL"


def test_unwritable_response_location_fails_cleanly(tmp_path):
    req = tmp_path / "request.json"
    req.write_text("{}", encoding="utf-8")
    assert sample.run_sample(str(req), str(tmp_path / "missing_dir" / "response.json")) == 1
