"""The single-sample wire contract end to end on the GPU: each generic FEMM fixture artifact
(gpu_femm_magnetostatic_mesh_v1) is solved from a gpu_femm_magnetostatic_sample_v1 request,
and the published response must carry the frozen FEMM answers of those fixtures
(gpu_solver/tests/test_femm_gpu_cli.py:223-270: flux linkage and nodal potential; the
axisymmetric potential is reported as the poloidal flux 2*pi*r*A, app/sample_adapter.cpp:301-308)."""

import hashlib
import json
import math
import shutil
import sys

import numpy as np
import pytest

import femm_fixtures as ff
from femm_gpu_amd import sample

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("name,pot_l2,flux_rel", [
    ("periodic_strip", 1e-10, 1e-10),
    ("antiperiodic_strip", 1e-10, 1e-10),
    ("axisymmetric_coil", 3e-3, 1e-3),
])
def test_generic_sample_request_reproduces_frozen_femm(tmp_path, name, pot_l2, flux_rel):
    say(f"case sample {name}")
    model, current, expected, flux = ff.generic_reference(name)
    art = tmp_path / f"{name}.gpu.json"
    shutil.copyfile(ff.GOLDEN / f"{name}_v1" / f"{name}.gpu.json", art)
    request = dict(protocol=sample.GENERIC, mesh_artifact_path=str(art),
                   mesh_artifact_sha256=hashlib.sha256(art.read_bytes()).hexdigest(),
                   circuit_currents_A=[current], force_group_number=-1,
                   stress_air_group_number=-1, airgap_radius_mm=0.0, airgap_angles_deg=[],
                   sliding_band_angle_deg=0.0, compute_force_torque=False,
                   include_field_solution=True)
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(request), encoding="utf-8")
    assert sample.run_sample(str(req), str(resp)) == 0
    doc = json.loads(resp.read_text(encoding="utf-8"))
    assert (doc["status"], doc["solve_status"], doc["postprocess_status"]) == \
        ("PASS", "OK", "NOT_REQUESTED")
    assert doc["error_identifier"] == "" and doc["field_solution_included"] is True
    assert doc["problem_type"] == ("axisymmetric" if name.startswith("axi") else "planar")
    assert doc["actual_circuit_currents_A"] == [current]
    assert doc["mesh_node_count"] == model.node_count
    assert doc["mesh_element_count"] == model.triangle_count
    assert len(doc["element_B_component_1_T"]) == model.triangle_count
    pot = np.asarray(doc["node_potential"])
    assert math.sqrt(np.sum((pot - expected) ** 2) / np.sum(expected ** 2)) <= pot_l2
    got = doc["circuit_flux_linkage_Wb"][0]
    assert abs(got - flux) / abs(flux) <= flux_rel
    assert doc["convergence"]["iterations"] >= 1
    assert doc["convergence"]["residual_l2"] < 1e-6
