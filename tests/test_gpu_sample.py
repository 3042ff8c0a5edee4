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


def _planar_dc_artifact(path, model, current, turns):
    """Write the frozen PM / coil model as a gpu_femm_planar_dc_mesh_v1 artifact: one region
    per label with group number = label, the coil's turns chosen so J_per_A = turns / area."""
    from femm_gpu_amd.model import MU0

    mats, regions = [], []
    for i, m in enumerate(model.materials):
        b, h = (model.curves[m.bh_curve] if m.bh_curve >= 0 else ([], []))
        mats.append(dict(id=f"m{i}", mu_x=1.0 / (MU0 * m.reluctivity_zero),
                         mu_y=1.0 / (MU0 * m.reluctivity_zero), lam_type=0, lam_fill=1,
                         H_c_A_per_m=m.hc, B_T=[float(x) for x in b],
                         H_A_per_m=[float(x) for x in h]))
        regions.append(dict(id=f"r{i}", material_id=f"m{i}", circuit_index=m.circuit,
                            turns=turns if m.circuit >= 0 else 0.0,
                            pm_magnetization_deg=m.magnetization_deg, group_number=i))
    doc = dict(schema_version="gpu_femm_planar_dc_mesh_v1", resolved=dict(
        model=dict(frequency_hz=0, problem_type="planar", depth_mm=model.depth_m * 1e3),
        nodes_mm=(model.node_xy * 1e3).tolist(), materials=mats,
        circuits=[dict(current_A=current)], regions=regions,
        triangles=dict(node_indices=model.triangles.tolist(),
                       region_ids=[f"r{k}" for k in model.triangle_material]),
        outer_dirichlet=dict(node_indices=model.dirichlet_nodes.tolist(),
                             A_Wb_per_m=model.dirichlet_values.tolist()),
        node_constraints=[]))
    path.write_text(json.dumps(doc), encoding="utf-8")


def test_planar_dc_request_with_force_torque_and_airgap_probes(tmp_path):
    """The planar-DC path with postprocessing: weighted-stress loads on the PM group against
    air and the four smoothed air-gap samples at r = 1 mm, against the frozen FEMM values
    (reference_commands.cu:144-208), at the reference's 1e-5 relative bar.  The adapter solves
    with the production sample options (undamped step < 1e-6); the oracle with those options
    lands within 4.3e-7 of every frozen value."""
    say("case sample planar_dc pm_coil")
    model, _, current, flux = ff.nonlinear_pm_coil()
    post = ff.nonlinear_pm_coil_postprocess()
    art = tmp_path / "pm_coil.gpu.json"
    _planar_dc_artifact(art, model, current, 50.0)
    request = dict(protocol=sample.PLANAR_DC, mesh_artifact_path=str(art),
                   mesh_artifact_sha256=hashlib.sha256(art.read_bytes()).hexdigest(),
                   circuit_currents_A=[current], force_group_number=post["selected_label"],
                   stress_air_group_number=3, airgap_radius_mm=post["radius_m"] * 1e3,
                   airgap_angles_deg=[a * 180.0 / math.pi for a in post["angles"]],
                   sliding_band_angle_deg=0.0, compute_force_torque=True,
                   include_field_solution=False)
    req, resp = tmp_path / "request.json", tmp_path / "response.json"
    req.write_text(json.dumps(request), encoding="utf-8")
    assert sample.run_sample(str(req), str(resp)) == 0
    doc = json.loads(resp.read_text(encoding="utf-8"))
    assert (doc["status"], doc["solve_status"], doc["postprocess_status"]) == ("PASS", "OK", "OK")
    assert doc["field_solution_included"] is False and doc["node_A_Wb_per_m"] == []
    assert doc["mesh_node_count"] == model.node_count

    def near(got, want, atol):
        return abs(got - want) <= atol + 1e-5 * abs(want)

    assert near(doc["circuit_flux_linkage_Wb"][0], flux, 1e-12)
    assert near(doc["Fx_N"], post["fx"], 1e-8) and near(doc["Fy_N"], post["fy"], 1e-8)
    assert near(doc["torque_Nm"], post["torque"], 1e-10)
    assert len(doc["airgap_radial_flux_density_T"]) == len(post["radial_b"])
    for got, want in zip(doc["airgap_radial_flux_density_T"], post["radial_b"]):
        assert near(got, want, 1e-9)
