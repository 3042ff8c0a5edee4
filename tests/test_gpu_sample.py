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
    sha = ff.write_artifact(art, model, [current], turns={2: 50.0})
    request = dict(protocol=sample.PLANAR_DC, mesh_artifact_path=str(art),
                   mesh_artifact_sha256=sha,
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


def _near(got, want, atol):
    return abs(got - want) <= atol + 1e-5 * abs(want)


def test_motor_batch_reproduces_frozen_femm(tmp_path):
    """gpu_femm_motor_batch_v1 end to end: the frozen PM / coil model as a posed motor artifact
    (gpu_femm_mesh_v1), five items in chunks of three: the frozen current (twice, and once
    through the single-sample adapter), its negative and half, and one item whose artifact SHA
    does not match.  The frozen-current items carry the frozen FEMM flux, weighted-stress loads
    and air-gap samples at the reference's 1e-5 bar; the mismatched item fails alone (batch
    PARTIAL_FAILURE); solved items equal the single-sample adapter's numbers bit for bit."""
    say("case motor batch pm_coil")
    model, _, current, flux = ff.nonlinear_pm_coil()
    post = ff.nonlinear_pm_coil_postprocess()
    art = tmp_path / "pm_coil.motor.gpu.json"
    sha = ff.write_artifact(art, model, [current], schema="gpu_femm_mesh_v1", turns={2: 50.0})
    base = dict(protocol=sample.MOTOR, mesh_artifact_path=str(art), mesh_artifact_sha256=sha,
                base_motor_fem_sha256="b" * 64, source_fem_sha256="a" * 64,
                circuit_currents_A=[current], selected_group_number=post["selected_label"],
                air_group_number=3, airgap_radius_mm=post["radius_m"] * 1e3,
                airgap_angles_deg=[a * 180.0 / math.pi for a in post["angles"]],
                rotor_angle_deg=0.0, displacement_mm=[0.0, 0.0])
    items = [("frozen", base), ("negative", dict(base, circuit_currents_A=[-current])),
             ("bad-sha", dict(base, mesh_artifact_sha256="0" * 64)),
             ("half", dict(base, circuit_currents_A=[0.5 * current])),
             ("frozen.again", base)]
    batch = dict(protocol=sample.MOTOR_BATCH, max_items_per_chunk=3,
                 items=[dict(task_id=t, request=r) for t, r in items])
    req, resp = tmp_path / "batch.json", tmp_path / "batch.response.json"
    req.write_text(json.dumps(batch), encoding="utf-8")
    assert sample.run_motor_batch(str(req), str(resp)) == 1
    doc = json.loads(resp.read_text(encoding="utf-8"))
    assert (doc["status"], doc["solve_status"]) == ("FAIL", "PARTIAL_FAILURE")
    got = {i["task_id"]: i["response"] for i in doc["items"]}
    assert [i["task_id"] for i in doc["items"]] == [t for t, _ in items]
    assert got["bad-sha"]["solve_status"] == "INVALID_ARGUMENT"
    for t in ("frozen", "negative", "half", "frozen.again"):
        assert (got[t]["status"], got[t]["solve_status"]) == ("PASS", "OK"), t
        assert got[t]["mesh_element_count"] == model.triangle_count
    r = got["frozen"]
    say(f"  flux {r['circuit_flux_linkage_Wb'][0]:.12e} (frozen {flux:.12e}), torque "
        f"{r['torque_Nm']:.12e} (frozen {post['torque']:.12e})")
    assert _near(r["circuit_flux_linkage_Wb"][0], flux, 1e-12)
    assert _near(r["Fx_N"], post["fx"], 1e-8) and _near(r["Fy_N"], post["fy"], 1e-8)
    assert _near(r["torque_Nm"], post["torque"], 1e-10)
    for g, w in zip(r["airgap_radial_flux_density_T"], post["radial_b"]):
        assert _near(g, w, 1e-9)
    # the same item in another chunk, and through the single-sample adapter: same numbers
    assert got["frozen.again"] == r
    single = tmp_path / "single.json"
    req.write_text(json.dumps(base), encoding="utf-8")
    assert sample.run_motor_sample(str(req), str(single)) == 0
    assert json.loads(single.read_text(encoding="utf-8")) == r
    assert doc["cache"]["chunk_count"] == 2 and doc["cache"]["mesh_cache_hits"] == 3


def test_sliding_band_motor_sample_and_batch(tmp_path):
    """gpu_femm_motor_sample_v1 / gpu_femm_motor_batch_v1 on a sliding-band artifact
    (gpu_femm_mesh_v2, a full 12-slot / 4-pole machine whose air gap is a native air-gap
    element): the adapter turns the element to each item's rotor angle
    (ApplySlidingBandRotorAngle) and reports the element's Maxwell-stress torque and Fourier
    air-gap B (app/sample_adapter.cpp:59-61).  Against the oracle on the same turned model with
    the production options: flux linkage within 1e-7, torque and forces within 1e-7 of the
    torque, radial B within 1e-7; batch items equal the single-sample responses bit for bit,
    and items at the same angle share one operator."""
    import oracle_lib as ol
    from femm_gpu_amd import meshes, wire
    from femm_gpu_amd.model import production_options
    say("case sliding-band motor sample and batch")
    fam = meshes.AgeMachine(96, "age")
    model = fam.model()
    currents = [3e6, -1.5e6, -1.5e6]
    art = tmp_path / "machine.v2.json"
    sha = ff.write_artifact(art, model, currents, schema="gpu_femm_mesh_v2")
    angles_deg = [5.0, 40.0, 95.0, 200.0]
    base = dict(protocol=sample.MOTOR, mesh_artifact_path=str(art), mesh_artifact_sha256=sha,
                base_motor_fem_sha256="b" * 64, source_fem_sha256="a" * 64,
                circuit_currents_A=currents, selected_group_number=1, air_group_number=0,
                airgap_radius_mm=34.5, airgap_angles_deg=angles_deg, rotor_angle_deg=0.0,
                displacement_mm=[0.0, 0.0])
    singles = {}
    for rotor in (0.0, 1.3, 2.6):
        req, resp = tmp_path / f"m{rotor}.json", tmp_path / f"m{rotor}.out.json"
        req.write_text(json.dumps(dict(base, rotor_angle_deg=rotor)), encoding="utf-8")
        assert sample.run_motor_sample(str(req), str(resp)) == 0
        doc = json.loads(resp.read_text(encoding="utf-8"))
        assert (doc["status"], doc["solve_status"]) == ("PASS", "OK")
        singles[rotor] = doc
        turned = wire.rotate_air_gaps(model, rotor)
        ref = ol.newton_solve(turned, np.asarray(currents), production_options())
        assert ref["status"] == 0
        st, loads, radial = ol.age_postprocess(turned, ref["a"],
                                               np.radians(np.asarray(angles_deg)))
        assert st == 0
        flux = np.asarray(doc["circuit_flux_linkage_Wb"])
        got = np.array([doc["Fx_N"], doc["Fy_N"], doc["torque_Nm"]])
        say(f"  rotor {rotor} deg: torque {got[2]:.9e} (oracle {loads[2]:.9e}), flux "
            f"{np.abs(flux - ref['flux']).max() / np.abs(ref['flux']).max():.2e}")
        assert np.abs(flux - ref["flux"]).max() <= 1e-7 * np.abs(ref["flux"]).max()
        assert np.abs(got - loads).max() <= 1e-7 * abs(loads[2])
        rb = np.asarray(doc["airgap_radial_flux_density_T"])
        assert np.abs(rb - radial).max() <= 1e-7 * np.abs(radial).max()
    assert singles[0.0]["torque_Nm"] != singles[1.3]["torque_Nm"]
    items = [("a", 1.3), ("b", 0.0), ("c", 2.6), ("d", 1.3)]
    batch = dict(protocol=sample.MOTOR_BATCH, max_items_per_chunk=4,
                 items=[dict(task_id=t, request=dict(base, rotor_angle_deg=r)) for t, r in items])
    req, resp = tmp_path / "batch.json", tmp_path / "batch.out.json"
    req.write_text(json.dumps(batch), encoding="utf-8")
    assert sample.run_motor_batch(str(req), str(resp)) == 0
    doc = json.loads(resp.read_text(encoding="utf-8"))
    # batch level: PASS or PARTIAL_FAILURE (app/batch_response.cpp:297)
    assert (doc["status"], doc["solve_status"]) == ("PASS", "PASS")
    for item, (t, rotor) in zip(doc["items"], items):
        assert item["task_id"] == t and item["response"] == singles[rotor], t
    assert doc["cache"]["chunk_count"] == 1
