"""Native air-gap element (AGE) on the host: the element matrix, its assembly into the plan,
validation, the sliding-band rotor angle and the AGE postprocess.

What pins each part:
* The cell matrix (native_airgap_matrix.cpp:5-22, text from fkn/prob1big.cpp:158-213) is
  derived here, not copied: bilinear between Catmull-Rom corners with the mixed-mass rectangle
  stiffness.  It is held to what the mathematics fixes (symmetry, constants in the null
  space, positive semi-definiteness, the energy K of a field linear along the arc, the
  rectangle stiffness alone at zero shifts) and to closed forms of single entries, and the
  product's closed form to the oracle's quadrature form.
* The stencil and the per-cell field are pinned by the reference's own postprocess
  (airgap_postprocess.cpp:41-160, compiled from the reference sources into oracle/_ref):
  its Br / Bt formulas are the gradient of the same corner interpolation, so force, torque
  and Fourier radial B agree to rounding.
* The sliding-band remap equals the reference's ApplySlidingBandAngle bit for bit, and a
  turn by whole cells equals moving the magnets by as many cells.
* The assembled operator is pinned physically: the same machine with its gap meshed by
  triangles gives the same flux linkage, torque and air-gap B to discretisation accuracy.
The reference's assembly entry points are not run on AGE models (their matrix translation
unit needs build-generated code and is not compiled), so the plan's AGE terms are checked
against the oracle's.
"""

import math

import numpy as np
import pytest

import oracle_lib as ol
from femm_gpu_amd import meshes, wire
from femm_gpu_amd.model import accepted_options
from femm_gpu_amd.solver import FemmError, Plan

needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference bridge not built")
CURRENTS = np.array([3e6, -1.5e6, -1.5e6])


def _cr(t):
    """Catmull-Rom weights of nodes -1, 0, 1, 2 at t in [0, 1] (textbook form)."""
    return 0.5 * np.array([-t + 2 * t * t - t ** 3, 2 - 5 * t * t + 3 * t ** 3,
                           t + 4 * t * t - 3 * t ** 3, t ** 3 - t * t])


@pytest.mark.parametrize("ci,co", [(0.25, 0.0), (0.7, 1.0), (0.0, 1.0), (1.0, 1.0), (0.25, 0.75)])
def test_age_matrix_invariants(ci, co):
    K = 0.3
    M = ol.age_matrix(K, 1 / K, ci, co)
    assert np.all(np.isfinite(M)) and np.allclose(M, M.T, rtol=0, atol=1e-15)
    assert np.abs(M.sum(axis=1)).max() < 1e-14          # constants cost nothing
    assert np.linalg.eigvalsh(M).min() > -1e-14          # positive semi-definite
    assert np.all(M.diagonal() >= 0.0)
    # A potential rising by one per cell along the arc on both rings (ring nodes at
    # j - 2 + shift): the corner interpolation reproduces it, the energy is K.
    a = np.concatenate([np.arange(5) - 2 + ci, np.arange(5) - 2 + co])
    assert abs(a @ M @ a - K) < 1e-13
    # Single entries in closed form: node 0 couples through the inner u = 0 corner only,
    # node 4 through u = 1, node 5 / 9 through the outer corners.
    S00 = 5 * (K + 1 / K) / 12
    w0, w3 = _cr(1 - ci)[0], _cr(1 - ci)[3]
    v0, v3 = _cr(1 - co)[0], _cr(1 - co)[3]
    assert M[0, 0] == pytest.approx(w0 * w0 * S00, abs=1e-15)
    assert M[0, 4] == pytest.approx(w0 * w3 * (-5 * K + 1 / K) / 12, abs=1e-15)
    assert M[0, 5] == pytest.approx(w0 * v0 * (K - 5 / K) / 12, abs=1e-15)
    assert M[0, 9] == pytest.approx(w0 * v3 * -(K + 1 / K) / 12, abs=1e-15)


def test_age_matrix_at_zero_shift_is_the_rectangle_stiffness():
    """With both rings unshifted the corners are nodes 2, 3 (inner) and 7, 8 (outer): the
    matrix is the rectangle's K A (x) M + K^-1 M (x) A and zero elsewhere; the reference's
    own test case (0.75, 1/0.75, 0.25, 0.75) is finite with a positive diagonal."""
    K = 0.75
    M = ol.age_matrix(K, 1 / K, 0.0, 0.0)
    A = np.array([[1.0, -1.0], [-1.0, 1.0]])
    Mm = np.array([[5.0, 1.0], [1.0, 5.0]]) / 12
    S = K * np.kron(Mm, A) + (1 / K) * np.kron(A, Mm)   # corners (u, v) with u fastest
    idx = [2, 3, 7, 8]
    assert np.allclose(M[np.ix_(idx, idx)], S, rtol=0, atol=1e-15)
    rest = np.ones((10, 10), bool)
    rest[np.ix_(idx, idx)] = False
    assert np.abs(M[rest]).max() < 1e-15
    R = ol.age_matrix(0.75, 1 / 0.75, 0.25, 0.75)
    assert np.all(R.diagonal() > 0.0) and np.allclose(R, R.T, rtol=0, atol=1e-15)
    assert ol.age_matrix(0.0, 1.0, 0.0, 0.0) is None
    assert ol.age_matrix(1.0, 1.0, math.nan, 0.0) is None


@pytest.mark.parametrize("sector", [False, True])
def test_plan_air_gap_terms_match_oracle(sector):
    """CSR pattern (triangle and AGE couplings, nonlinear_topology.cpp:37-85) and the fixed
    AGE values per slot equal the oracle's; the product computes the matrix in closed form,
    the oracle by quadrature."""
    model = meshes.age_machine(48, "age", sector=sector)
    plan = Plan(model)
    mine = plan.symbolic()
    plan.close()
    st, ref = ol.prepare_symbolic(model)
    assert st == 0
    for key in ("rows", "cols", "diag_slots", "free_nodes", "node_to_free", "node_sign"):
        assert np.array_equal(mine[key], ref[key]), key
    scale = np.abs(ref["age_values"]).max()
    assert scale > 0 and np.abs(mine["age_values"] - ref["age_values"]).max() <= 1e-14 * scale
    assert np.array_equal(mine["age_values"] != 0, ref["age_values"] != 0)


def test_fixed_ring_nodes_load_the_rhs():
    """An element whose outer ring is a Dirichlet circle with A != 0 moves those couplings
    into the right-hand side (nonlinear_topology.cpp:228-230)."""
    fam = meshes.AgeMachine(48, "age")
    model = fam.model()
    ring = fam.outer_ring
    model.dirichlet_nodes = np.concatenate([model.dirichlet_nodes, ring]).astype(np.int32)
    values = 1e-3 * np.sin(np.arange(len(ring)) * 2 * math.pi / len(ring))
    model.dirichlet_values = np.concatenate([model.dirichlet_values, values])
    model._c = None
    plan = Plan(model)
    mine = plan.symbolic()
    plan.close()
    st, ref = ol.prepare_symbolic(model)
    assert st == 0 and np.abs(ref["age_rhs"]).max() > 0
    assert np.abs(mine["age_rhs"] - ref["age_rhs"]).max() <= 1e-14 * np.abs(ref["age_rhs"]).max()


@needs_ref
@pytest.mark.parametrize("change,field,value", [
    ("few_cells", None, None), ("inner_radius", "inner_radius_m", 0.0),
    ("outer_radius", "outer_radius_m", 34e-3), ("arc", "arc_length_deg", 0.0),
    ("shift", "inner_shift", math.nan), ("centre", "center_xy", (math.inf, 0.0)),
    ("node", None, None), ("weight", None, None)])
def test_air_gap_validation_codes_match_reference(change, field, value):
    """ValidateNonlinearModel's AGE rules (nonlinear_model.cpp:281-297): the plan, the
    oracle and the reference return the same status."""
    model = meshes.age_machine(48, "age")
    age = model.air_gaps[0]
    if change == "few_cells":
        age.nodes, age.weights = age.nodes[:2], age.weights[:2]
    elif change == "node":
        age.nodes[3, 1] = model.node_count
    elif change == "weight":
        age.weights[5, 2] = math.inf
    else:
        setattr(age, field, value)
    model._c = None
    want = ol.validate(model, lib="ref")
    assert want != 0
    assert ol.validate(model) == want
    with pytest.raises(FemmError) as err:
        Plan(model)
    assert err.value.code == want


@needs_ref
@pytest.mark.parametrize("angle", [0.0, 1.2, 3.75, 7.5, -5.0, 400.0, 1e-9, -3.75, 359.99])
@pytest.mark.parametrize("sector", [False, True])
def test_sliding_band_remap_matches_reference(angle, sector):
    model = meshes.age_machine(96, "age", sector=sector)
    mine = wire.rotate_air_gaps(model, angle).air_gaps[0]
    shifts, nodes, weights = ol.ref_sliding_band(model, angle)
    assert mine.inner_shift == shifts[0]
    assert np.array_equal(mine.nodes, nodes) and np.array_equal(mine.weights, weights)
    assert mine.outer_shift == model.air_gaps[0].outer_shift
    with pytest.raises(ValueError):
        wire.rotate_air_gaps(model, math.nan)


def test_whole_cell_turn_equals_moved_magnets():
    """Turning the rotor by two cells through the sliding band solves the same problem as
    moving the magnets two cells along the unchanged element (the opposite turn does not)."""
    fam = meshes.AgeMachine(96, "age")
    opts = accepted_options()
    turned = ol.newton_solve(wire.rotate_air_gaps(fam.model(), 2 * 3.75), CURRENTS, opts)
    moved = ol.newton_solve(fam.model(rotor_steps=2), CURRENTS, opts)
    back = ol.newton_solve(fam.model(rotor_steps=-2), CURRENTS, opts)
    scale = np.abs(moved["flux"]).max()
    assert turned["status"] == moved["status"] == 0
    assert np.abs(turned["flux"] - moved["flux"]).max() <= 1e-12 * scale
    assert np.abs(back["flux"] - moved["flux"]).max() > 0.1 * scale


@needs_ref
@pytest.mark.parametrize("sector,angle", [(False, 0.0), (False, 1.3), (True, 0.0), (True, 0.7)])
def test_age_postprocess_matches_reference(sector, angle):
    model = wire.rotate_air_gaps(meshes.age_machine(48, "age", sector=sector), angle)
    sol = ol.newton_solve(model, CURRENTS, accepted_options())
    assert sol["status"] == 0
    angles = (np.arange(12) + 0.5) * (math.pi / 24 if sector else math.pi / 6)
    st, loads, radial = ol.age_postprocess(model, sol["a"], angles)
    st_r, loads_r, radial_r = ol.ref_age_postprocess(model, sol["a"], angles)
    assert st == st_r == 0
    assert np.abs(loads - loads_r).max() <= 1e-12 * np.abs(loads_r).max()
    assert np.abs(radial - radial_r).max() <= 1e-13 * np.abs(radial_r).max()
    st, _, _ = ol.age_postprocess(model, sol["a"], [math.nan])
    st_r, _, _ = ol.ref_age_postprocess(model, sol["a"], [math.nan])
    assert st == st_r == 2


@pytest.mark.parametrize("sector", [False, True])
def test_air_gap_element_agrees_with_a_triangulated_gap(sector):
    """The same machine with its gap meshed by two layers of triangles: flux linkage within
    0.1 %, torque (AGE Maxwell stress vs weighted stress tensor) within 3 %, radial air-gap B
    (Fourier series vs smoothed nodal B) within 4 %."""
    cells = 96 if sector else 192
    age = meshes.age_machine(cells, "age", sector=sector)
    tri = meshes.age_machine(cells, "triangles", sector=sector)
    opts = accepted_options()
    ra, rt = ol.newton_solve(age, CURRENTS, opts), ol.newton_solve(tri, CURRENTS, opts)
    assert ra["status"] == rt["status"] == 0
    assert np.abs(ra["flux"] - rt["flux"]).max() <= 1e-3 * np.abs(rt["flux"]).max()
    angles = (np.arange(8) + 0.5) * (math.pi / 16 if sector else math.pi / 4)
    _, la, radial_a = ol.age_postprocess(age, ra["a"], angles)
    _, bx, by, _ = ol.finalize(tri, rt["a"])
    _, lw, radial_w, _ = ol.weighted_stress(tri, rt["a"], bx, by, selected_group=1, air_group=0,
                                            radius_m=34.5e-3, angles_rad=angles)
    assert abs(la[2] - lw[2]) <= 0.03 * abs(lw[2])
    assert np.abs(radial_a - radial_w).max() <= 0.04 * np.abs(radial_w).max()


# ---- sliding-band artifacts (gpu_femm_mesh_v2) ------------------------------------------------

def _v2(tmp_path, name="machine.v2.json", **kw):
    import femm_fixtures as ff
    fam = meshes.AgeMachine(48, "age")
    model = fam.model()
    path = tmp_path / name
    sha = ff.write_artifact(path, model, CURRENTS, schema="gpu_femm_mesh_v2", **kw)
    return model, path, sha


@needs_ref
@pytest.mark.parametrize("scalar", [False, True])
def test_v2_artifact_parses_like_the_reference(tmp_path, scalar):
    model, path, _ = _v2(tmp_path, age_scalar=scalar)
    mine, currents = wire.load_artifact(path)
    ref = ol.ref_parse_artifact(path)
    assert len(mine.air_gaps) == 1 and ref["air_gaps"] is not None
    age, want = mine.air_gaps[0], model.air_gaps[0]
    assert np.array_equal(age.nodes, want.nodes) and np.array_equal(age.weights, want.weights)
    ri, ro, arc, ci, co, cx, cy, anti = ref["air_gaps"]["params"][0]
    assert (age.inner_radius_m, age.outer_radius_m, age.arc_length_deg) == (ri, ro, arc)
    assert (age.inner_shift, age.outer_shift, age.center_xy) == (ci, co, (cx, cy))
    assert age.antiperiodic == bool(anti)
    assert np.array_equal(age.nodes, ref["air_gaps"]["nodes"])
    assert np.array_equal(age.weights, ref["air_gaps"]["weights"])
    assert np.array_equal(mine.node_xy, ref["node_xy"]) and np.array_equal(currents, CURRENTS)


def _edits():
    def pose(r):
        r["pose"]["rotor_angle_deg"] = 2.0
    def shift(r):
        r["pose"]["displacement_mm"] = [0.1, 0.0]
    def empty(r):
        r["air_gap_elements"] = []
    def missing_points(r):
        del r["air_gap_elements"][0]["quad_points"]
    def extra(r):
        r["air_gap_elements"][0]["extra"] = 1
    def count(r):
        r["air_gap_elements"][0]["sector_count"] = 47
    def periodicity(r):
        r["air_gap_elements"][0]["periodicity"] = "cyclic"
    def one_cell(r):
        r["air_gap_elements"][0]["sector_count"] = 1
        r["air_gap_elements"][0]["quad_points"] = r["air_gap_elements"][0]["quad_points"][:2]
    def fractional_node(r):
        r["air_gap_elements"][0]["quad_points"][3]["n1"] = 2.5
    def point_extra(r):
        r["air_gap_elements"][0]["quad_points"][3]["n4"] = 1
    def no_name(r):
        r["air_gap_elements"][0]["name"] = ""
    def inner_radius(r):
        r["air_gap_elements"][0]["ri_mm"] = 0.0
    def node_range(r):
        r["air_gap_elements"][0]["quad_points"][3]["n1"] = 10 ** 6
    def wrong_type(r):
        r["air_gap_elements"] = "gap"
    return [pose, shift, empty, missing_points, extra, count, periodicity, one_cell,
            fractional_node, point_extra, no_name, inner_radius, node_range, wrong_type]


@needs_ref
@pytest.mark.parametrize("edit", _edits(), ids=[e.__name__ for e in _edits()])
def test_v2_artifact_rejections_match_the_reference(tmp_path, edit):
    """Each rule of mesh_artifact.cpp:594-709 on a sliding-band artifact: both parsers refuse
    the document with the same reason."""
    _, path, _ = _v2(tmp_path, edit=edit)
    with pytest.raises(ValueError) as ref_err:
        ol.ref_parse_artifact(path)
    with pytest.raises(wire.ArtifactError) as err:
        wire.load_artifact(path)
    assert str(err.value) == str(ref_err.value)


def test_motor_request_on_a_sliding_band_artifact(tmp_path):
    """gpu_femm_motor_sample_v1 on a v2 artifact (RequestMatchesArtifact,
    io/sample_request.cpp:418-446): any rotor angle, no displacement, the selected and air
    groups present, probes inside the gap; on a host without a GPU a matching request reaches
    the solver (GPU_UNAVAILABLE), the others are refused before it."""
    import json
    from femm_gpu_amd import sample
    _, path, sha = _v2(tmp_path)
    base = dict(protocol=sample.MOTOR, mesh_artifact_path=str(path), mesh_artifact_sha256=sha,
                base_motor_fem_sha256="b" * 64, source_fem_sha256="a" * 64,
                circuit_currents_A=CURRENTS.tolist(), selected_group_number=1,
                air_group_number=0, airgap_radius_mm=34.5, airgap_angles_deg=[10.0, 20.0],
                rotor_angle_deg=7.3, displacement_mm=[0.0, 0.0])
    cases = [(base, None), (dict(base, displacement_mm=[0.5, 0.0]), "mismatch"),
             (dict(base, airgap_radius_mm=36.0), "mismatch"),
             (dict(base, selected_group_number=42), "mismatch")]
    for k, (req, expect) in enumerate(cases):
        rq, rs = tmp_path / f"r{k}.json", tmp_path / f"s{k}.json"
        rq.write_text(json.dumps(req), encoding="utf-8")
        code = sample.run_motor_sample(str(rq), str(rs))
        doc = json.loads(rs.read_text(encoding="utf-8"))
        assert code == 1
        if expect is None:
            if doc["solve_status"] != "GPU_UNAVAILABLE":
                pytest.skip("a GPU is present")
        else:
            assert doc["solve_status"] == "INVALID_ARGUMENT", (k, doc["solve_status"])
