"""Single-sample wire contract: one ``gpu_femm_planar_dc_sample_v1`` or
``gpu_femm_magnetostatic_sample_v1`` request in, one response JSON out.

    python -m femm_gpu_amd.sample REQUEST.json RESPONSE.json

Mirrors the reference's single-sample adapter:

* request parsing and validation: exactly the 11 fields, no unknown or duplicate keys,
  lower-case SHA-256, group pairing and air-gap rules (io/sample_request.cpp:152-210);
* artifact checks in order: open (INPUT_IO), SHA-256 of the bytes, parse, request/artifact
  match (app/sample_adapter.cpp:183-224, 364-392);
* the solve: the Newton driver with the production sample options from the Dirichlet start
  (:394-414), here ``femm_newton_batch`` with one item; the weighted-stress force / torque and
  air-gap probes when requested (:415-433), here ``femm_postprocess_batch``;
* the response: the same keys in the same order with 17 significant digits
  (``setprecision(17)``), PASS / FAIL, status names, ``GPU_FEMM_<NAME>`` error identifiers
  (:226-327); written to a temporary file and renamed into place, refusing to overwrite the
  request or the artifact (:442-487).

Exit status 0 when the response says PASS, 1 otherwise (also when no response could be
published), 2 for a wrong command line.  Sliding-band (motor v2 / v4) artifacts are not part
of these two schemas, so the frozen (weighted-stress) postprocess is the only one here.
"""

from __future__ import annotations

import hashlib
import itertools
import json
import math
import os
import sys
import time

import numpy as np

PLANAR_DC = "gpu_femm_planar_dc_sample_v1"
GENERIC = "gpu_femm_magnetostatic_sample_v1"
GENERIC_MESH = "gpu_femm_magnetostatic_mesh_v1"
REQUEST_FIELDS = ("protocol", "mesh_artifact_path", "mesh_artifact_sha256", "circuit_currents_A",
                  "force_group_number", "stress_air_group_number", "airgap_radius_mm",
                  "airgap_angles_deg", "sliding_band_angle_deg", "compute_force_torque",
                  "include_field_solution")

OK, INPUT_IO, INVALID_ARGUMENT = 0, 1, 2
_sequence = itertools.count()


class RequestError(ValueError):
    pass


def _strict_object(pairs):
    keys = [k for k, _ in pairs]
    if len(set(keys)) != len(keys):
        raise RequestError("duplicate JSON object key")
    return dict(pairs)


def _reject_constant(name):
    raise RequestError(f"non-standard JSON literal {name}")


def _is_number(v):
    return isinstance(v, (int, float)) and not isinstance(v, bool) and math.isfinite(v)


def _integer(v):
    if not _is_number(v) or float(v) != math.floor(float(v)) or abs(v) > 2**31 - 1:
        return None
    return int(v)


def parse_request(text: str) -> dict:
    """Strict parse of a sample request (ReadPlanarDcSampleRequestJson); raises RequestError."""
    try:
        root = json.loads(text, object_pairs_hook=_strict_object, parse_constant=_reject_constant)
    except RequestError:
        raise
    except ValueError as err:
        raise RequestError(f"invalid JSON: {err}") from None
    if not isinstance(root, dict):
        raise RequestError("request must be a JSON object")
    missing = [k for k in REQUEST_FIELDS if k not in root]
    extra = sorted(set(root) - set(REQUEST_FIELDS))
    if missing or extra:
        raise RequestError(f"request fields: missing {missing}, unknown {extra}")
    r = root
    selected, air = _integer(r["force_group_number"]), _integer(r["stress_air_group_number"])
    sha = r["mesh_artifact_sha256"]
    arrays_ok = all(isinstance(r[k], list) and all(_is_number(x) for x in r[k])
                    for k in ("circuit_currents_A", "airgap_angles_deg"))
    ok = (r["protocol"] in (PLANAR_DC, GENERIC)
          and isinstance(r["mesh_artifact_path"], str) and r["mesh_artifact_path"] != ""
          and isinstance(sha, str) and len(sha) == 64 and all(c in "0123456789abcdef" for c in sha)
          and selected is not None and air is not None and selected >= -1 and air >= -1
          and (selected == -1) == (air == -1) and not (selected >= 0 and selected == air)
          and _is_number(r["airgap_radius_mm"]) and r["airgap_radius_mm"] >= 0.0
          and _is_number(r["sliding_band_angle_deg"])
          and isinstance(r["compute_force_torque"], bool)
          and isinstance(r["include_field_solution"], bool)
          and arrays_ok and (not r["airgap_angles_deg"] or r["airgap_radius_mm"] > 0.0))
    if not ok:
        raise RequestError(f"invalid {PLANAR_DC} request")
    out = dict(r)
    out["force_group_number"], out["stress_air_group_number"] = selected, air
    out["circuit_currents_A"] = [float(x) for x in r["circuit_currents_A"]]
    out["airgap_angles_deg"] = [float(x) for x in r["airgap_angles_deg"]]
    return out


def request_matches_artifact(request: dict, schema: str, model) -> bool:
    """PlanarDcRequestMatchesArtifact (app/sample_adapter.cpp:183-224) for artifacts without a
    sliding band."""
    generic = request["protocol"] == GENERIC
    if generic != (schema == GENERIC_MESH):
        return False
    if len(request["circuit_currents_A"]) != model.circuit_count:
        return False
    groups = {m.group for m in model.materials}
    selected, air = request["force_group_number"], request["stress_air_group_number"]
    groups_enabled = selected >= 0
    if groups_enabled and (selected not in groups or air not in groups):
        return False
    needs_post = request["compute_force_torque"] or bool(request["airgap_angles_deg"])
    if generic:
        return (not needs_post and selected == -1 and air == -1
                and request["airgap_radius_mm"] == 0.0 and not request["airgap_angles_deg"]
                and request["sliding_band_angle_deg"] == 0.0)
    return request["sliding_band_angle_deg"] == 0.0 and (not needs_post or groups_enabled)


def _num(x) -> str:
    return "%.17g" % x


def _array(values) -> str:
    return ", ".join(_num(v) for v in values)


def format_response(status, solve_status, post_requested, post_status, request=None,
                    solution=None, post=None, diagnostic="", model=None) -> str:
    """The response text of WritePlanarDcSampleResponse (app/sample_adapter.cpp:226-327).
    ``solution``: dict(a, bx, by, flux, currents, iterations, residual_l2) or None when no solve
    was attempted; ``post``: dict(loads, radial_b) or None."""
    from .model import AXISYMMETRIC, status_name

    solved = solve_status == OK and request is not None and solution is not None
    posted = solved and post_requested and post_status == OK and post is not None
    message = "" if status == OK else (diagnostic or status_name(status))
    protocol = PLANAR_DC if request is None else request["protocol"]
    generic = protocol == GENERIC
    axisym = model is not None and model.problem_type == AXISYMMETRIC
    node_name = "node_potential" if generic else "node_A_Wb_per_m"
    first = "element_B_component_1_T" if generic else "element_Bx_T"
    second = "element_B_component_2_T" if generic else "element_By_T"
    out = ["{\n", f'  "protocol": "{protocol}",\n',
           '  "mesh_artifact_sha256": "%s",\n' % ("" if request is None
                                                  else request["mesh_artifact_sha256"])]
    if generic:
        out.append('  "problem_type": "%s",\n' % ("axisymmetric" if axisym else "planar"))
        out.append('  "field_components": [%s],\n' % ('"Br", "Bz"' if axisym else '"Bx", "By"'))
        out.append('  "node_potential_quantity": "%s",\n'
                   % ("poloidal_flux_function" if axisym else "magnetic_vector_potential"))
        out.append('  "node_potential_unit": "%s",\n' % ("Wb" if axisym else "Wb/m"))
    post_name = ("NOT_REQUESTED" if not post_requested
                 else (status_name(post_status) if solve_status == OK else "NOT_RUN"))
    out += ['  "status": "%s",\n' % ("PASS" if status == OK else "FAIL"),
            '  "solve_status": "%s",\n' % status_name(solve_status),
            '  "postprocess_status": "%s",\n' % post_name,
            '  "error_identifier": "%s",\n' % ("" if status == OK
                                               else "GPU_FEMM_" + status_name(status)),
            '  "error_message": %s,\n' % json.dumps(message)]
    if posted and request["compute_force_torque"]:
        fx, fy, tq = post["loads"]
        out.append(f'  "Fx_N": {_num(fx)},\n  "Fy_N": {_num(fy)},\n  "torque_Nm": {_num(tq)},\n')
    else:
        out.append('  "Fx_N": null,\n  "Fy_N": null,\n  "torque_Nm": null,\n')
    include = solved and request["include_field_solution"]
    potential = np.asarray(solution["a"]) if solved else np.zeros(0)
    if generic and axisym and solved:
        potential = potential * (2.0 * math.pi * model.node_xy[:, 0])
    out += ['  "actual_circuit_currents_A": [%s],\n' % (_array(solution["currents"]) if solved
                                                         else ""),
            '  "circuit_flux_linkage_Wb": [%s],\n' % (_array(solution["flux"]) if solved else ""),
            '  "airgap_sample_angles_deg": [%s],\n' % (_array(request["airgap_angles_deg"])
                                                        if posted else ""),
            '  "airgap_radial_flux_density_T": [%s],\n' % (_array(post["radial_b"]) if posted
                                                            else ""),
            '  "mesh_node_count": %d,\n' % (len(solution["a"]) if solved else 0),
            '  "mesh_element_count": %d,\n' % (len(solution["bx"]) if solved else 0),
            '  "field_solution_included": %s,\n' % ("true" if include else "false"),
            f'  "{node_name}": [%s],\n' % (_array(potential) if include else ""),
            f'  "{first}": [%s],\n' % (_array(solution["bx"]) if include else ""),
            f'  "{second}": [%s],\n' % (_array(solution["by"]) if include else "")]
    iterations = 0 if solution is None else int(solution["iterations"])
    residual = None if solution is None else solution["residual_l2"]
    out.append('  "convergence": {"iterations": %d, "residual_l2": %s}\n}\n'
               % (iterations, _num(residual) if residual is not None and math.isfinite(residual)
                  else "null"))
    return "".join(out)


def _solve(request, model):
    """One Newton solve (+ postprocess) through the C ABI; returns the response pieces."""
    from .model import production_options
    from .solver import FemmError, Plan, reference_config

    plan = Plan(model)
    try:
        currents = np.asarray(request["circuit_currents_A"], np.float64).reshape(1, model.circuit_count)
        out = plan.solve_batch(currents, production_options(), reference_config(), fields=True)
        r = out["results"][0]
        solution = dict(a=out["a"][0], bx=out["bx"][0], by=out["by"][0], flux=out["flux"][0],
                        currents=request["circuit_currents_A"],
                        iterations=int(r["newton_iterations"]),
                        residual_l2=float(r["relative_step"]))
        solve_status = int(r["status"])
        post_requested = request["compute_force_torque"] or bool(request["airgap_angles_deg"])
        post, post_status = None, OK
        if solve_status == OK and post_requested:
            try:
                plan.set_postprocess(
                    selected_group=request["force_group_number"],
                    air_group=request["stress_air_group_number"],
                    radius_m=request["airgap_radius_mm"] * 1e-3,
                    angles_rad=[a * math.pi / 180.0 for a in request["airgap_angles_deg"]],
                    max_mask_iterations=4096)
                loads, radial = plan.postprocess_batch(solution["a"])
                post = dict(loads=loads[0], radial_b=radial[0])
            except FemmError as err:
                post_status = err.code
        return solution, solve_status, post_requested, post, post_status
    finally:
        plan.close()


def _same_file(first, second):
    try:
        if os.path.exists(first) and os.path.exists(second):
            return os.path.samefile(first, second)
    except OSError:
        pass
    return os.path.realpath(first) == os.path.realpath(second)


def run_sample(request_path, response_path, log=sys.stderr) -> int:
    """PlanarDcSingleSampleAdapter: solve one request and publish its response atomically."""
    if _same_file(request_path, response_path):
        print("FAIL magnetostatic solve: response must not overwrite request", file=log)
        return 1
    try:
        with open(request_path, "r", encoding="utf-8") as f:
            request_text = f.read()
    except OSError:
        request_text = None
    request, parse_error = None, ""
    if request_text is not None:
        try:
            request = parse_request(request_text)
        except RequestError as err:
            parse_error = str(err)
    else:
        parse_error = "cannot open request"
    if request is not None and _same_file(request["mesh_artifact_path"], response_path):
        print("FAIL magnetostatic solve: response must not overwrite mesh artifact", file=log)
        return 1
    temporary = f"{response_path}.tmp.{time.time_ns()}.{next(_sequence)}"
    try:
        code = _run_direct(request, parse_error, temporary, log)
    except OSError:            # the response could not be written at all
        code = 1
    try:
        os.replace(temporary, response_path)
    except OSError:
        try:
            os.remove(temporary)
        except OSError:
            pass
        print("FAIL magnetostatic solve: could not atomically publish response", file=log)
        return 1
    return code


def _run_direct(request, parse_error, path, log) -> int:
    from .model import load_artifact, status_name

    if request is None:
        text = format_response(INVALID_ARGUMENT, INVALID_ARGUMENT, False, OK,
                               diagnostic=parse_error)
        with open(path, "w", encoding="utf-8") as f:
            f.write(text)
        print(f"FAIL planar-DC request: {parse_error}", file=log)
        return 1
    status, error, model, schema = OK, "", None, ""
    try:
        with open(request["mesh_artifact_path"], "rb") as f:
            data = f.read()
    except OSError:
        data = None
    if data is None:
        status, error = INPUT_IO, "cannot open mesh artifact"
    elif hashlib.sha256(data).hexdigest() != request["mesh_artifact_sha256"]:
        status, error = INVALID_ARGUMENT, "mesh artifact SHA mismatch"
    else:
        try:
            schema = json.loads(data.decode("utf-8"))["schema_version"]
            model, _ = load_artifact(request["mesh_artifact_path"])
        except Exception as err:   # any malformed artifact is an invalid argument
            status, error, model = INVALID_ARGUMENT, f"invalid mesh artifact: {err}", None
        if status == OK and not request_matches_artifact(request, schema, model):
            status = INVALID_ARGUMENT
            error = "artifact physics, circuit count, group, or sliding-band request mismatch"
    solution, post, post_status = None, None, OK
    post_requested = request["compute_force_torque"] or bool(request["airgap_angles_deg"])
    solve_status = status
    if status == OK:
        from .solver import FemmError

        try:
            solution, solve_status, post_requested, post, post_status = _solve(request, model)
        except FemmError as err:    # the plan or the solve call failed as a whole
            solve_status = err.code
        status = solve_status
        if solve_status == OK and post_requested:
            status = post_status
    text = format_response(status, solve_status, post_requested, post_status, request, solution,
                           post if post_status == OK else None, error, model)
    with open(path, "w", encoding="utf-8") as f:
        f.write(text)
    if status != OK:
        print(f"FAIL magnetostatic solve: {error or status_name(status)}", file=log)
        return 1
    return 0


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("usage: python -m femm_gpu_amd.sample REQUEST.json RESPONSE.json", file=sys.stderr)
        return 2
    return run_sample(argv[0], argv[1])


if __name__ == "__main__":
    sys.exit(main())
