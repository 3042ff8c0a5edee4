"""ctypes binding of the wire contract (include/femm_wire.h): the command line
(``femm_cli_main``), the strict mesh-artifact parser and SHA-256.

The adapters themselves are C++ (femm-gpu_amd/csrc/host); this module only forwards to the
library, so a Python caller and the ``bin/femm_gpu_amd`` executable run the same code.
"""

from __future__ import annotations

import ctypes
import hashlib
import json
from pathlib import Path

import numpy as np

from .model import MU0, AirGap, CAgePoint, CAirGap, CMaterial, CModel, Material, Model
from .solver import library

P = ctypes.POINTER


class ArtifactError(ValueError):
    """The artifact parser rejected a document (the message is the parser's reason)."""


def cli(args) -> int:
    """Runs the command line with ``args`` (without the program name); returns its exit
    status (0 PASS, 1 FAIL / not published, 2 usage)."""
    argv = [b"femm_gpu_amd"] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * (len(argv) + 1))(*argv, None)
    return int(library().femm_cli_main(len(argv), arr))


def sha256_hex(data: bytes) -> str:
    out = ctypes.create_string_buffer(65)
    code = library().femm_sha256_hex(data, len(data), out)
    if code != 0:
        raise ArtifactError(f"femm_sha256_hex: status {code}")
    return out.value.decode()


def parse_artifact(text, name="artifact"):
    """(Model, circuit currents) of an artifact document (bytes or str), parsed by the
    library's strict parser; raises ArtifactError with the parser's reason."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    lib = library()
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    code = lib.femm_artifact_parse(data, len(data), ctypes.byref(handle), err, 512)
    if code != 0:
        raise ArtifactError(err.value.decode() or f"status {code}")
    try:
        m = CModel()
        if lib.femm_artifact_model(handle, ctypes.byref(m)) != 0:
            raise ArtifactError("femm_artifact_model failed")
        N, T, M = m.node_count, m.triangle_count, m.material_count
        xy = np.ctypeslib.as_array(m.node_xy, (2 * N,)).reshape(N, 2).copy()
        tri = np.ctypeslib.as_array(m.triangle_nodes, (3 * T,)).reshape(T, 3).copy()
        mat = np.ctypeslib.as_array(m.triangle_material, (T,)).copy()
        mats = []
        for k in range(M):
            c = m.materials[k]
            mats.append(Material(c.reluctivity_zero, c.alpha, c.source_j_per_a, c.hc,
                                 c.magnetization_deg, c.bh_curve, c.circuit, c.group))
        offsets = np.ctypeslib.as_array(m.curve_offsets, (m.curve_count + 1,)).copy()
        curves = []
        if m.curve_count:
            b = np.ctypeslib.as_array(m.curve_b, (int(offsets[-1]),)).copy()
            h = np.ctypeslib.as_array(m.curve_h, (int(offsets[-1]),)).copy()
            curves = [(b[offsets[k]:offsets[k + 1]], h[offsets[k]:offsets[k + 1]])
                      for k in range(m.curve_count)]
        D = m.dirichlet_count
        dn = np.ctypeslib.as_array(m.dirichlet_nodes, (D,)).copy() if D else np.zeros(0, np.int32)
        dv = np.ctypeslib.as_array(m.dirichlet_values, (D,)).copy() if D else np.zeros(0)
        cons = np.asarray([(m.constraints[k].node_a, m.constraints[k].node_b,
                            m.constraints[k].sign) for k in range(m.constraint_count)],
                          np.int32).reshape(-1, 3)
        currents = np.zeros(max(1, m.circuit_count))
        lib.femm_artifact_currents(handle, currents.ctypes.data_as(P(ctypes.c_double)))
        gaps = [_air_gap(m.air_gaps[g]) for g in range(m.air_gap_count)]
        model = Model(xy, tri, mat, mats, curves, dn, dv, cons, m.circuit_count, m.depth_m,
                      m.problem_type, name=str(name), air_gaps=gaps)
        return model, currents[: m.circuit_count]
    finally:
        lib.femm_artifact_destroy(handle)


def _air_gap(c) -> AirGap:
    rec = [c.points[k] for k in range(c.cells + 1)]
    return AirGap(inner_radius_m=c.inner_radius_m, outer_radius_m=c.outer_radius_m,
                  arc_length_deg=c.arc_length_deg, nodes=[list(r.node) for r in rec],
                  weights=[list(r.weight) for r in rec], inner_shift=c.inner_shift,
                  outer_shift=c.outer_shift, antiperiodic=bool(c.antiperiodic),
                  center_xy=(c.center_xy[0], c.center_xy[1]))


def rotate_air_gaps(model: Model, angle_deg: float, posed_deg: float = 0.0) -> Model:
    """The sliding-band model turned from `posed_deg` to `angle_deg` (femm_air_gaps_rotate:
    the reference's ApplySlidingBandAngle, io/sample_request.cpp:448-491): a copy of `model`
    whose air-gap elements carry the rotated inner records and shift."""
    c = model.c_struct()
    count = c.air_gap_count
    total = sum(g.cells + 1 for g in model.air_gaps)
    out = (CAirGap * max(1, count))()
    pts = (CAgePoint * max(1, total))()
    code = library().femm_air_gaps_rotate(count, c.air_gaps, ctypes.c_double(posed_deg),
                                          ctypes.c_double(angle_deg), out, pts,
                                          ctypes.c_size_t(total))
    if code != 0:
        raise ValueError(f"femm_air_gaps_rotate: status {code}")
    rotated = Model(model.node_xy, model.triangles, model.triangle_material, model.materials,
                    model.curves, model.dirichlet_nodes, model.dirichlet_values,
                    model.constraints, model.circuit_count, model.depth_m, model.problem_type,
                    name=f"{model.name}@{angle_deg:g}deg",
                    air_gaps=[_air_gap(out[g]) for g in range(count)])
    return rotated


def write_artifact(path, model, currents, schema="gpu_femm_planar_dc_mesh_v1", turns=None,
                   source_sha="a" * 64, base_sha="b" * 64, rotor_angle_deg=0.0,
                   displacement_mm=(0.0, 0.0), age_scalar=False, edit=None):
    """Write `model` as a strictly valid mesh artifact (the layout the reference's preparer
    writes, io/mesh_artifact.cpp:14-440): one material and one region per material label
    (group number = label, FEMM's region id = label), integer ids, named series circuits,
    node constraints for the magnetostatic schema, identities for the motor schema and, for the
    neutral schemas, canonical_identity_sha256 = SHA-256 of the canonical `resolved` (sorted
    keys, no whitespace, numbers as written).  `turns[label]` sets a coil region's turns
    (default: its J per ampere x its area).  The sliding-band schema gpu_femm_mesh_v2 also
    carries the model's air-gap elements (mesh_artifact.cpp:610-697; `age_scalar` writes a
    single element as an object, the spelling MATLAB's jsonencode produces).  `edit(resolved)`
    may change the document before it is written.  Returns the SHA-256 of the bytes written."""
    areas = model.element_areas()
    mats, regions = [], []
    for i, m in enumerate(model.materials):
        b, h = (model.curves[m.bh_curve] if m.bh_curve >= 0 else ([], []))
        mu = 1.0 / (MU0 * m.reluctivity_zero)
        mats.append(dict(id=i, name=f"material_{i}", mu_x=mu, mu_y=mu, H_c_A_per_m=float(m.hc),
                         B_T=[float(x) for x in b], H_A_per_m=[float(x) for x in h],
                         lam_type=0, lam_fill=1))
        if m.circuit >= 0:
            t = turns[i] if turns is not None else m.source_j_per_a * float(
                areas[model.triangle_material == i].sum())
        else:
            t = 0.0
        regions.append(dict(id=i, group_number=i, material_id=i, circuit_index=int(m.circuit),
                            turns=float(t), pm_magnetization_deg=float(m.magnetization_deg)))
    motor = schema in ("gpu_femm_mesh_v1", "gpu_femm_mesh_v2")
    resolved = dict(
        source_fem_sha256=source_sha,
        model=dict(depth_mm=model.depth_m * 1e3,
                   problem_type="axisymmetric" if model.problem_type == 1 else "planar",
                   frequency_hz=0),
        nodes_mm=(model.node_xy * 1e3).tolist(),
        triangles=dict(node_indices=model.triangles.tolist(),
                       region_ids=[int(k) for k in model.triangle_material]),
        regions=regions, materials=mats,
        circuits=[dict(index=k, name=f"circuit_{k}", type="series", current_A=float(c))
                  for k, c in enumerate(currents)],
        outer_dirichlet=dict(node_indices=model.dirichlet_nodes.tolist(),
                             A_Wb_per_m=model.dirichlet_values.tolist()))
    if schema == "gpu_femm_magnetostatic_mesh_v1":
        resolved["node_constraints"] = [
            dict(node_a=int(a), node_b=int(b), relation="periodic" if s > 0 else "antiperiodic")
            for a, b, s in model.constraints]
    if motor:
        resolved["base_motor_fem_sha256"] = base_sha
        resolved["pose"] = dict(rotor_angle_deg=rotor_angle_deg,
                                displacement_mm=list(displacement_mm))
    if schema == "gpu_femm_mesh_v2":
        ages = []
        for g, age in enumerate(model.air_gaps):
            points = [dict(n0=int(n[0]), w0=float(w[0]), n1=int(n[1]), w1=float(w[1]),
                           n2=int(n[2]), w2=float(w[2]), n3=int(n[3]), w3=float(w[3]))
                      for n, w in zip(age.nodes, age.weights)]
            ages.append(dict(name=f"gap_{g}",
                             periodicity="antiperiodic" if age.antiperiodic else "periodic",
                             center_mm=[1e3 * age.center_xy[0], 1e3 * age.center_xy[1]],
                             ri_mm=1e3 * age.inner_radius_m, ro_mm=1e3 * age.outer_radius_m,
                             arc_length_deg=age.arc_length_deg, sector_count=age.cells,
                             inner_shift=age.inner_shift, outer_shift=age.outer_shift,
                             quad_points=points))
        resolved["air_gap_elements"] = ages[0] if age_scalar and len(ages) == 1 else ages
    if edit is not None:
        edit(resolved)
    canonical = json.dumps(resolved, sort_keys=True, separators=(",", ":"))
    doc = dict(schema_version=schema, source_fem_sha256=source_sha,
               canonical_identity_sha256=hashlib.sha256(canonical.encode()).hexdigest(),
               resolved=resolved)
    if motor:
        doc["base_motor_fem_sha256"] = base_sha
    data = json.dumps(doc).encode()
    Path(path).write_bytes(data)
    return hashlib.sha256(data).hexdigest()


def load_artifact(path):
    """(Model, circuit currents) of the artifact file at ``path`` (strict parse)."""
    return parse_artifact(Path(path).read_bytes(), name=str(path))


__all__ = ["ArtifactError", "cli", "sha256_hex", "parse_artifact", "load_artifact",
           "write_artifact", "rotate_air_gaps"]
