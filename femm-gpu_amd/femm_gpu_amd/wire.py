"""ctypes binding of the wire contract (include/femm_wire.h): the command line
(``femm_cli_main``), the strict mesh-artifact parser and SHA-256.

The adapters themselves are C++ (femm-gpu_amd/csrc/host); this module only forwards to the
library, so a Python caller and the ``bin/femm_gpu_amd`` executable run the same code.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .model import CMaterial, CModel, Material, Model
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
        model = Model(xy, tri, mat, mats, curves, dn, dv, cons, m.circuit_count, m.depth_m,
                      m.problem_type, name=str(name))
        return model, currents[: m.circuit_count]
    finally:
        lib.femm_artifact_destroy(handle)


def load_artifact(path):
    """(Model, circuit currents) of the artifact file at ``path`` (strict parse)."""
    return parse_artifact(Path(path).read_bytes(), name=str(path))


__all__ = ["ArtifactError", "cli", "sha256_hex", "parse_artifact", "load_artifact"]
