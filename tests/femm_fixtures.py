"""Readers for the frozen stock-FEMM fixtures and the small analytic fixtures.

The readers follow the reference's test readers: gpu_solver/tests/native/reference.cpp
(ReadFemmAns :14-120, ReadNonlinearFemmReference :227-383, ReadFrozenPostprocessReference
:160-225) and gpu_solver/tests/native/fixtures.cpp (analytic models).
"""

import math
from pathlib import Path

import numpy as np

from femm_gpu_amd.model import MU0, Material, Model, load_artifact
from femm_gpu_amd.wire import write_artifact  # noqa: F401  (the fixtures' artifact writer)

GOLDEN = Path(__file__).resolve().parent / "golden" / "femm"


def _solution_block(path):
    lines = Path(path).read_text(encoding="utf-8").splitlines()
    start = next(i for i, l in enumerate(lines) if l.strip() == "[Solution]")
    tokens = " ".join(lines[start + 1:]).split()
    return tokens


def read_ans(path, scale=1.0):
    """Nodes (x, y, A), elements (n0, n1, n2, label, e0, e1, e2) of a FEMM .ans [Solution]."""
    tok = _solution_block(path)
    pos = 0
    n = int(tok[pos]); pos += 1
    xy, a = np.zeros((n, 2)), np.zeros(n)
    for i in range(n):
        xy[i] = float(tok[pos]) * scale, float(tok[pos + 1]) * scale
        a[i] = float(tok[pos + 2])
        pos += 4
    t = int(tok[pos]); pos += 1
    tri, label, edges = np.zeros((t, 3), np.int32), np.zeros(t, np.int32), np.zeros((t, 3), np.int32)
    for e in range(t):
        tri[e] = int(tok[pos]), int(tok[pos + 1]), int(tok[pos + 2])
        label[e] = int(tok[pos + 3])
        edges[e] = int(tok[pos + 4]), int(tok[pos + 5]), int(tok[pos + 6])
        pos += 8
    return xy, a, tri, label, edges


def _boundary_nodes(tri, edges, n):
    boundary = np.zeros(n, bool)
    for e in range(len(tri)):
        for side in range(3):
            if edges[e, side] >= 0:
                boundary[tri[e, side]] = True
                boundary[tri[e, (side + 1) % 3]] = True
    return np.nonzero(boundary)[0].astype(np.int32)


def read_key_values(path):
    out = {}
    for line in Path(path).read_text(encoding="utf-8").splitlines():
        if "=" in line and not line.startswith("airgap_sample_deg"):
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def read_bh_table(path=GOLDEN / "35PN230_v1" / "35PN230.tab"):
    """FEMM .tab: header then (H, B) pairs (nonlinear_model.cpp:150-164)."""
    lines = Path(path).read_text(encoding="utf-8").splitlines()
    assert "H (A_per_meter)" in lines[0] and "B (tesla)" in lines[0]
    h, b = [], []
    for line in lines[1:]:
        parts = line.split()
        if len(parts) == 2:
            h.append(float(parts[0]))
            b.append(float(parts[1]))
    return np.asarray(b), np.asarray(h)


def nonlinear_pm_coil():
    """The frozen FEMM nonlinear PM/coil model (reference.cpp:227-383).

    Returns (Model, expected A, current, expected flux)."""
    stem = GOLDEN / "nonlinear_pm_coil_v1" / "nonlinear_pm_coil_planar"
    xy, a, tri, label, edges = read_ans(str(stem) + ".ans", scale=1e-3)
    p = xy[tri]
    area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                  - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    coil_area = 0.0
    for e in range(len(tri)):   # element order, as the reference accumulates
        if label[e] == 2:
            coil_area += area[e]
    b, h = read_bh_table()
    materials = [
        Material(1.0, bh_curve=0),                          # label 0: steel 35PN230
        Material(1.0 / (MU0 * 1.05), hc=900000.0),          # label 1: PM
        Material(1.0 / MU0, source_j_per_a=50.0 / coil_area, circuit=0),   # label 2: coil
        Material(1.0 / MU0),                                # label 3: air
    ]
    dn = _boundary_nodes(tri, edges, len(xy))
    model = Model(xy, tri, label, materials, [(b, h)], dn, np.zeros(len(dn)), None, 1, 0.020,
                  name="nonlinear_pm_coil")
    summary = read_key_values(str(stem) + ".summary.txt")
    return model, a, float(summary["circuit_current_A"]), float(summary["flux_linkage_Wb"])


def nonlinear_pm_coil_postprocess():
    """Frozen WST loads and smoothed air-gap samples (reference.cpp:160-225)."""
    stem = GOLDEN / "nonlinear_pm_coil_v1" / "nonlinear_pm_coil_planar"
    text = Path(str(stem) + ".postprocess.txt").read_text(encoding="utf-8").splitlines()
    kv, angles, radial = {}, [], []
    for line in text:
        if line.startswith("airgap_sample_deg="):
            fields = dict(f.split("=") for f in line.split())
            angles.append(float(fields["airgap_sample_deg"]) * math.pi / 180.0)
            radial.append(float(fields["radial_B_T"]))
        elif "=" in line:
            k, v = line.split("=", 1)
            kv[k] = v
    return dict(selected_label=int(kv["selected_material_label"]), fx=float(kv["Fx_N"]),
                fy=float(kv["Fy_N"]), torque=float(kv["torque_Nm"]),
                radius_m=float(kv["airgap_radius_mm"]) * 1e-3, angles=np.asarray(angles),
                radial_b=np.asarray(radial), flux=float(kv["flux_linkage_Wb"]),
                current=float(kv["circuit_current_A"]))


def linear_square():
    """Frozen linear FEMM square (reference.cpp:14-120): nu = 1/mu0, J_per_A = 1, depth 1 m."""
    stem = GOLDEN / "linear_square_v1" / "linear_square"
    xy, a, tri, label, edges = read_ans(str(stem) + ".ans")
    dn = _boundary_nodes(tri, edges, len(xy))
    model = Model(xy, tri, np.zeros(len(tri), np.int32),
                  [Material(1.0 / MU0, source_j_per_a=1.0, circuit=0)], [], dn,
                  np.zeros(len(dn)), None, 1, 1.0, name="linear_square")
    parts = Path(str(stem) + ".circuit.txt").read_text(encoding="utf-8").split()
    return model, a, float(parts[1]), float(parts[5])


def generic_reference(name):
    """A generic artifact fixture with its frozen FEMM answer (test_femm_gpu_cli.py:223-270)."""
    root = GOLDEN / f"{name}_v1"
    model, currents = load_artifact(root / f"{name}.gpu.json")
    xy, a, _, _, _ = read_ans(root / f"{name}.ans")
    by_coord = {(round(x, 12), round(y, 12)): v for (x, y), v in zip(xy, a)}
    expected = np.asarray([by_coord[(round(x * 1e3, 12), round(y * 1e3, 12))]
                           for x, y in model.node_xy])
    summary = read_key_values(root / f"{name}.summary.txt")
    return model, float(summary["circuit_current_A"]), expected, float(summary["flux_linkage_Wb"])


def unit_square_nonlinear(source=1.0):
    """UnitSquareFixture (fixtures.cpp:6-16) as a NonlinearModel: nu = 1, J_per_A = 1."""
    xy = [[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]]
    tri = [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]
    return Model(xy, tri, [0, 0, 0, 0], [Material(1.0, source_j_per_a=source, circuit=0)], [],
                 [0, 1, 2, 3], [0.0] * 4, None, 1, 1.0, name="unit_square")


def three_region():
    """NonlinearThreeRegionFixture (fixtures.cpp:86-103): PM, coil and B-H steel."""
    xy = [[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]]
    tri = [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]
    mats = [Material(1.0, hc=10.0), Material(1.0, source_j_per_a=12.0, circuit=0),
            Material(0.5, alpha=1.0, bh_curve=0)]
    return Model(xy, tri, [0, 1, 2, 2], mats, [(np.array([0.0, 1.0, 2.0]),
                                                np.array([0.0, 1.0, 4.0]))],
                 [0, 1, 2, 3], [0.0] * 4, None, 1, 1.0, name="three_region")


def two_circuit():
    """NonlinearTwoCircuitFixture (fixtures.cpp:105-120)."""
    xy = [[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]]
    tri = [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]
    mats = [Material(1.0), Material(1.0, source_j_per_a=12.0, circuit=0),
            Material(1.0, source_j_per_a=6.0, circuit=1)]
    return Model(xy, tri, [1, 0, 2, 0], mats, [], [0, 1, 2, 3], [0.0] * 4, None, 2, 1.0,
                 name="two_circuit")


def postprocess_ring():
    """PostprocessRingFixture (fixtures.cpp:122-138)."""
    xy = [[-1, -1], [1, -1], [1, 1], [-1, 1], [-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]]
    tri = [[0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7],
           [4, 5, 6], [4, 6, 7]]
    mats = [Material(1.0 / MU0) for _ in range(4)]
    return Model(xy, tri, [3, 3, 3, 3, 3, 3, 3, 3, 1, 1], mats, [], [0, 1, 2, 3],
                 [-1.0, 1.0, 1.0, -1.0], None, 1, 1.0, name="postprocess_ring")


