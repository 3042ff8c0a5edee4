"""Model container and its C view (the femm_model_t contract of include/femm_model.h).

A :class:`Model` holds the fields of the reference's ``NonlinearModel``
(gpu_solver/src/physics/nonlinear_model.h:74-97) as numpy arrays in SI units.
``Model.c_struct()`` returns a ``femm_model_t`` whose pointers stay valid while
the Model object is alive.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

MU0 = 4.0e-7 * math.pi
PLANAR = 0
AXISYMMETRIC = 1

STATUS_NAMES = [
    "OK", "INPUT_IO", "INVALID_ARGUMENT", "UNSUPPORTED_FEATURE", "MESH_INVALID",
    "BOUNDARY_INVALID", "GPU_UNAVAILABLE", "GPU_ALLOCATION_FAILED", "ASSEMBLY_FAILED",
    "LINEAR_SOLVE_NOT_CONVERGED", "LINEAR_SOLVE_BREAKDOWN", "NUMERICAL_NONFINITE",
    "OUTPUT_IO", "INTERNAL_ERROR", "NONLINEAR_SOLVE_NOT_CONVERGED", "INVALID_MATERIAL",
]


def status_name(code: int) -> str:
    """Stable status names of gpu_solver/src/core/types.cpp:6-43."""
    return STATUS_NAMES[code] if 0 <= code < len(STATUS_NAMES) else "UNKNOWN"


class CMaterial(ctypes.Structure):
    _fields_ = [
        ("reluctivity_zero", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("source_j_per_a", ctypes.c_double),
        ("hc", ctypes.c_double),
        ("magnetization_deg", ctypes.c_double),
        ("bh_curve", ctypes.c_int32),
        ("circuit", ctypes.c_int32),
        ("group", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class CConstraint(ctypes.Structure):
    _fields_ = [("node_a", ctypes.c_int32), ("node_b", ctypes.c_int32), ("sign", ctypes.c_int32)]


class CAgePoint(ctypes.Structure):
    _fields_ = [("node", ctypes.c_int32 * 4), ("weight", ctypes.c_double * 4)]


class CAirGap(ctypes.Structure):
    _fields_ = [
        ("antiperiodic", ctypes.c_int32),
        ("cells", ctypes.c_int32),
        ("center_xy", ctypes.c_double * 2),
        ("inner_radius_m", ctypes.c_double),
        ("outer_radius_m", ctypes.c_double),
        ("arc_length_deg", ctypes.c_double),
        ("inner_shift", ctypes.c_double),
        ("outer_shift", ctypes.c_double),
        ("points", ctypes.POINTER(CAgePoint)),
    ]


class CModel(ctypes.Structure):
    _fields_ = [
        ("problem_type", ctypes.c_int32),
        ("node_count", ctypes.c_int32),
        ("node_xy", ctypes.POINTER(ctypes.c_double)),
        ("triangle_count", ctypes.c_int32),
        ("triangle_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("triangle_material", ctypes.POINTER(ctypes.c_int32)),
        ("material_count", ctypes.c_int32),
        ("materials", ctypes.POINTER(CMaterial)),
        ("curve_count", ctypes.c_int32),
        ("curve_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("curve_b", ctypes.POINTER(ctypes.c_double)),
        ("curve_h", ctypes.POINTER(ctypes.c_double)),
        ("dirichlet_count", ctypes.c_int32),
        ("dirichlet_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("dirichlet_values", ctypes.POINTER(ctypes.c_double)),
        ("constraint_count", ctypes.c_int32),
        ("constraints", ctypes.POINTER(CConstraint)),
        ("circuit_count", ctypes.c_int32),
        ("depth_m", ctypes.c_double),
        ("air_gap_count", ctypes.c_int32),
        ("air_gaps", ctypes.POINTER(CAirGap)),
    ]


class CNewtonOptions(ctypes.Structure):
    _fields_ = [
        ("relative_tolerance", ctypes.c_double),
        ("linear_relative_tolerance", ctypes.c_double),
        ("max_newton_iterations", ctypes.c_int32),
        ("max_linear_iterations", ctypes.c_int32),
        ("recover_nonconvergence", ctypes.c_int32),
        ("adaptive_linear_tolerance", ctypes.c_int32),
        ("require_equation_residual", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


def production_options(**overrides) -> CNewtonOptions:
    """ProductionSampleOptions (gpu_solver/src/physics/nonlinear_model.cpp:195-205)."""
    values = dict(relative_tolerance=1e-8, linear_relative_tolerance=1e-10,
                  max_newton_iterations=128, max_linear_iterations=16384,
                  recover_nonconvergence=1, adaptive_linear_tolerance=1)
    values.update(overrides)
    return CNewtonOptions(**values)


def accepted_options(**overrides) -> CNewtonOptions:
    """ProductionSampleOptions plus the equation-residual gate: an item is accepted only when
    its undamped step is < 1e-6 and its equation residual is <= 1e-8."""
    values = dict(require_equation_residual=1)
    values.update(overrides)
    return production_options(**values)


def plain_options(**overrides) -> CNewtonOptions:
    """Default NonlinearOptions (nonlinear_model.h:99-106) with a chosen tolerance."""
    values = dict(relative_tolerance=1e-12, linear_relative_tolerance=1e-13,
                  max_newton_iterations=32, max_linear_iterations=128,
                  recover_nonconvergence=0, adaptive_linear_tolerance=0)
    values.update(overrides)
    return CNewtonOptions(**values)


def _ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(ctypes.POINTER(ctype))


@dataclass
class Material:
    reluctivity_zero: float
    alpha: float = 0.0
    source_j_per_a: float = 0.0
    hc: float = 0.0
    magnetization_deg: float = 0.0
    bh_curve: int = -1
    circuit: int = -1
    group: int = -1


@dataclass
class AirGap:
    """FEMM's native air-gap element (femm_air_gap_t): the annulus between an inner and an
    outer ring cut into `cells` equal angular cells.  nodes / weights are [cells + 1, 4]:
    per cell boundary the inner ring's (start, end) nodes and the outer ring's (start, end)
    nodes, with weights +-1."""
    inner_radius_m: float
    outer_radius_m: float
    arc_length_deg: float
    nodes: np.ndarray
    weights: np.ndarray
    inner_shift: float = 0.0
    outer_shift: float = 0.0
    antiperiodic: bool = False
    center_xy: tuple = (0.0, 0.0)

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.int32).reshape(-1, 4)
        self.weights = np.ascontiguousarray(self.weights, dtype=np.float64).reshape(-1, 4)

    @property
    def cells(self) -> int:
        return int(self.nodes.shape[0]) - 1


@dataclass
class Model:
    node_xy: np.ndarray                      # [N, 2] float64, metres
    triangles: np.ndarray                    # [T, 3] int32, CCW
    triangle_material: np.ndarray            # [T] int32
    materials: list                          # list[Material]
    curves: list = field(default_factory=list)   # list[(B array, H array)] raw knots
    dirichlet_nodes: np.ndarray = None
    dirichlet_values: np.ndarray = None
    constraints: np.ndarray = None           # [K, 3] int32 (node_a, node_b, sign)
    circuit_count: int = 1
    depth_m: float = 1.0
    problem_type: int = PLANAR
    name: str = "model"
    air_gaps: list = field(default_factory=list)  # list[AirGap]

    def __post_init__(self):
        self.node_xy = np.ascontiguousarray(self.node_xy, dtype=np.float64).reshape(-1, 2)
        self.triangles = np.ascontiguousarray(self.triangles, dtype=np.int32).reshape(-1, 3)
        self.triangle_material = np.ascontiguousarray(self.triangle_material, dtype=np.int32)
        if self.dirichlet_nodes is None:
            self.dirichlet_nodes = np.zeros(0, np.int32)
        self.dirichlet_nodes = np.ascontiguousarray(self.dirichlet_nodes, dtype=np.int32)
        if self.dirichlet_values is None:
            self.dirichlet_values = np.zeros(len(self.dirichlet_nodes))
        self.dirichlet_values = np.ascontiguousarray(self.dirichlet_values, dtype=np.float64)
        if self.constraints is None:
            self.constraints = np.zeros((0, 3), np.int32)
        self.constraints = np.ascontiguousarray(self.constraints, dtype=np.int32).reshape(-1, 3)
        self._c = None

    @property
    def node_count(self) -> int:
        return int(self.node_xy.shape[0])

    @property
    def triangle_count(self) -> int:
        return int(self.triangles.shape[0])

    def c_struct(self) -> CModel:
        """Build (once) and return the femm_model_t view of this model."""
        if self._c is not None:
            return self._c
        mats = (CMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = CMaterial(m.reluctivity_zero, m.alpha, m.source_j_per_a, m.hc,
                                m.magnetization_deg, m.bh_curve, m.circuit, m.group, 0)
        offsets = [0]
        bs, hs = [], []
        for b, h in self.curves:
            bs.append(np.asarray(b, np.float64))
            hs.append(np.asarray(h, np.float64))
            offsets.append(offsets[-1] + len(b))
        self._curve_offsets = np.asarray(offsets, np.int32)
        self._curve_b = np.ascontiguousarray(np.concatenate(bs) if bs else np.zeros(1))
        self._curve_h = np.ascontiguousarray(np.concatenate(hs) if hs else np.zeros(1))
        cons = (CConstraint * max(1, len(self.constraints)))()
        for i, (a, b, s) in enumerate(self.constraints):
            cons[i] = CConstraint(int(a), int(b), int(s))
        self._mats, self._cons = mats, cons
        gaps = (CAirGap * max(1, len(self.air_gaps)))()
        self._age_points = []
        for g, age in enumerate(self.air_gaps):
            pts = (CAgePoint * max(1, age.nodes.shape[0]))()
            for k in range(age.nodes.shape[0]):
                pts[k].node[:] = [int(v) for v in age.nodes[k]]
                pts[k].weight[:] = [float(w) for w in age.weights[k]]
            self._age_points.append(pts)
            gaps[g] = CAirGap(int(bool(age.antiperiodic)), age.cells,
                              (ctypes.c_double * 2)(*age.center_xy), age.inner_radius_m,
                              age.outer_radius_m, age.arc_length_deg, age.inner_shift,
                              age.outer_shift, pts)
        self._gaps = gaps
        self._c = CModel(
            self.problem_type, self.node_count, _ptr(self.node_xy, ctypes.c_double),
            self.triangle_count, _ptr(self.triangles, ctypes.c_int32),
            _ptr(self.triangle_material, ctypes.c_int32), len(self.materials), mats,
            len(self.curves), _ptr(self._curve_offsets, ctypes.c_int32),
            _ptr(self._curve_b, ctypes.c_double), _ptr(self._curve_h, ctypes.c_double),
            len(self.dirichlet_nodes), _ptr(self.dirichlet_nodes, ctypes.c_int32),
            _ptr(self.dirichlet_values, ctypes.c_double), len(self.constraints), cons,
            self.circuit_count, self.depth_m, len(self.air_gaps), gaps)
        return self._c

    def element_areas(self) -> np.ndarray:
        p = self.node_xy[self.triangles]
        return 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                      - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))


def load_artifact(path):
    """Load a mesh artifact (``gpu_femm_magnetostatic_mesh_v1``, ``gpu_femm_planar_dc_mesh_v1``
    or ``gpu_femm_mesh_v1``) with the library's strict parser (include/femm_wire.h,
    femm-gpu_amd/csrc/host/artifact.cpp; reference gpu_solver/src/io/mesh_artifact.cpp:14-440):
    mm -> m, nu0 = 1/(mu0 mu_r), base materials first and then one material label per region
    (group, circuit, PM angle, J_per_A = turns / region area), zero outer Dirichlet, signed node
    constraints; duplicate keys, unknown fields, a wrong canonical identity and every other
    rule violation raise ``wire.ArtifactError``.  Returns (Model, circuit currents)."""
    from .wire import load_artifact as strict_load
    return strict_load(path)
