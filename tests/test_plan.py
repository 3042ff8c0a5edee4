"""Host-side plan of the MI355X library (no GPU needed): the symbolic CSR, DOF map and
validation codes must equal the oracle's restatement of the reference bit for bit
(nonlinear_topology.cpp:37-85, signed_node_constraints.h:33-106, nonlinear_model.cpp:223-319).
Also checks that the C-ABI library loads and exports every declared symbol."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.solver import FemmError, Plan, library

ROOT = Path(__file__).resolve().parents[1]


def _models():
    yield "three_region", ff.three_region()
    yield "two_circuit", ff.two_circuit()
    yield "pm_coil", ff.nonlinear_pm_coil()[0]
    yield "linear_square", ff.linear_square()[0]
    for name in ("periodic_strip", "antiperiodic_strip", "axisymmetric_coil"):
        yield name, ff.generic_reference(name)[0]
    yield "ring", ff.postprocess_ring()
    yield "square_coil_ring_32", meshes.square_coil_ring(cells=32)[0]
    yield "motor_small", meshes.motor_sector(angular_cells=160)[0]
    yield "solenoid", meshes.solenoid_rz(radial_cells=8, axial_cells=12)[0]


def test_library_exports_every_declared_symbol():
    lib = library()
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        names |= set(re.findall(r"^(?:int|void|const char \*)\s*\*?\s*(femm_\w+)\s*\(", text, re.M))
    assert len(names) >= 10
    for name in sorted(names):
        assert hasattr(lib, name), name


@pytest.mark.parametrize("name,model", list(_models()), ids=lambda v: v if isinstance(v, str) else "")
def test_symbolic_plan_equals_oracle(name, model):
    plan = Plan(model)
    mine = plan.symbolic()
    st, ref = ol.prepare_symbolic(model)
    assert st == 0
    for key in ("rows", "cols", "diag_slots", "free_nodes", "node_to_free", "node_sign", "boundary"):
        assert np.array_equal(mine[key], ref[key]), key
    plan.close()


def test_motor_plan_sizes_and_pattern_invariants():
    model, _ = meshes.motor_sector()
    plan = Plan(model)
    s = plan.symbolic()
    assert (plan.N, plan.T) == (58682, 116160)
    n = plan.n
    rows, cols = s["rows"], s["cols"]
    # sorted unique columns, diagonal present, symmetric pattern, nnz <= 9T
    for r in range(0, n, 997):
        c = cols[rows[r]:rows[r + 1]]
        assert np.all(np.diff(c) > 0) and r in c
    dense_pairs = set(zip(np.repeat(np.arange(n), np.diff(rows)).tolist(), cols.tolist()))
    assert all((c, r) in dense_pairs for r, c in list(dense_pairs)[:20000])
    assert plan.nnz <= 9 * plan.T
    # anti-periodic pairs fold into one signed DOF
    for a, b, sign in model.constraints[:50]:
        assert s["node_to_free"][a] == s["node_to_free"][b] >= 0
        assert s["node_sign"][b] == sign * s["node_sign"][a]
    plan.close()


@pytest.mark.parametrize("mutation,code", [
    ("bad_material", 15), ("clockwise", 4), ("no_dirichlet", 2), ("bad_circuit", 15),
    ("contradictory_cycle", 5), ("self_pair", 5), ("all_fixed", 5),
])
def test_validation_codes_match_oracle(mutation, code):
    m = ff.two_circuit()
    if mutation == "bad_material":
        m.triangle_material[0] = 9
    elif mutation == "clockwise":
        m.triangles[0] = m.triangles[0][[0, 2, 1]]
    elif mutation == "no_dirichlet":
        m.dirichlet_nodes = np.zeros(0, np.int32)
        m.dirichlet_values = np.zeros(0)
    elif mutation == "bad_circuit":
        m.materials[2].circuit = 2
    elif mutation == "contradictory_cycle":
        m.constraints = np.array([[0, 4, 1], [4, 2, 1], [2, 0, -1]], np.int32)
    elif mutation == "self_pair":
        m.constraints = np.array([[4, 4, -1]], np.int32)
    elif mutation == "all_fixed":
        m.dirichlet_nodes = np.arange(5, dtype=np.int32)
        m.dirichlet_values = np.zeros(5)
    m._c = None
    assert ol.validate(m) == code
    with pytest.raises(FemmError) as err:
        Plan(m)
    assert err.value.code == code


def test_spline_equals_oracle_through_plan():
    """The product's tridiagonal spline equals the oracle's dense elimination bit for bit:
    a model whose only nonlinear curve is the 35PN230 table yields identical plans and the
    oracle-spline evaluation is reproduced by the device path (checked on the GPU)."""
    b, h = ff.read_bh_table()
    ok, pb, ph, ps, cold, passes = ol.bh_spline(b, h)
    assert ok and passes == 0
    # smoothing branch (generator's synthetic curve) must also be accepted by the plan
    m = ff.three_region()
    m.curves = [(np.array([0.0, 1.0, 2.0, 3.0]), np.array([0.0, 1.0, 1.01, 10.0]))]
    m._c = None
    Plan(m).close()


def test_compute_without_gpu_reports_gpu_unavailable():
    text = ctypes.create_string_buffer(256)
    if library().femm_device_info(text, 256) == 0:   # the library's own device query
        pytest.skip("a GPU is present")
    plan = Plan(ff.three_region())
    with pytest.raises(FemmError) as err:
        plan.solve_batch(np.array([[2.0]]))
    assert err.value.code == 6
    plan.close()


def test_group_entry_points_validate_without_gpu():
    """femm_group_create checks its entries on the host; solving needs the GPU."""
    from femm_gpu_amd.solver import Group

    model, _ = meshes.square_coil_ring(cells=4, ring_bh=True)
    plan = Plan(model)
    spec = dict(plan=plan, items=1, currents=8, a=8, flux=8, results=8)
    with pytest.raises(FemmError) as err:
        Group([spec, dict(spec)])            # the same plan twice
    assert err.value.code == 2
    with pytest.raises(FemmError) as err:
        Group([dict(spec, partitions=0)])
    assert err.value.code == 2
    group = Group([spec])
    if library().femm_device_info(ctypes.create_string_buffer(256), 256) != 0:   # no GPU
        with pytest.raises(FemmError) as err:
            group.solve(0)
        assert err.value.code == 6           # GPU_UNAVAILABLE
    group.close()
    plan.close()


def test_empty_batch_is_an_empty_result_without_gpu():
    """SolveBatch returns an empty result for empty currents (nonlinear_solver.cu:189-192):
    items = 0 succeeds on every Newton batch entry point without touching a pointer or the
    device (so it holds on hosts without a GPU), while negative counts stay invalid."""
    plan = Plan(ff.three_region())
    for safeguarded in (False, True):
        out = plan.solve_batch(np.zeros((0, plan.C)), safeguarded=safeguarded, fields=True)
        assert out["a"].shape == (0, plan.N) and out["flux"].shape == (0, plan.C)
        assert out["results"].shape == (0,) and out["bx"].shape == (0, plan.T)
        plan.solve_batch_device(0, 0, 0, 0, 0, 0, safeguarded=safeguarded)
    lib = library()
    null = ctypes.c_void_p(0)
    for name in ("femm_newton_batch", "femm_newton_batch_safeguarded"):
        fn = getattr(lib, name)
        assert fn(plan.handle, 0, null, null, null, null, null, null, null, null, null) == 0
        assert fn(plan.handle, -1, null, null, null, null, null, null, null, null, null) == 2
    assert lib.femm_newton_batch(None, 0, null, null, null, null, null, null, null, null,
                                 null) == 2
    plan.close()


_ALLOCATION_PROBE = r"""
import ctypes, re, resource, sys
sys.path[:0] = [sys.argv[1]]
from femm_gpu_amd import meshes
from femm_gpu_amd.solver import library
lib = library()
model, _ = meshes.square_coil_ring(cells=256)
cm = model.c_struct()
vm = int(re.search(r"VmSize:\s+(\d+) kB", open("/proc/self/status").read()).group(1)) * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (16 << 20), resource.RLIM_INFINITY))
handle = ctypes.c_void_p()
limited = lib.femm_plan_create(ctypes.byref(cm), ctypes.byref(handle))
resource.setrlimit(resource.RLIMIT_AS, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
print("limited", limited, handle.value is None)
again = lib.femm_plan_create(ctypes.byref(cm), ctypes.byref(handle))
print("unlimited", again, handle.value is not None)
lib.femm_plan_destroy(handle)
"""


def test_host_allocation_failure_is_a_status_not_an_abort():
    """A std::bad_alloc inside femm_plan_create (a 66k-node model under an address-space
    limit that leaves 16 MiB of headroom) must come back as FEMM_GPU_ALLOCATION_FAILED with
    no handle, not cross the C ABI; the same call then succeeds without the limit."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, "-c", _ALLOCATION_PROBE, str(ROOT / "femm-gpu_amd")],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = dict(line.split(" ", 1) for line in out.stdout.strip().splitlines())
    assert lines["limited"] == "7 True"
    assert lines["unlimited"] == "0 True"
