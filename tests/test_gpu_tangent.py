"""Shared-operator tangent solves on the GPU (femm_tangent_multi_rhs) against the oracle.

Bars: per direction status OK, dA within 1e-9 relative L2 and dflux within 1e-9 relative of
the oracle's restatement of SolveTangentMultiRhs at the same base state and options, PCG
iteration counts within 2 %; element B of the derivative equals finalizing dA.
"""

import sys

import numpy as np
import pytest

import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.model import plain_options
from femm_gpu_amd.solver import Plan

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _rel(x, y):
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))


@pytest.mark.parametrize("case", ["pm_coil", "motor_small", "ring_linear"])
def test_tangent_matches_oracle(case):
    say(f"case tangent {case}")
    if case == "pm_coil":
        model, _, current, _ = ff.nonlinear_pm_coil()
        cur = np.array([current])
        dirs = np.array([[1.0], [-2.5]])
    elif case == "motor_small":
        model, _ = meshes.motor_sector(angular_cells=160)
        cur = meshes.motor_currents(1, seed=12)[0]
        dirs = np.vstack([np.eye(3), [[1.0, -0.5, -0.5]]])
    else:
        model, cur = meshes.square_coil_ring(cells=24, seed=5)
        dirs = np.array([[1.0]])
    o = plain_options(relative_tolerance=1e-10, linear_relative_tolerance=1e-12,
                      max_newton_iterations=200, max_linear_iterations=100000)
    base = ol.newton_solve(model, cur, o)
    assert base["status"] == 0
    # Tangent solves at the production linear tolerance 1e-10.
    ot = plain_options(relative_tolerance=1e-10, linear_relative_tolerance=1e-10,
                       max_linear_iterations=20000)
    ref = ol.tangent(model, base["a"], cur, dirs, ot)
    plan = Plan(model)
    out = plan.tangent_multi_rhs(base["a"], cur, dirs, ot, fields=True)
    bx, by, flux = plan.finalize_batch(out["da"])
    for d in range(len(dirs)):
        r = out["results"][d]
        say(f"  direction {d}: status {r['status']} pcg {r['linear_iterations']} (oracle "
            f"{ref['results'][d]['linear']}) dA {_rel(out['da'][d], ref['da'][d]):.2e} "
            f"dflux {_rel(out['dflux'][d], ref['dflux'][d]):.2e}")
        assert r["status"] == 0 and ref["results"][d]["status"] == 0
        it_ref = ref["results"][d]["linear"]
        assert abs(int(r["linear_iterations"]) - it_ref) <= max(3, it_ref // 50)
        assert _rel(out["da"][d], ref["da"][d]) <= 1e-9
        assert _rel(out["dflux"][d], ref["dflux"][d]) <= 1e-9
        assert np.array_equal(out["dbx"][d], bx[d]) and np.array_equal(out["dby"][d], by[d])
        assert np.array_equal(out["dflux"][d], flux[d])
    again = plan.tangent_multi_rhs(base["a"], cur, dirs, ot)
    assert np.array_equal(again["da"], out["da"])
    plan.close()


def test_tangent_rejects_non_finite_directions():
    model, _, current, _ = ff.nonlinear_pm_coil()
    plan = Plan(model)
    o = plain_options(relative_tolerance=1e-10)
    out = plan.tangent_multi_rhs(np.zeros(plan.N), [current], [[1.0], [float("inf")]], o)
    assert list(out["results"]["status"]) == [2, 2]
    plan.close()
