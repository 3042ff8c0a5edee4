"""Block-Jacobi preconditioner on the GPU (config block_size 2, 4, 8) against the oracle.

* PCG alone (femm_pcg_batch_block): status, iteration count within max(3, 2 %) of the
  oracle's block-Jacobi PCG, x within 1e-7 relative, verified true residual <= 1e-10 ||b||;
  fewer iterations than point Jacobi.
* Newton solves: both gates per item, A within 1e-9 relative L2 of the oracle run with the
  same block size, flux within 1e-7, the same Newton count in the FP64 form; the mixed form
  (the reference driver with FP32 iterations) at the oracle driver's own distance from the
  converged solution + 1e-9 (tests/accuracy.py); repeats bitwise equal; the frozen FEMM
  pm-coil state within 1e-9.
"""

import sys

import numpy as np
import pytest

import accuracy as acc
import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.model import accepted_options, plain_options
from femm_gpu_amd.solver import Plan, mixed_config, reference_config

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _rel_l2(x, y):
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))


@pytest.mark.parametrize("bs", [2, 4, 8])
def test_block_pcg_matches_oracle(bs):
    say(f"case block pcg motor_small bs={bs}")
    model, cur = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    sym = plan.symbolic()
    state = np.zeros(model.node_count)
    code, values, rhs, diag = plan.assemble_batch(state[None], cur[None], False)
    assert code == 0
    x, info = plan.pcg_batch(values, rhs, 1e-10, 16384, block_size=bs)
    xr, st, it, res = ol.pcg_block(sym["rows"], sym["cols"], values[0], diag[0], rhs[0], 1e-10,
                                   16384, bs)
    _, _, it_point, _ = ol.pcg(sym["rows"], sym["cols"], values[0], diag[0], rhs[0], 1e-10, 16384)
    say(f"  gpu {int(info[0, 1])} iterations, oracle {it} (point Jacobi {it_point})")
    assert int(info[0, 0]) == 0 and st == 0
    assert abs(int(info[0, 1]) - it) <= max(3, it // 50)
    assert it < it_point
    assert np.linalg.norm(x[0] - xr) <= 1e-7 * np.linalg.norm(xr)
    rows, cols = sym["rows"], sym["cols"]
    y = np.array([values[0][rows[r]:rows[r + 1]] @ x[0][cols[rows[r]:rows[r + 1]]]
                  for r in range(len(x[0]))])
    assert np.linalg.norm(rhs[0] - y) <= 1.0001e-10 * np.linalg.norm(rhs[0])
    plan.close()


@pytest.mark.parametrize("form,bs", [("fp64", 4), ("fp64", 8), ("mixed", 4)])
def test_block_newton_matches_oracle(form, bs):
    say(f"case block newton {form} bs={bs}: pm_coil, motor_small x2")
    cfg = reference_config if form == "fp64" else mixed_config
    model, a_exp, current, _ = ff.nonlinear_pm_coil()
    plan = Plan(model)
    o = plain_options(relative_tolerance=1e-8, max_newton_iterations=128)
    out = plan.solve_batch(np.array([[current]]), o, cfg(block_size=bs))
    with ol.block_size(bs):
        ref = ol.newton_solve(model, [current], o)
    r = out["results"][0]
    say(f"  pm_coil: status {r['status']} newton {r['newton_iterations']} (oracle {ref['newton']})"
        f" pcg {r['linear_iterations']} (oracle {ref['linear']}) dA {_rel_l2(out['a'][0], ref['a']):.2e}")
    assert r["status"] == 0 and ref["status"] == 0
    if form == "fp64":
        assert _rel_l2(out["a"][0], ref["a"]) <= 1e-9
        assert np.abs(out["a"][0] - a_exp).max() <= 1e-9
        assert r["newton_iterations"] == ref["newton"]
    else:
        acc.check_driver_form(say, "pm_coil", out["a"][0], out["flux"][0], ref["a"],
                              acc.converged(model, [[current]])[0])
    plan.close()
    model, _ = meshes.motor_sector(angular_cells=160)
    currents = meshes.motor_currents(2, seed=5)
    plan = Plan(model)
    opts = accepted_options()
    out = plan.solve_batch(currents, opts, cfg(block_size=bs))
    with ol.block_size(bs):
        a_ref, flux_ref, res_ref = ol.newton_batch(model, currents, opts, threads=8)
    point = plan.solve_batch(currents, opts, cfg())
    for b in range(2):
        r = out["results"][b]
        say(f"  motor item {b}: status {r['status']} newton {r['newton_iterations']} (oracle "
            f"{res_ref[b]['newton']}) pcg {r['linear_iterations']} (oracle {res_ref[b]['linear']},"
            f" point {point['results'][b]['linear_iterations']}) "
            f"dA {_rel_l2(out['a'][b], a_ref[b]):.2e}")
        assert r["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        if form == "fp64":
            assert _rel_l2(out["a"][b], a_ref[b]) <= 1e-9
            assert np.abs(out["flux"][b] - flux_ref[b]).max() <= 1e-7 * np.abs(flux_ref[b]).max()
            assert r["newton_iterations"] == res_ref[b]["newton"]
        else:
            acc.check_driver_form(say, f"motor item {b}", out["a"][b], out["flux"][b], a_ref[b],
                                  acc.converged(model, currents)[b])
    again = plan.solve_batch(currents, opts, cfg(block_size=bs))
    assert np.array_equal(again["a"], out["a"])
    plan.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_block_partitioned_matches_oracle(form):
    say(f"case block partitioned {form}: motor_small x1, 8 workgroups, bs 4")
    cfg = reference_config if form == "fp64" else mixed_config
    model, _ = meshes.motor_sector(angular_cells=160)
    currents = meshes.motor_currents(1, seed=9)
    plan = Plan(model)
    opts = accepted_options()
    out = plan.solve_batch(currents, opts, cfg(block_size=4, partitions=8))
    with ol.block_size(4):
        a_ref, flux_ref, res_ref = ol.newton_batch(model, currents, opts, threads=1)
    r = out["results"][0]
    say(f"  status {r['status']} newton {r['newton_iterations']} (oracle {res_ref[0]['newton']}) "
        f"dA {_rel_l2(out['a'][0], a_ref[0]):.2e}")
    assert r["status"] == 0 and r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
    if form == "fp64":
        assert _rel_l2(out["a"][0], a_ref[0]) <= 1e-9
        assert np.abs(out["flux"][0] - flux_ref[0]).max() <= 1e-7 * np.abs(flux_ref[0]).max()
    else:
        acc.check_driver_form(say, "motor item 0", out["a"][0], out["flux"][0], a_ref[0],
                              acc.converged(model, currents)[0])
    again = plan.solve_batch(currents, opts, cfg(block_size=4, partitions=8))
    assert np.array_equal(again["a"], out["a"])
    plan.close()


def test_mixed_block_size_other_than_four_is_rejected():
    from femm_gpu_amd.solver import FemmError
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    with pytest.raises(FemmError) as err:
        plan.solve_batch(meshes.motor_currents(1, seed=1), accepted_options(),
                         mixed_config(block_size=8))
    assert err.value.code == 2
    plan.close()
