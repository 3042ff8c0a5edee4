"""GPU parity of the MI355X path against the CPU oracle and the frozen FEMM answers.

Every call goes through the C ABI (libfemm_gpu_amd.so).  Tolerances:
* assembly: the reference's host/device contract NearMixed(1e-12 abs, 5e-11 rel)
  (protocol_contract_tests.cu:964-1041), CSR pattern identical, 10 repeats bitwise equal;
* Newton solution: A within 1e-9 relative L2 of the oracle, flux within 1e-7 relative,
  undamped step < 1e-6 and equation residual <= 1e-8 per item; repeats bitwise equal;
  forms that follow their own Newton path are held to the oracle's solution converged past
  the gates (tests/accuracy.py);
* frozen FEMM fixtures: the reference's own limits (reference_commands.cu:99-142,
  test_femm_gpu_cli.py:223-292).
"""

import math
import sys

import numpy as np
import pytest

import accuracy as acc
import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.model import accepted_options, plain_options, production_options
from femm_gpu_amd.solver import Plan, mixed_config, reference_config

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def near_mixed(a, b, abs_tol, rel_tol):
    return np.abs(a - b) <= abs_tol + rel_tol * np.maximum(np.abs(a), np.abs(b))


def _assembly_cases():
    yield "three_region", ff.three_region(), [2.0]
    yield "two_circuit", ff.two_circuit(), [2.0, -3.0]
    m, _, cur, _ = ff.nonlinear_pm_coil()
    yield "pm_coil", m, [cur]
    m, cur, _, _ = ff.generic_reference("antiperiodic_strip")
    yield "antiperiodic", m, [cur]
    m, cur, _, _ = ff.generic_reference("axisymmetric_coil")
    yield "axisymmetric", m, [cur]
    m, cur = meshes.motor_sector(angular_cells=160)
    yield "motor_small", m, cur


@pytest.mark.parametrize("name,model,currents", list(_assembly_cases()),
                         ids=[c[0] for c in _assembly_cases()])
@pytest.mark.parametrize("newton", [False, True])
def test_device_assembly_matches_oracle(name, model, currents, newton):
    say(f"case assembly {name} newton={newton}")
    rng = np.random.default_rng(11)
    plan = Plan(model)
    sym = plan.symbolic()
    B = 3
    states = rng.normal(0.0, 2e-3, (B, model.node_count))
    fixed = sym["node_to_free"] < 0
    states[:, fixed] = sym["boundary"][fixed]
    cur = np.tile(np.asarray(currents, np.float64), (B, 1)) * np.array([[1.0], [-0.5], [0.25]])
    code, values, rhs, diag = plan.assemble_batch(states, cur, newton)
    assert code == 0
    for b in range(B):
        st, ref = ol.assemble(model, states[b], cur[b], newton)
        assert st == 0
        assert np.array_equal(ref["rows"], sym["rows"]) and np.array_equal(ref["cols"], sym["cols"])
        assert np.all(near_mixed(values[b], ref["values"], 1e-12, 5e-11)), name
        assert np.all(near_mixed(diag[b], ref["diag"], 1e-12, 5e-11))
        ref_rhs = ref["rhs_per_amp"] + ref["rhs_offset"]
        scale = max(1.0, np.abs(ref_rhs).max())
        assert np.all(near_mixed(rhs[b], ref_rhs, 1e-12 * scale, 5e-11)), name
    for _ in range(10):
        code2, v2, r2, d2 = plan.assemble_batch(states, cur, newton)
        assert code2 == 0 and np.array_equal(v2, values) and np.array_equal(r2, rhs)
    plan.close()


def test_device_pcg_matches_oracle_pcg():
    say("case pcg motor_small")
    model, cur = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    sym = plan.symbolic()
    state = np.zeros(model.node_count)
    code, values, rhs, diag = plan.assemble_batch(state[None], cur[None], False)
    assert code == 0
    x, info = plan.pcg_batch(values, rhs, 1e-10, 16384)
    xr, st, it, res = ol.pcg(sym["rows"], sym["cols"], values[0], diag[0], rhs[0], 1e-10, 16384)
    assert int(info[0, 0]) == 0 and st == 0
    assert abs(int(info[0, 1]) - it) <= max(3, it // 50)
    assert np.linalg.norm(x[0] - xr) <= 1e-7 * np.linalg.norm(xr)
    # verified true residual
    y = np.zeros_like(x[0])
    rows, cols = sym["rows"], sym["cols"]
    for r in range(len(y)):
        y[r] = values[0][rows[r]:rows[r + 1]] @ x[0][cols[rows[r]:rows[r + 1]]]
    assert np.linalg.norm(rhs[0] - y) <= 1.0001e-10 * np.linalg.norm(rhs[0])
    # zero right-hand side returns immediately
    x0, info0 = plan.pcg_batch(values, np.zeros_like(rhs), 1e-10, 100)
    assert info0[0, 0] == 0 and info0[0, 1] == 0 and not x0.any()
    plan.close()


def test_newton_matches_frozen_femm_pm_coil():
    """reference_commands.cu:99-142 through the GPU path."""
    say("case newton pm_coil frozen FEMM")
    model, a_exp, current, flux = ff.nonlinear_pm_coil()
    plan = Plan(model)
    out = plan.solve_batch(np.array([[current]]), plain_options(relative_tolerance=1e-8,
                                                                 max_newton_iterations=128),
                           reference_config(), fields=True)
    r = out["results"][0]
    assert r["status"] == 0
    assert np.abs(out["a"][0] - a_exp).max() <= 1e-9
    _, bx, by, _ = ol.finalize(model, a_exp)
    assert max(np.abs(out["bx"][0] - bx).max(), np.abs(out["by"][0] - by).max()) <= 1e-6
    assert abs(out["flux"][0, 0] - flux) <= 1e-10 + 1e-6 * abs(flux)
    assert r["equation_residual"] <= 1e-8
    plan.close()


def test_batched_safeguarded_matches_oracle():
    """femm_newton_batch_safeguarded (SolveSafeguarded, nonlinear_solver.cu:534-659) from the
    plain start: per item the oracle's Newton count and state (A within 1e-9, flux 1e-7) with
    exact linear solves; with inexact forcing 0.01 the oracle's state under the same policy
    (A within 1e-9; counts within 2 Newton steps / 20 % PCG, the tolerance following the
    computed residual) and the frozen FEMM pm_coil answer at its bars."""
    say("case batched safeguarded pm_coil: 3 currents, exact and forcing 0.01")
    model, a_exp, current, flux = ff.nonlinear_pm_coil()
    plan = Plan(model)
    o = production_options(require_equation_residual=1)
    cur = np.array([[current], [0.5 * current], [1.5 * current]])
    exact = plan.solve_batch(cur, o, reference_config(), safeguarded=True)
    inexact = plan.solve_batch(cur, o, reference_config(inexact_forcing=0.01), safeguarded=True)
    for b in range(3):
        for out, forcing in ((exact, 0.0), (inexact, 0.01)):
            ref = ol.safeguarded_solve(model, cur[b], o, forcing=forcing)
            r = out["results"][b]
            say(f"  item {b} forcing {forcing}: newton {r['newton_iterations']} (oracle "
                f"{ref['newton']}) pcg {r['linear_iterations']} (oracle {ref['linear']})")
            assert r["status"] == ref["status"] == 0
            assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
            assert np.linalg.norm(out["a"][b] - ref["a"]) <= 1e-9 * np.linalg.norm(ref["a"])
            if forcing == 0.0:
                assert r["newton_iterations"] == ref["newton"]
            else:
                assert abs(int(r["newton_iterations"]) - ref["newton"]) <= 2
                assert abs(int(r["linear_iterations"]) - ref["linear"]) <= 0.2 * ref["linear"]
    for out in (exact, inexact):
        assert np.abs(out["a"][0] - a_exp).max() <= 1e-9
        assert abs(out["flux"][0, 0] - flux) <= 1e-10 + 1e-6 * abs(flux)
    plan.close()


@pytest.mark.parametrize("name,problem,pot_l2,flux_rel", [
    ("periodic_strip", 0, 1e-10, 1e-10),
    ("antiperiodic_strip", 0, 1e-10, 1e-10),
    ("axisymmetric_coil", 1, 3e-3, 1e-3),
])
def test_newton_matches_frozen_femm_generic(name, problem, pot_l2, flux_rel):
    say(f"case newton {name}")
    model, current, expected, flux = ff.generic_reference(name)
    plan = Plan(model)
    out = plan.solve_batch(np.array([[current]]), production_options())
    assert out["results"][0]["status"] == 0
    pot = out["a"][0].copy()
    if problem == 1:
        pot *= 2.0 * math.pi * model.node_xy[:, 0]
    assert math.sqrt(np.sum((pot - expected) ** 2) / np.sum(expected ** 2)) <= pot_l2
    assert abs(out["flux"][0, 0] - flux) / abs(flux) <= flux_rel
    for a, b, s in model.constraints:
        assert abs(out["a"][0][b] - s * out["a"][0][a]) <= 1e-15
    plan.close()


def test_newton_analytic_unit_square_and_superposition():
    say("case newton analytic")
    plan = Plan(ff.unit_square_nonlinear())
    out = plan.solve_batch(np.array([[12.0], [-12.0]]), plain_options(relative_tolerance=1e-8),
                           fields=True)
    assert list(out["results"]["status"]) == [0, 0]
    np.testing.assert_allclose(out["a"][0], [0, 0, 0, 0, 1.0], atol=1e-12)
    np.testing.assert_allclose(out["a"][1], -out["a"][0], atol=1e-12)
    np.testing.assert_allclose(out["bx"][0], [2, 0, -2, 0], atol=1e-12)
    assert out["flux"][0, 0] == pytest.approx(1.0 / 3.0, abs=1e-12)
    plan.close()
    plan = Plan(ff.two_circuit())
    out = plan.solve_batch(np.array([[2.0, 0.0], [0.0, -3.0], [2.0, -3.0]]),
                           plain_options(relative_tolerance=1e-8))
    np.testing.assert_allclose(out["a"][:, 4], [0.5, -0.375, 0.125], atol=1e-12)
    np.testing.assert_allclose(out["flux"][2], [0.125, 0.0625], atol=1e-12)
    plan.close()


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_newton_motor_matches_oracle_and_is_deterministic():
    """Anti-periodic 12-slot motor sector (19.6k nodes), 4 current samples incl. zero."""
    say("case newton motor_small vs oracle")
    model, _ = meshes.motor_sector(angular_cells=160)
    currents = np.vstack([meshes.motor_currents(3, seed=5), np.zeros((1, 3))])
    plan = Plan(model)
    opts = accepted_options()
    out = plan.solve_batch(currents, opts, reference_config())
    a_ref, flux_ref, res_ref = ol.newton_batch(model, currents, opts, threads=16)
    for b in range(len(currents)):
        r = out["results"][b]
        say(f"  item {b}: status {r['status']} newton {r['newton_iterations']} "
            f"pcg {r['linear_iterations']} step {r['relative_step']:.3e} "
            f"eq {r['equation_residual']:.3e} | oracle newton {res_ref[b]['newton']} "
            f"pcg {res_ref[b]['linear']} | dA {_rel_l2(out['a'][b], a_ref[b]):.3e} "
            f"dflux {np.abs(out['flux'][b] - flux_ref[b]).max() / np.abs(flux_ref[b]).max():.3e}")
        assert r["status"] == 0 and res_ref[b]["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        assert _rel_l2(out["a"][b], a_ref[b]) <= 1e-9
        assert np.abs(out["flux"][b] - flux_ref[b]).max() <= 1e-7 * np.abs(flux_ref[b]).max()
    again = plan.solve_batch(currents, opts, reference_config())
    assert np.array_equal(again["a"], out["a"]) and np.array_equal(again["flux"], out["flux"])
    # batch composition does not change an item's result
    single = plan.solve_batch(currents[2:3], opts, reference_config())
    assert np.array_equal(single["a"][0], out["a"][2])
    plan.close()


def test_newton_invalid_inputs_are_isolated_per_item():
    say("case newton invalid inputs")
    plan = Plan(ff.three_region())
    cur = np.array([[2.0], [np.nan], [-2.0]])
    out = plan.solve_batch(cur, production_options())
    st = list(out["results"]["status"])
    assert st[1] == 2 and st[0] == 0 and st[2] == 0
    limited = plan.solve_batch(np.array([[2.0]]), plain_options(max_newton_iterations=0))
    assert limited["results"][0]["status"] == 14
    plan.close()


def test_postprocess_matches_frozen_femm_and_oracle():
    """Weighted-stress loads and smoothed air-gap B on the GPU: frozen FEMM limits
    (reference_commands.cu:144-208) and the oracle's mask and loads."""
    say("case postprocess pm_coil")
    model, _, current, _ = ff.nonlinear_pm_coil()
    post = ff.nonlinear_pm_coil_postprocess()
    plan = Plan(model)
    opts = plain_options(relative_tolerance=1e-8, max_newton_iterations=128)
    out = plan.solve_batch(np.array([[current], [0.5 * current]]), opts, reference_config(),
                           fields=True)
    plan.set_postprocess(selected_label=post["selected_label"], air_label=3,
                         radius_m=post["radius_m"], angles_rad=post["angles"])
    loads, radial = plan.postprocess_batch(out["a"])
    f = loads[0]
    assert near_mixed(f[0], post["fx"], 1e-8, 1e-5) and near_mixed(f[1], post["fy"], 1e-8, 1e-5)
    assert near_mixed(f[2], post["torque"], 1e-10, 1e-5)
    assert np.all(near_mixed(radial[0], post["radial_b"], 1e-9, 1e-5))
    for b in range(2):
        st, ref_loads, ref_br, ref_mask = ol.weighted_stress(
            model, out["a"][b], out["bx"][b], out["by"][b], post["selected_label"], 3, -1, -1,
            post["radius_m"], post["angles"])
        assert st == 0
        assert np.array_equal(plan.postprocess_mask(), ref_mask)
        assert np.all(near_mixed(loads[b], ref_loads, 1e-12, 1e-10))
        assert np.all(near_mixed(radial[b], ref_br, 1e-14, 1e-12))
    again = plan.postprocess_batch(out["a"])
    assert np.array_equal(again[0], loads) and np.array_equal(again[1], radial)
    plan.close()


def test_postprocess_uniform_field_cancels():
    """physics_regression.cu:88-149: A = x, uniform B = (0, -1): zero loads, Br = 0 / -1."""
    say("case postprocess uniform field")
    model = ff.postprocess_ring()
    plan = Plan(model)
    plan.set_postprocess(selected_label=1, air_label=3, radius_m=0.75,
                         angles_rad=[0.0, 0.5 * math.pi])
    loads, radial = plan.postprocess_batch(model.node_xy[:, 0][None].copy())
    np.testing.assert_allclose(loads[0], 0.0, atol=1e-8)
    np.testing.assert_allclose(radial[0], [0.0, -1.0], atol=1e-12)
    plan.close()


def test_mixed_precision_form_matches_oracle():
    """FP32 PCG with FP64 reliable updates from the current state, same acceptance gates.
    The reference driver in this precision follows its own Newton path (its relaxation
    decisions see different rounding), so it is held to the oracle's converged solution at
    the oracle driver's own distance from it + 1e-9.  The safeguarded policy in the same
    precision (the chosen form's solver, inexact forcing 0.01, block-Jacobi 4) ends in an
    exact acceptance and is held to 1e-9 of the converged solution, flux 1e-7."""
    say("case newton motor_small mixed precision: driver and safeguarded forms")
    model, _ = meshes.motor_sector(angular_cells=160)
    currents = np.vstack([meshes.motor_currents(3, seed=5), np.zeros((1, 3))])
    plan = Plan(model)
    opts = accepted_options()
    out = plan.solve_batch(currents, opts, mixed_config())
    sg = plan.solve_batch(currents, opts, mixed_config(block_size=4, inexact_forcing=0.01),
                          safeguarded=True)
    a_ref, flux_ref, res_ref = ol.newton_batch(model, currents, opts, threads=16)
    conv = acc.converged(model, currents)
    for b in range(len(currents)):
        r, q = out["results"][b], sg["results"][b]
        say(f"  item {b}: status {r['status']} newton {r['newton_iterations']} "
            f"(oracle {res_ref[b]['newton']}) pcg {r['linear_iterations']} "
            f"(oracle {res_ref[b]['linear']}) replacements {r['refinements']} "
            f"eq {r['equation_residual']:.3e}; safeguarded newton {q['newton_iterations']} "
            f"pcg {q['linear_iterations']} eq {q['equation_residual']:.3e}")
        assert r["status"] == 0 and q["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        assert q["relative_step"] < 1e-6 and q["equation_residual"] <= 1e-8
        acc.check_driver_form(say, f"driver item {b}", out["a"][b], out["flux"][b], a_ref[b],
                              conv[b])
        acc.check_converged(say, f"safeguarded item {b}", sg["a"][b], sg["flux"][b], conv[b])
    again = plan.solve_batch(currents, opts, mixed_config())
    assert np.array_equal(again["a"], out["a"])
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_partitioned_solve_matches_oracle(precision):
    """Several workgroups per item (rows, nodes and triangles split, team reductions in
    workgroup order): same acceptance as the one-workgroup path.  The reference form: A within
    1e-9 relative L2 of the FP64 oracle, the oracle's Newton counts.  The mixed form: the
    driver at the oracle driver's own distance from the converged solution + 1e-9, and the
    safeguarded policy (inexact forcing 0.01, block-Jacobi 4) within 1e-9 of it.  Repeatable
    bitwise for a fixed partition count."""
    say(f"case partitioned solve ({precision}): pm_coil x 3 partitions, motor_small 2 x 16")
    cfg = reference_config if precision == "fp64" else mixed_config
    model, a_exp, current, _ = ff.nonlinear_pm_coil()
    plan = Plan(model)
    o = plain_options(relative_tolerance=1e-8, max_newton_iterations=128)
    out = plan.solve_batch(np.array([[current]]), o, cfg(partitions=3), fields=True)
    ref = ol.newton_solve(model, [current], o)
    r = out["results"][0]
    say(f"  pm_coil: status {r['status']} newton {r['newton_iterations']} (oracle {ref['newton']}) "
        f"dA {_rel_l2(out['a'][0], ref['a']):.2e}")
    assert r["status"] == 0
    if precision == "fp64":
        assert _rel_l2(out["a"][0], ref["a"]) <= 1e-9
        assert np.abs(out["a"][0] - a_exp).max() <= 1e-9
    else:
        conv = acc.converged(model, [[current]])[0]
        acc.check_driver_form(say, "pm_coil driver", out["a"][0], out["flux"][0], ref["a"], conv)
    plan.close()
    model, _ = meshes.motor_sector(angular_cells=160)
    currents = meshes.motor_currents(2, seed=5)
    plan = Plan(model)
    opts = accepted_options()
    out = plan.solve_batch(currents, opts, cfg(partitions=16))
    a_ref, flux_ref, res_ref = ol.newton_batch(model, currents, opts, threads=16)
    conv = acc.converged(model, currents)
    for b in range(2):
        r = out["results"][b]
        say(f"  motor item {b}: status {r['status']} newton {r['newton_iterations']} "
            f"(oracle {res_ref[b]['newton']}) pcg {r['linear_iterations']} "
            f"(oracle {res_ref[b]['linear']}) dA {_rel_l2(out['a'][b], a_ref[b]):.2e}")
        assert r["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        if precision == "fp64":
            assert _rel_l2(out["a"][b], a_ref[b]) <= 1e-9
            assert np.abs(out["flux"][b] - flux_ref[b]).max() <= 1e-7 * np.abs(flux_ref[b]).max()
            assert r["newton_iterations"] == res_ref[b]["newton"]
        else:
            acc.check_driver_form(say, f"driver item {b}", out["a"][b], out["flux"][b],
                                  a_ref[b], conv[b])
    again = plan.solve_batch(currents, opts, cfg(partitions=16))
    assert np.array_equal(again["a"], out["a"])
    auto = plan.solve_batch(currents[:1], opts, cfg(partitions=0))
    assert auto["results"][0]["status"] == 0
    if precision == "fp64":
        assert _rel_l2(auto["a"][0], a_ref[0]) <= 1e-9
    else:
        acc.check_driver_form(say, "autotuned driver item 0", auto["a"][0], auto["flux"][0],
                              a_ref[0], conv[0])
        sg = plan.solve_batch(currents, opts, mixed_config(partitions=16, block_size=4,
                                                           inexact_forcing=0.01),
                              safeguarded=True)
        for b in range(2):
            assert sg["results"][b]["status"] == 0
            assert sg["results"][b]["equation_residual"] <= 1e-8
            acc.check_converged(say, f"safeguarded item {b}", sg["a"][b], sg["flux"][b], conv[b])
    plan.close()


def test_c_core_single_rhs_partitioned_matches_oracle():
    """the 62k-node C-core (2 mm gap, NdFeB block, two coils, B-H iron),
    one right-hand side, solved by the autotuned partitioned FP64 path: A within 1e-9
    relative L2 of the oracle, flux within 1e-7, same Newton count, both gates."""
    model, currents = meshes.c_core()
    say(f"case c_core single RHS: N={model.node_count}")
    plan = Plan(model)
    opts = accepted_options()
    parts = plan.autotune(1, reference_config(partitions=0))
    out = plan.solve_batch(currents[None, :], opts, reference_config(partitions=0))
    ref = ol.newton_solve(model, currents, opts)
    r = out["results"][0]
    say(f"  {parts} workgroups: status {r['status']} newton {r['newton_iterations']} "
        f"(oracle {ref['newton']}) pcg {r['linear_iterations']} (oracle {ref['linear']}) "
        f"dA {_rel_l2(out['a'][0], ref['a']):.2e}")
    assert r["status"] == ref["status"] == 0
    assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
    assert r["newton_iterations"] == ref["newton"]
    assert _rel_l2(out["a"][0], ref["a"]) <= 1e-9
    assert np.abs(out["flux"][0] - ref["flux"]).max() <= 1e-7 * np.abs(ref["flux"]).max()
    plan.close()
