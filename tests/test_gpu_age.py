"""Native air-gap element (AGE) models on the GPU against the oracle.

The full 360 deg machine at a fractional rotor angle (sliding band turned by 1.3 deg) and the
anti-periodic sector: device assembly (AGE values added after the triangle terms) at the
usual 1e-12 bar, the reference Newton driver in FP64 at 1e-9 (A) / 1e-7 (flux) with the same
Newton counts, bitwise repeats, the chosen mixed form against the oracle's converged solution,
and the AGE postprocess (Maxwell stress force / torque and the Fourier air-gap B of
airgap_postprocess.cpp:41-160) against the oracle's.
"""

import math
import sys

import numpy as np
import pytest

import accuracy as acc
import oracle_lib as ol
from femm_gpu_amd import meshes, wire
from femm_gpu_amd.model import accepted_options
from femm_gpu_amd.solver import Plan, mixed_config, reference_config

pytestmark = pytest.mark.gpu

CURRENTS = np.array([[3e6, -1.5e6, -1.5e6], [-2e6, 4e6, -2e6], [0.5e6, 0.5e6, -1e6]])


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def _model(kind):
    if kind == "full":
        return wire.rotate_air_gaps(meshes.age_machine(96, "age"), 1.3)
    return wire.rotate_air_gaps(meshes.age_machine(96, "age", sector=True), 0.7)


def _angles(kind):
    return (np.arange(16) + 0.5) * (math.pi / 8 if kind == "full" else math.pi / 32)


@pytest.mark.parametrize("kind", ["full", "sector"])
def test_age_assembly_matches_oracle(kind):
    say(f"case age assembly {kind}")
    model = _model(kind)
    plan = Plan(model)
    sym = plan.symbolic()
    rng = np.random.default_rng(3)
    states = rng.normal(0.0, 2e-3, (2, model.node_count))
    fixed = sym["node_to_free"] < 0
    states[:, fixed] = sym["boundary"][fixed]
    code, values, rhs, diag = plan.assemble_batch(states, CURRENTS[:2], True)
    assert code == 0
    for b in range(2):
        st, ref = ol.assemble(model, states[b], CURRENTS[b], True)
        assert st == 0 and np.array_equal(ref["cols"], sym["cols"])
        scale = np.abs(ref["values"]).max()
        dv = np.abs(values[b] - ref["values"]).max() / scale
        ref_rhs = ref["rhs_per_amp"] + ref["rhs_offset"]
        dr = np.abs(rhs[b] - ref_rhs).max() / np.abs(ref_rhs).max()
        say(f"  item {b}: values {dv:.2e} rhs {dr:.2e}")
        assert dv <= 1e-12 and dr <= 1e-12
    plan.close()


@pytest.mark.parametrize("kind", ["full", "sector"])
def test_age_newton_matches_oracle_and_repeats(kind):
    say(f"case age newton {kind}")
    model = _model(kind)
    opts = accepted_options()
    plan = Plan(model)
    out = plan.solve_batch(CURRENTS, opts, reference_config())
    for b in range(len(CURRENTS)):
        ref = ol.newton_solve(model, CURRENTS[b], opts)
        r = out["results"][b]
        da = acc.rel_l2(out["a"][b], ref["a"])
        df = acc.flux_rel(out["flux"][b], ref["flux"])
        say(f"  item {b}: status {r['status']} newton {r['newton_iterations']} (oracle "
            f"{ref['newton']}) dA {da:.2e} dflux {df:.2e}")
        assert r["status"] == 0 and ref["status"] == 0
        assert r["newton_iterations"] == ref["newton"]
        assert da <= 1e-9 and df <= 1e-7
    again = plan.solve_batch(CURRENTS, opts, reference_config())
    assert np.array_equal(again["a"], out["a"]) and np.array_equal(again["flux"], out["flux"])
    plan.close()


@pytest.mark.parametrize("kind", ["full", "sector"])
def test_age_chosen_form_matches_converged_oracle(kind):
    """FP32 PCG with 4-row block-Jacobi and FP64 reliable updates, safeguarded inexact Newton
    with the tight acceptance solve: A within 1e-9 and flux within 1e-7 of the oracle's
    converged solution."""
    say(f"case age chosen form {kind}")
    model = _model(kind)
    opts = accepted_options()
    plan = Plan(model)
    cfg = mixed_config(block_size=4, inexact_forcing=0.01)
    out = plan.solve_batch(CURRENTS, opts, cfg, safeguarded=True)
    conv = acc.converged(model, CURRENTS, threads=4)
    for b in range(len(CURRENTS)):
        assert out["results"][b]["status"] == 0
        acc.check_converged(say, f"item {b}", out["a"][b], out["flux"][b], conv[b])
    plan.close()


@pytest.mark.parametrize("kind", ["full", "sector"])
def test_age_postprocess_matches_oracle(kind):
    say(f"case age postprocess {kind}")
    model = _model(kind)
    angles = _angles(kind)
    opts = accepted_options()
    plan = Plan(model)
    out = plan.solve_batch(CURRENTS, opts, reference_config())
    plan.set_postprocess(angles_rad=angles)
    loads, radial = plan.postprocess_batch(out["a"])
    for b in range(len(CURRENTS)):
        st, ref_loads, ref_radial = ol.age_postprocess(model, out["a"][b], angles)
        assert st == 0
        dl = np.abs(loads[b] - ref_loads).max() / np.abs(ref_loads).max()
        dp = np.abs(radial[b] - ref_radial).max() / np.abs(ref_radial).max()
        say(f"  item {b}: torque {loads[b][2]:.9e} (oracle {ref_loads[2]:.9e}) loads {dl:.2e} "
            f"radial {dp:.2e}")
        assert dl <= 1e-12 and dp <= 1e-12
    again, radial2 = plan.postprocess_batch(out["a"])
    assert np.array_equal(again, loads) and np.array_equal(radial2, radial)
    plan.close()
