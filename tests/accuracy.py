"""Accuracy bars of the GPU forms against the oracle (test infrastructure).

Two references, both from the CPU oracle:
* the oracle run with the same algorithm (the reference driver, FP64, x0 = 0): a form that
  follows that path (the plain FP64 HIP form) must agree with it to 1e-9 relative L2;
* the oracle's solution converged past the gates (``oracle_lib.converged_solve``): a form
  that follows its own Newton path but ends in the safeguarded policy's exact acceptance
  (learned start, inexact forcing, FP32 PCG with FP64 refinement, partitioned teams) must
  lie within 1e-9 relative L2 of it, flux linkage and torque within 1e-7.  The reference
  driver's own accepted state is only fixed to the gates' resolution (its relaxed steps
  stop at equation residuals of a few 1e-9), so a form running that driver in another
  precision is held to the driver's own distance from the converged solution plus 1e-9.
"""

import numpy as np

import oracle_lib as ol

A_BAR = 1e-9
FLUX_BAR = 1e-7
TORQUE_BAR = 1e-7

_cache = {}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def flux_rel(f, g):
    return float(np.abs(f - g).max() / max(np.abs(g).max(), 1e-300))


def converged(model, currents, threads=16):
    """converged_solve per row of currents (cached per model and currents)."""
    currents = np.ascontiguousarray(currents, np.float64).reshape(-1, model.circuit_count)
    key = (model.name, model.node_count, currents.tobytes())
    if key not in _cache:
        rows = ol.converged_batch(model, currents, threads)
        for r in rows:
            assert r["status"] == 0 and r["eq"] <= 1e-11, r["eq"]
        _cache[key] = rows
    return _cache[key]


def check_converged(say, label, a, flux, conv):
    """A within A_BAR and flux within FLUX_BAR of one converged oracle item."""
    da, df = rel_l2(a, conv["a"]), flux_rel(flux, conv["flux"])
    say(f"  {label}: A vs converged {da:.2e}, flux {df:.2e}")
    assert da <= A_BAR, (label, da)
    assert df <= FLUX_BAR, (label, df)
    return da


def check_driver_form(say, label, a, flux, driver_a, conv):
    """A form running the reference driver in another precision: at least as close to the
    converged solution as the oracle's own driver state, to A_BAR; flux within FLUX_BAR."""
    da, own = rel_l2(a, conv["a"]), rel_l2(driver_a, conv["a"])
    df = flux_rel(flux, conv["flux"])
    say(f"  {label}: A vs converged {da:.2e} (oracle driver {own:.2e}), flux {df:.2e}")
    assert da <= own + A_BAR, (label, da, own)
    assert df <= FLUX_BAR, (label, df)
    return da


def torque_of(model, a, post):
    """Oracle weighted-stress torque of state a with the options of meshes.motor_postprocess."""
    st, bx, by, _ = ol.finalize(model, a)
    assert st == 0
    st, loads, _, _ = ol.weighted_stress(model, a, bx, by, selected_label=-1, air_label=-1,
                                         selected_group=post["selected_group"],
                                         air_group=post["air_group"], radius_m=post["radius_m"],
                                         angles_rad=post["angles_rad"])
    assert st == 0
    return loads[2]
