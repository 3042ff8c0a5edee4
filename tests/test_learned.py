"""CPU checks for the learned warm start and the safeguarded Newton oracle.

* The MLP restatement (oracle/mlp_oracle.py) is pinned by central finite differences of
  its exact FP64 mode and by known BF16 round-to-nearest-even answers.
* The safeguarded Newton oracle (SolveSafeguarded, nonlinear_solver.cu:534-659) is pinned
  against the frozen FEMM pm_coil solution (reference_commands.cu:99-142 limits), the
  unit-square closed form and a linear problem, where one full Newton step is exact.
* Host-side entry points of the library report GPU_UNAVAILABLE without a GPU.
"""

import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd.model import plain_options

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402

# sha256 prefix of the default motor_sector mesh (nodes, triangles, labels).
DEFAULT_MOTOR_DIGEST = "872cb79793dbea9a"


def test_bf16_round_to_nearest_even_known_answers():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)          # BF16 spacing just above 1.0
    cases = [
        (one, one),
        (one + ulp, one + ulp),
        (one + ulp / 2, one),                # tie -> even (1.0)
        (one + 3 * ulp / 2, one + 2 * ulp),  # tie -> even (1 + 2 ulp)
        (one + ulp / 2 + np.float32(2.0 ** -20), one + ulp),  # above the tie
        (np.float32(-2.5), np.float32(-2.5)),
        (np.float32(3.0e38), np.float32(226 * 2.0 ** 120)),   # spacing 2^120 near 2^127
    ]
    for x, expected in cases:
        got = mo.bf16(np.array([x], np.float32))[0]
        assert got == pytest.approx(expected, rel=2e-3), (x, got)
        assert (got.view(np.uint32) & 0xFFFF) == 0
    assert mo.bf16(np.array([one + ulp / 2], np.float32))[0] == one
    assert mo.bf16(np.array([one + 3 * ulp / 2], np.float32))[0] == one + 2 * ulp


def _random_network(layout, seed):
    rng = np.random.default_rng(seed)
    flat = np.zeros(layout.count)
    W1, b1, W2, b2, W3, b3 = layout.split(flat)
    W1[: layout.inputs] = rng.normal(0, 0.5, (layout.inputs, layout.hidden))
    b1[:] = rng.normal(0, 0.1, layout.hidden)
    W2[:] = rng.normal(0, 1.0 / np.sqrt(layout.hidden), W2.shape)
    b2[:] = rng.normal(0, 0.1, layout.hidden)
    W3[:, : layout.outputs] = rng.normal(0, 1.0 / np.sqrt(layout.hidden),
                                         (layout.hidden, layout.outputs))
    b3[: layout.outputs] = rng.normal(0, 0.1, layout.outputs)
    return flat


def test_mlp_oracle_gradient_matches_finite_differences():
    layout = mo.Layout(5, 32, 7)
    flat = _random_network(layout, 3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 7))
    mean, std = x.mean(0), x.std(0)
    loss, grad = mo.loss_and_gradients(flat, x, t, mean, std, layout, quantized=False)
    live = np.flatnonzero(grad)
    assert len(live) > 0.5 * (layout.inputs * 32 + 32 * 32 + 32 * 7)
    picks = rng.choice(live, 60, replace=False)
    h = 1e-6
    for k in picks:
        e = np.zeros_like(flat)
        e[k] = h
        lp, _ = mo.loss_and_gradients(flat + e, x, t, mean, std, layout, quantized=False)
        lm, _ = mo.loss_and_gradients(flat - e, x, t, mean, std, layout, quantized=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, grad[k])
    # BF16 rounding points perturb, but do not change, the gradient
    _, gq = mo.loss_and_gradients(flat, x, t, mean, std, layout, quantized=True)
    assert np.linalg.norm(gq - grad) <= 5e-2 * np.linalg.norm(grad)


def test_mlp_oracle_forward_is_the_loss_model():
    layout = mo.Layout(3, 16, 4)
    flat = _random_network(layout, 8)
    x = np.random.default_rng(9).normal(size=(5, 3))
    y = mo.forward(flat, x, np.zeros(3), np.ones(3), layout, quantized=False)
    loss, _ = mo.loss_and_gradients(flat, x, y, np.zeros(3), np.ones(3), layout, quantized=False)
    assert loss == 0.0


def test_safeguarded_oracle_matches_frozen_femm_pm_coil():
    """Full Newton + Armijo reaches the frozen FEMM state within the reference's limits."""
    model, a_exp, current, flux = ff.nonlinear_pm_coil()
    o = plain_options(relative_tolerance=1e-8, max_newton_iterations=128)
    r = ol.safeguarded_solve(model, [current], o)
    assert r["status"] == 0 and r["newton"] >= 2
    assert np.abs(r["a"] - a_exp).max() <= 1e-9
    assert r["eq"] <= 1e-8 and r["step"] < 1e-6
    # from its own answer: accepted at the first iteration
    w = ol.safeguarded_solve(model, [current], o, warm_start=r["a"])
    assert w["status"] == 0 and w["newton"] == 1
    np.testing.assert_array_equal(w["a"], r["a"])


def test_safeguarded_oracle_linear_problems_are_exact():
    """Unit square (self_test.cu:163, flux 1/3) and the two-circuit superposition: one
    full step is exact, the second iteration only confirms it."""
    o = plain_options(relative_tolerance=1e-8)
    r = ol.safeguarded_solve(ff.unit_square_nonlinear(), [12.0], o)
    assert r["status"] == 0 and r["newton"] == 2
    np.testing.assert_allclose(r["a"], [0, 0, 0, 0, 1.0], atol=1e-12)
    r = ol.safeguarded_solve(ff.two_circuit(), [2.0, -3.0], o)
    assert r["status"] == 0 and r["newton"] == 2
    assert r["a"][4] == pytest.approx(0.125, abs=1e-12)


def test_learned_entry_points_report_gpu_unavailable_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from femm_gpu_amd.solver import library
    lib = library()
    h = ctypes.c_void_p()
    assert lib.femm_mlp_create(3, 32, 8, ctypes.c_uint64(1), ctypes.byref(h)) == 6
    assert not h.value


def test_motor_family_variants_share_the_mesh():
    from femm_gpu_amd import meshes
    fam = meshes.study_family(angular_cells=160)
    x = meshes.study_samples(3, seed=1)
    models = [meshes.study_model(fam, s) for s in x]
    base = models[0]
    for m in models[1:]:
        assert np.shares_memory(m.node_xy, base.node_xy)
        assert np.shares_memory(m.triangles, base.triangles)
        assert len(m.materials) == len(base.materials) == 9 and m.circuit_count == 3
        assert np.array_equal(m.dirichlet_nodes, base.dirichlet_nodes)
        assert np.array_equal(m.constraints, base.constraints)
    # turning the rotor moves only magnet-band labels and keeps each magnet's area
    area = base.element_areas()
    l0, l6 = fam.labels(0), fam.labels(6)
    moved = l0 != l6
    assert moved.any()
    assert set(np.unique(l0[moved])) | set(np.unique(l6[moved])) <= {fam.AIR, fam.MAG1, fam.MAG2}
    for mag in (fam.MAG1, fam.MAG2):
        assert area[l6 == mag].sum() == pytest.approx(area[l0 == mag].sum(), rel=1e-12)
    # fill factors scale the coil current densities; coercivities and B-H scale pass through
    s = x[1]
    m = models[1]
    assert [mt.source_j_per_a for mt in m.materials[5:]] == pytest.approx(
        [s[6], -s[7], s[8], -s[9]])
    assert (m.materials[3].hc, m.materials[4].hc) == (s[3], s[4])
    b, _ = m.curves[0]
    b0, _ = meshes.tanh_bh_curve()
    assert b[-1] == pytest.approx(b0[-1] + (s[5] - 1.0) * 1.9 * np.tanh(2e5 / 300.0), rel=1e-12)


def test_default_motor_mesh_is_pinned():
    """The benchmark mesh (58,682 nodes) must not drift: nodes, triangles and labels."""
    import hashlib
    from femm_gpu_amd import meshes
    m, _ = meshes.motor_sector()
    h = hashlib.sha256()
    for arr in (m.node_xy, m.triangles, m.triangle_material):
        h.update(np.ascontiguousarray(arr).tobytes())
    assert (m.node_count, m.triangle_count) == (58682, 116160)
    assert h.hexdigest()[:16] == DEFAULT_MOTOR_DIGEST
