"""GPU checks of the safeguarded Newton kernel and the learned warm start (C ABI).

* MLP forward / gradients: the device's BF16 MFMA kernels against the CPU restatement
  with the same BF16 rounding points (oracle/mlp_oracle.py, itself pinned by finite
  differences); FP32 accumulation order differs, so the bar is 2e-3 relative L2 per
  parameter block (a BF16 rounding of an intermediate may flip by one unit).
* Safeguarded Newton (nonlinear_solver.cu:534-659) and the recovery path: A within 1e-9
  relative L2 of the oracle, same iteration counts.
* Learned start: every item passes the 1e-6 step and 1e-8 equation-residual gates; A lies
  within 1e-9 relative L2 of the oracle's solution converged past the gates, flux linkage
  and weighted-stress torque within 1e-7 (tests/accuracy.py); the mean Newton iteration
  count drops; a zero warm budget solves every item from the plain start by the same
  safeguarded policy.
"""

import sys
from pathlib import Path

import numpy as np
import pytest

import accuracy as acc
import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.learned import Mlp, WarmStart, train_settings
from femm_gpu_amd.model import accepted_options, plain_options
from femm_gpu_amd.solver import Plan, mixed_config, reference_config

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _network(inputs, hidden, outputs, seed):
    net = Mlp(inputs, hidden, outputs, seed=seed)
    layout = mo.Layout(inputs, hidden, outputs)
    assert (net.count, net.in_p, net.out_p) == (layout.count, layout.in_p, layout.out_p)
    assert (net.off_b1, net.off_w2, net.off_w3, net.off_b3) == (layout.b1, layout.w2, layout.w3,
                                                                layout.b3)
    return net, layout


def test_mlp_forward_matches_cpu_restatement():
    say("case mlp forward 11->256->256->64, 37 rows")
    net, layout = _network(11, 256, 64, seed=5)
    rng = np.random.default_rng(1)
    x = rng.normal(3.0, 2.0, (37, 11))
    mean, std = x.mean(0), x.std(0)
    net.set_standardization(mean, std)
    flat = net.parameters().astype(np.float64)
    y = net.forward(x)
    ref = mo.forward(flat, x, mean, std, layout, quantized=True)
    say(f"  rel L2 {_rel_l2(y, ref):.2e}")
    assert _rel_l2(y, ref) <= 2e-3
    assert np.array_equal(net.forward(x), y)


def test_mlp_gradients_match_cpu_restatement():
    say("case mlp gradients 11->256->256->64, 53 rows")
    net, layout = _network(11, 256, 64, seed=7)
    rng = np.random.default_rng(2)
    flat = net.parameters().astype(np.float64)
    W1, b1, W2, b2, W3, b3 = layout.split(flat)
    b1[:] = rng.normal(0, 0.2, b1.shape)        # non-zero biases exercise the bias paths
    b2[:] = rng.normal(0, 0.2, b2.shape)
    b3[: layout.outputs] = rng.normal(0, 0.2, layout.outputs)
    net.set_parameters(flat.astype(np.float32))
    flat = net.parameters().astype(np.float64)
    x = rng.normal(size=(53, 11))
    t = rng.normal(size=(53, 64)).astype(np.float32)
    mean, std = np.zeros(11), np.ones(11)
    loss, grad = net.gradients(x, t)
    ref_loss, ref_grad = mo.loss_and_gradients(flat, x, t.astype(np.float64), mean, std, layout,
                                               quantized=True)
    say(f"  loss {loss:.6e} vs {ref_loss:.6e}")
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    for name, g, r in zip(names, layout.split(grad.astype(np.float64)), layout.split(ref_grad)):
        err = _rel_l2(g, r)
        say(f"  {name}: rel L2 {err:.2e}")
        assert err <= 2e-3, name
    # padded weights never receive gradient
    gW1 = layout.split(grad)[0]
    assert not gW1[11:].any()
    _, again = net.gradients(x, t)
    assert np.array_equal(again, grad)


def test_mlp_training_fits_a_smooth_map():
    say("case mlp training 3->64->64->8")
    net, _ = _network(3, 64, 8, seed=3)
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (512, 3))
    t = np.stack([np.sin(x[:, 0] + k * x[:, 1]) * np.cos(x[:, 2]) for k in range(8)], 1)
    net.set_standardization(np.zeros(3), np.ones(3))
    hist = net.train(x, t, steps=1500, batch=64, learning_rate=3e-3, seed=1)
    say(f"  loss {hist[0]:.3e} -> {hist[-100:].mean():.3e}")
    assert np.all(np.isfinite(hist))
    assert hist[-100:].mean() <= 0.02 * hist[:10].mean()
    y = net.forward(x[:64])
    assert np.mean((y - t[:64]) ** 2) <= 5e-3


def test_safeguarded_kernel_matches_oracle():
    """Recovery path (SolveSafeguarded after an exhausted reference driver) against the
    oracle: an exhausted recovery on the PM/coil model (cap 3) and the small motor with a
    cap of 16, where item 0 needs the recovery and item 1 does not."""
    say("case safeguarded pm_coil + motor_small vs oracle")
    model, _, current, _ = ff.nonlinear_pm_coil()
    o = plain_options(relative_tolerance=1e-8, max_newton_iterations=3,
                      recover_nonconvergence=1)
    ref = ol.newton_solve(model, [current], o)
    plan = Plan(model)
    out = plan.solve_batch(np.array([[current]]), o, reference_config())
    r = out["results"][0]
    say(f"  pm_coil: status {r['status']} newton {r['newton_iterations']} "
        f"(oracle {ref['status']} / {ref['newton']}) pcg {r['linear_iterations']} "
        f"(oracle {ref['linear']})")
    assert r["status"] == ref["status"] == 14
    assert r["newton_iterations"] == ref["newton"] == 6
    assert r["linear_iterations"] == ref["linear"]
    assert _rel_l2(out["a"][0], ref["a"]) <= 1e-9
    plan.close()

    model, _ = meshes.motor_sector(angular_cells=160)
    currents = meshes.motor_currents(2, seed=9)
    o = accepted_options(max_newton_iterations=16, recover_nonconvergence=1)
    plan = Plan(model)
    out = plan.solve_batch(currents, o, reference_config())
    for b in range(2):
        ref = ol.newton_solve(model, currents[b], o)
        r = out["results"][b]
        say(f"  motor item {b}: status {r['status']} newton {r['newton_iterations']} "
            f"(oracle {ref['newton']}) dA {_rel_l2(out['a'][b], ref['a']):.2e}")
        assert r["status"] == ref["status"] == 0
        assert r["newton_iterations"] == ref["newton"]
        assert _rel_l2(out["a"][b], ref["a"]) <= 1e-9
        assert r["equation_residual"] <= 1e-8 and r["relative_step"] < 1e-6
    assert out["results"][0]["newton_iterations"] > 16     # recovered
    assert out["results"][1]["newton_iterations"] <= 16
    plan.close()


def _small_motor_data(plan, count, seed):
    currents = meshes.motor_currents(count, seed=seed)
    out = plan.solve_batch(currents, accepted_options(), mixed_config())
    assert np.all(out["results"]["status"] == 0)
    return currents, out


def test_learned_warm_start_gains_and_matches_plain():
    say("case learned warm start motor_small: 160 training solves, 16 held out")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 160, seed=777)
    ws = WarmStart(plan, inputs=3, modes=32, hidden=256, seed=2)
    rep = ws.fit(train_cur, train["a"], train_settings(steps=3000, batch=64, learning_rate=2e-3))
    say(f"  fit: {rep}")
    assert rep["pod_energy_fraction"] > 0.999
    assert rep["final_loss"] < rep["first_loss"]
    test_cur = meshes.motor_currents(16, seed=20251203)
    plain = plan.solve_batch(test_cur, accepted_options(), reference_config())
    a0 = ws.predict(test_cur)
    err = np.array([_rel_l2(a0[b], plain["a"][b]) for b in range(16)])
    say(f"  prediction rel error: mean {err.mean():.2e} max {err.max():.2e}")
    assert err.max() < 1e-2
    learned = ws.solve_batch(test_cur, test_cur, accepted_options(), mixed_config(),
                             warm_iterations=12)
    res, pres = learned["results"], plain["results"]
    conv = acc.converged(model, test_cur)
    post = meshes.motor_postprocess()
    plan.set_postprocess(**post)
    loads, _ = plan.postprocess_batch(learned["a"])
    for b in range(16):
        assert res[b]["status"] == 0
        assert res[b]["relative_step"] < 1e-6 and res[b]["equation_residual"] <= 1e-8
        acc.check_converged(say, f"item {b}", learned["a"][b], learned["flux"][b], conv[b])
        if b < 4:
            torque = acc.torque_of(model, conv[b]["a"], post)
            dt = abs(loads[b, 2] - torque) / abs(torque)
            say(f"  item {b}: torque {loads[b, 2]:.9e} vs converged {torque:.9e} ({dt:.2e})")
            assert dt <= acc.TORQUE_BAR
    say(f"  newton mean: learned {res['newton_iterations'].mean():.2f} plain "
        f"{pres['newton_iterations'].mean():.2f}; pcg mean learned "
        f"{res['linear_iterations'].mean():.0f} plain {pres['linear_iterations'].mean():.0f}")
    assert res["newton_iterations"].mean() <= 0.75 * pres["newton_iterations"].mean()
    again = ws.solve_batch(test_cur, test_cur, accepted_options(), mixed_config(),
                           warm_iterations=12)
    assert np.array_equal(again["a"], learned["a"])
    # A zero warm budget: every item is solved from the plain start by the fallback (the
    # safeguarded policy, here in FP64 with the 1e-11 acceptance solve), which runs the
    # rejected items in teams of workgroups (summation order differs from the one-workgroup
    # safeguarded solve only).
    tight = reference_config(acceptance_linear_tolerance=1e-11)
    fb = ws.solve_batch(test_cur[:4], test_cur[:4], accepted_options(), tight, warm_iterations=0)
    sg = plan.solve_batch(test_cur[:4], accepted_options(), tight, safeguarded=True)
    for b in range(4):
        assert fb["results"][b]["status"] == 0
        assert abs(int(fb["results"][b]["newton_iterations"]) -
                   int(sg["results"][b]["newton_iterations"])) <= 1
        assert _rel_l2(fb["a"][b], sg["a"][b]) <= 1e-10
        acc.check_converged(say, f"fallback item {b}", fb["a"][b], fb["flux"][b], conv[b])
    ws.close()
    plan.close()


def test_warm_safeguarded_kernel_matches_oracle_from_the_same_start():
    say("case warm safeguarded motor_small vs oracle")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 64, seed=31)
    ws = WarmStart(plan, inputs=3, modes=16, hidden=64, seed=4)
    ws.fit(train_cur, train["a"], train_settings(steps=800, batch=32))
    cur = meshes.motor_currents(2, seed=41)
    a0 = ws.predict(cur)
    out = ws.solve_batch(cur, cur, accepted_options(), reference_config(), warm_iterations=32)
    for b in range(2):
        ref = ol.safeguarded_solve(model, cur[b], accepted_options(), warm_start=a0[b])
        r = out["results"][b]
        say(f"  item {b}: newton {r['newton_iterations']} (oracle {ref['newton']}) "
            f"dA {_rel_l2(out['a'][b], ref['a']):.2e}")
        assert r["status"] == ref["status"] == 0
        assert r["newton_iterations"] == ref["newton"]
        assert _rel_l2(out["a"][b], ref["a"]) <= 1e-9
    ws.close()
    plan.close()


def test_variants_match_per_item_oracle_models():
    """Per-item materials (coercivities, B-H scale, fill factors, rotor position) on one
    mesh: each item equals the oracle's solve of its own model (A 1e-9 rel L2, flux 1e-7,
    same Newton count); an index outside the set fails only that item."""
    say("case variants: study family 160 cells, 4 samples")
    fam = meshes.study_family(angular_cells=160)
    x = meshes.study_samples(4, seed=3)
    models = [meshes.study_model(fam, s) for s in x]
    plan = Plan(models[0])
    plan.set_variants(models)
    opts = accepted_options()
    index = np.array([2, 0, 3, 1, 7], np.int32)
    cur = np.vstack([x[i, :3] for i in (2, 0, 3, 1, 1)])
    out = plan.solve_batch_variants(index, cur, opts, reference_config())
    for b in range(4):
        ref = ol.newton_solve(models[index[b]], cur[b], opts)
        r = out["results"][b]
        say(f"  item {b} (variant {index[b]}, rotor {int(x[index[b], 10])}): status {r['status']} "
            f"newton {r['newton_iterations']} (oracle {ref['newton']}) "
            f"dA {_rel_l2(out['a'][b], ref['a']):.2e}")
        assert r["status"] == ref["status"] == 0
        assert r["newton_iterations"] == ref["newton"]
        assert _rel_l2(out["a"][b], ref["a"]) <= 1e-9
        assert np.abs(out["flux"][b] - ref["flux"]).max() <= 1e-7 * np.abs(ref["flux"]).max()
    assert out["results"][4]["status"] == 2
    # a variant set built from the plan's own model reproduces the plain solve bitwise
    plan.set_variants([models[0]])
    v = plan.solve_batch_variants(np.zeros(1, np.int32), x[:1, :3], opts, reference_config())
    p = plan.solve_batch(x[:1, :3], opts, reference_config())
    assert np.array_equal(v["a"], p["a"])
    plan.close()


def test_learned_partitioned_matches_one_workgroup_path():
    """The learned solve with 8 workgroups per item (two launches: warm attempt, then the
    plain-start fallback) passes the same gates; both it and the one-workgroup path lie
    within 1e-9 of the oracle's converged solution."""
    say("case learned partitioned: motor_small 2 items x 8 partitions")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 64, seed=31)
    ws = WarmStart(plan, inputs=3, modes=16, hidden=64, seed=4)
    ws.fit(train_cur, train["a"], train_settings(steps=800, batch=32))
    cur = meshes.motor_currents(2, seed=41)
    one = ws.solve_batch(cur, cur, accepted_options(), mixed_config(), warm_iterations=16)
    part = ws.solve_batch(cur, cur, accepted_options(), mixed_config(partitions=8),
                          warm_iterations=16)
    for b in range(2):
        r = part["results"][b]
        say(f"  item {b}: newton {r['newton_iterations']} (one-workgroup "
            f"{one['results'][b]['newton_iterations']}) dA {_rel_l2(part['a'][b], one['a'][b]):.2e}")
        assert r["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        conv = acc.converged(model, cur)[b]
        acc.check_converged(say, f"partitioned item {b}", part["a"][b], part["flux"][b], conv)
        acc.check_converged(say, f"one-workgroup item {b}", one["a"][b], one["flux"][b], conv)
    ws.close()
    plan.close()


def test_learned_block_jacobi_matches_oracle():
    """The learned solve with FP32 block-Jacobi (4-row blocks) in the mixed form: both gates
    per item and A within 1e-9 relative L2 of the oracle's converged solution, flux within
    1e-7; fewer PCG iterations than the point-Jacobi learned solve."""
    say("case learned + block-Jacobi 4: motor_small 3 items")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 64, seed=31)
    ws = WarmStart(plan, inputs=3, modes=16, hidden=64, seed=4)
    ws.fit(train_cur, train["a"], train_settings(steps=800, batch=32))
    cur = meshes.motor_currents(3, seed=43)
    opts = accepted_options()
    point = ws.solve_batch(cur, cur, opts, mixed_config(), warm_iterations=16)
    block = ws.solve_batch(cur, cur, opts, mixed_config(block_size=4), warm_iterations=16)
    conv = acc.converged(model, cur)
    for b in range(3):
        r = block["results"][b]
        say(f"  item {b}: newton {r['newton_iterations']} pcg {r['linear_iterations']} (point "
            f"{point['results'][b]['linear_iterations']})")
        assert r["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        acc.check_converged(say, f"block item {b}", block["a"][b], block["flux"][b], conv[b])
        acc.check_converged(say, f"point item {b}", point["a"][b], point["flux"][b], conv[b])
    assert block["results"]["linear_iterations"].sum() < point["results"]["linear_iterations"].sum()
    ws.close()
    plan.close()


def test_learned_split_resumes_items_in_teams():
    """The learned solve with the warm attempt split after k iterations (items still
    iterating resume from their state as partitioned teams over the whole GPU): every item
    passes both gates, lies within 1e-9 of the oracle's converged solution (as the unsplit
    solve does), an item that finished within k iterations is bitwise the unsplit one, and
    repeated runs are bitwise identical.  A split beyond every item's count reproduces the unsplit solve bitwise."""
    say("case learned split: motor_small 6 items, block-Jacobi 4, splits 2 / median / beyond")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 64, seed=31)
    ws = WarmStart(plan, inputs=3, modes=16, hidden=64, seed=4)
    ws.fit(train_cur, train["a"], train_settings(steps=800, batch=32))
    cur = meshes.motor_currents(6, seed=47)
    opts = accepted_options()
    one = ws.solve_batch(cur, cur, opts, mixed_config(block_size=4), warm_iterations=16)
    counts = one["results"]["newton_iterations"]
    assert np.all(one["results"]["status"] == 0) and counts.max() < 16
    median = int(np.median(counts))
    for split in (2, median, int(counts.max()) + 1):
        cfg = mixed_config(block_size=4, split_newton_iterations=split)
        part = ws.solve_batch(cur, cur, opts, cfg, warm_iterations=16)
        again = ws.solve_batch(cur, cur, opts, cfg, warm_iterations=16)
        for b in range(6):
            r = part["results"][b]
            say(f"  split {split} item {b}: newton {r['newton_iterations']} (unsplit "
                f"{counts[b]}) pcg {r['linear_iterations']} "
                f"dA {_rel_l2(part['a'][b], one['a'][b]):.2e}")
            assert r["status"] == 0
            assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
            acc.check_converged(say, f"split {split} item {b}", part["a"][b], part["flux"][b],
                                acc.converged(model, cur)[b])
            assert abs(int(r["newton_iterations"]) - int(counts[b])) <= 1
            if counts[b] <= split:
                assert np.array_equal(part["a"][b], one["a"][b])
                assert r["linear_iterations"] == one["results"][b]["linear_iterations"]
        assert np.array_equal(part["a"], again["a"])
        assert np.array_equal(part["flux"], again["flux"])
    ws.close()
    plan.close()


def test_inexact_forcing_matches_oracle_and_the_plain_solve():
    """Safeguarded warm attempt with inexact linear solves (forcing 0.01 x relative equation
    residual): in the FP64 form from x0 = 0 it reaches the oracle's state under the same
    policy (A within 1e-9; the linear tolerance follows the computed residual, so last-bit
    summation differences may move a PCG stopping point: Newton counts within 2, PCG counts
    within 20 %); in the chosen mixed block-Jacobi form every
    item passes both gates, lies within 1e-9 of the oracle's converged solution (flux 1e-7)
    and needs fewer PCG iterations than exact solves from the same starts."""
    say("case inexact forcing: motor_small, FP64 vs oracle (2 items), mixed block-4 (6 items)")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    train_cur, train = _small_motor_data(plan, 64, seed=31)
    ws = WarmStart(plan, inputs=3, modes=16, hidden=64, seed=4)
    ws.fit(train_cur, train["a"], train_settings(steps=800, batch=32))
    opts = accepted_options()
    cur = meshes.motor_currents(2, seed=41)
    a0 = ws.predict(cur)
    out = ws.solve_batch(cur, cur, opts, reference_config(inexact_forcing=0.01),
                         warm_iterations=32)
    for b in range(2):
        ref = ol.safeguarded_solve(model, cur[b], opts, warm_start=a0[b], forcing=0.01)
        r = out["results"][b]
        say(f"  FP64 item {b}: newton {r['newton_iterations']} (oracle {ref['newton']}) pcg "
            f"{r['linear_iterations']} (oracle {ref['linear']}) dA {_rel_l2(out['a'][b], ref['a']):.2e}")
        assert r["status"] == ref["status"] == 0
        assert abs(int(r["newton_iterations"]) - ref["newton"]) <= 2
        assert abs(int(r["linear_iterations"]) - ref["linear"]) <= 0.2 * ref["linear"]
        assert _rel_l2(out["a"][b], ref["a"]) <= 1e-9
    cur = meshes.motor_currents(6, seed=53)
    conv = acc.converged(model, cur)
    exact = ws.solve_batch(cur, cur, opts, mixed_config(block_size=4), warm_iterations=16)
    inexact = ws.solve_batch(cur, cur, opts, mixed_config(block_size=4, inexact_forcing=0.01),
                             warm_iterations=16)
    for b in range(6):
        r = inexact["results"][b]
        say(f"  mixed item {b}: newton {r['newton_iterations']} pcg {r['linear_iterations']} "
            f"(exact {exact['results'][b]['linear_iterations']})")
        assert r["status"] == 0
        assert r["relative_step"] < 1e-6 and r["equation_residual"] <= 1e-8
        acc.check_converged(say, f"mixed item {b}", inexact["a"][b], inexact["flux"][b], conv[b])
    assert inexact["results"]["linear_iterations"].sum() < \
        exact["results"]["linear_iterations"].sum()
    # The same method without the network (safeguarded inexact Newton from the plain start)
    # needs more Newton steps and PCG iterations.
    no_net = plan.solve_batch(cur, opts, mixed_config(block_size=4, inexact_forcing=0.01),
                              safeguarded=True)
    assert np.all(no_net["results"]["status"] == 0)
    for b in range(6):
        acc.check_converged(say, f"without the network item {b}", no_net["a"][b],
                            no_net["flux"][b], conv[b])
    say(f"  without the network: newton {no_net['results']['newton_iterations'].sum()} pcg "
        f"{no_net['results']['linear_iterations'].sum()}; learned: newton "
        f"{inexact['results']['newton_iterations'].sum()} pcg "
        f"{inexact['results']['linear_iterations'].sum()}")
    assert inexact["results"]["newton_iterations"].sum() < \
        no_net["results"]["newton_iterations"].sum()
    assert inexact["results"]["linear_iterations"].sum() < \
        no_net["results"]["linear_iterations"].sum()
    again = ws.solve_batch(cur, cur, opts, mixed_config(block_size=4, inexact_forcing=0.01),
                           warm_iterations=16)
    assert np.array_equal(again["a"], inexact["a"])
    ws.close()
    plan.close()


def test_redo_of_more_items_than_workgroups_matches_plain_solve():
    """A zero warm budget sends every item to the plain-start redo (ListItemsKernel +
    TeamListKernel, the safeguarded policy).  With 300 items and at most 256 co-resident
    workgroups, each team is one workgroup and some teams take a second item; a one-workgroup
    team sums exactly as the one-workgroup kernel, so every item equals the safeguarded solve
    from the plain start (femm_newton_batch_safeguarded) bit for bit."""
    say("case redo of 300 items: saturating ring 8 cells, warm budget 0")
    model, base = meshes.square_coil_ring(cells=8, seed=3, ring_bh=True)
    plan = Plan(model)
    scale = np.linspace(0.2, 1.5, 300)[:, None]
    cur = np.ascontiguousarray(base[None, :] * scale)
    train = plan.solve_batch(cur[::10], accepted_options(), reference_config())
    ws = WarmStart(plan, inputs=cur.shape[1], modes=2, hidden=16, seed=3)
    ws.fit(cur[::10], train["a"], train_settings(steps=50, batch=16))
    plain = plan.solve_batch(cur, accepted_options(), reference_config(), safeguarded=True)
    out = ws.solve_batch(cur, cur, accepted_options(), reference_config(), warm_iterations=0)
    res, pres = out["results"], plain["results"]
    say(f"  statuses {np.unique(res['status'])}, newton {res['newton_iterations'].min()}.."
        f"{res['newton_iterations'].max()}")
    assert np.all(res["status"] == 0)
    assert np.array_equal(res["newton_iterations"], pres["newton_iterations"])
    assert np.array_equal(res["linear_iterations"], pres["linear_iterations"])
    assert np.array_equal(out["a"], plain["a"])
    assert np.array_equal(out["flux"], plain["flux"])
    ws.close()
    plan.close()
