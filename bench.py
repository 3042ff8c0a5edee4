"""Benchmark: batched nonlinear magnetostatic solves per second on one MI355X.

Workload: 256 samples of a ~60k-node 12-slot / 4-pole anti-periodic motor sector
(58,682 nodes, 116,160 P1 triangles, B-H iron, two NdFeB blocks, three phase coils),
phase current densities ~ U(-8e6, 8e6) A/m^2 from a fixed seed.  One step = one batched
nonlinear solve of all samples to the undamped 1e-6 step and the 1e-8 equation residual
(assembly, PCG, state updates, equation-residual checks and flux reductions all on the GPU),
followed by the weighted-stress force / torque on the rotor body and the smoothed radial air-gap
B at 32 probes for every item (femm_postprocess_batch_device).  An item counts only if it passes
both gates.  Inputs are resident in HBM before timing.

Forms (both timed, same items, same gates):
* plain_fp64   - the reference algorithm: FP64 Jacobi PCG from x0 = 0, Newton with the
                 reference relaxation policy from the plain start.
* learned_mixed (the reported value) - learned warm start (POD + linear path + BF16 MFMA
                 MLP, predicted on the GPU inside the timed step), safeguarded Newton from
                 it with inexact linear solves (tolerance forcing x relative equation
                 residual, exact before acceptance), FP32 PCG with FP64 reliable updates and
                 4-row block-Jacobi started from the current state, plain-start fallback for
                 any item not accepted within the warm budget.  The network is trained
                 beforehand, untimed, on solver solutions of a disjoint seeded sample set;
                 its cost is reported as `training`.
Optional forms (--forms): mixed_plain_start (the reference driver in the mixed form) and
inexact_plain_start (the chosen form without the network: safeguarded inexact Newton from
the plain start).

The plain FP64 form is timed over one step after an untimed call that allocates the batch
workspace (about 37 s per step); the chosen form over --steps steps after --warmup steps.
Accuracy: every form is compared with the same items solved past the gates (the reference's
safeguarded Newton in FP64 from the plain FP64 form's states, continued to a 1e-11 equation
residual with 1e-12 linear solves): A relative L2, flux linkage and torque relative errors.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "femm-gpu_amd"))

HBM_PEAK_GBS = 8000.0
TRAIN_SEED = 777
WARM_BUDGET = 16


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, when it was taken
    on this kernel and workload (FETCH_SIZE / WRITE_SIZE scaled by a copy dispatch)."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        rec = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    if rec.get("kernel") != kernel or rec.get("workload") != workload:
        return None
    return rec.get("hbm_bytes_per_launch")


def algorithmic_bytes(plan, results, B, mixed, block_size=1):
    """Algorithmic HBM bytes of one batched solve from the per-item kernel counters.

    PCG iteration (per item): s_v nnz + (4 nnz + 4 (n + 1)) / S + 13 s_x n with S = B (the
    CSR pattern is shared by every item); s_v = s_x = 4 for the FP32 iterations of the
    mixed form and 8 for FP64 ones (every iteration of the FP64 form; the FP64 finishes and
    acceptance re-solves of the mixed form, counted by the kernel as fp64_iterations).
    With block-Jacobi of b rows the vector passes are 12 + b (the b planes replace D^-1
    and z is stored), and each Newton iteration also rebuilds the planes (8 nnz + 8 n read,
    (8 + s_x) b n written).
    FP64 residual replacement (mixed form): 8 nnz + (4 nnz + 4 (n + 1)) / S + 76 n.
    Newton iteration (per item): element pass (read A: 8 N, write 12 doubles per
    triangle), slot and row gathers (read 12 doubles per triangle, write 8 nnz + 16 n),
    step and update passes (24 N), equation-residual SpMV (8 nnz + 32 n); a trial
    assembly of the safeguarded policy costs the same less the step pass (8 N read,
    16 N for the trial state)."""
    N, T, n, nnz = plan.N, plan.T, plan.n, plan.nnz
    shared = (4 * nnz + 4 * (n + 1)) / B
    passes = 13 if block_size <= 1 else 12 + block_size
    pcg64 = 8 * nnz + shared + passes * 8 * n
    pcg32 = 4 * nnz + shared + passes * 4 * n
    pcg_iter = pcg32 if mixed else pcg64
    replacement = 8 * nnz + shared + 76 * n
    assembly = 8 * N + 96 * T + 96 * T + 8 * nnz + 16 * n
    newton = assembly + 24 * N + 8 * nnz + 32 * n
    trial = assembly + 16 * N + 8 * nnz + 32 * n
    if block_size > 1:
        newton += 8 * nnz + 8 * n + (8 + (4 if mixed else 0)) * block_size * n
    total = 0.0
    for r in results:
        it64 = int(r["fp64_iterations"]) if mixed else int(r["linear_iterations"])
        it32 = int(r["linear_iterations"]) - it64 if mixed else 0
        total += (it32 * pcg32 + it64 * pcg64 + r["newton_iterations"] * newton
                  + r["trial_assemblies"] * trial
                  + (r["refinements"] * replacement if mixed else 0.0) + 2 * assembly)
    return total, pcg_iter


class Timed:
    """HIP events on the stream the library launches on.  A step is a sequence of phases
    (callables enqueueing work on that stream); an event is recorded between phases, so
    each phase's device time is known without synchronising inside the timed region."""

    def __init__(self, torch):
        self.torch = torch
        self.stream = torch.cuda.current_stream()

    def run(self, phases, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            for phase in phases:
                phase()
        torch.cuda.synchronize()
        marks = []
        t0 = time.perf_counter()
        for _ in range(steps):
            row = []
            for phase in phases:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(self.stream)
                row.append(ev)
                phase()
            marks.append(row)
        stop = torch.cuda.Event(enable_timing=True)
        stop.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per_phase = [0.0] * len(phases)
        for k, row in enumerate(marks):
            ends = row[1:] + [marks[k + 1][0] if k + 1 < steps else stop]
            for j, (a, b) in enumerate(zip(row, ends)):
                per_phase[j] += a.elapsed_time(b)
        total = marks[0][0].elapsed_time(stop) / steps
        return total, [v / steps for v in per_phase], wall


def host_cores():
    """(cores this process may run on, os.cpu_count()): the affinity mask, capped by a
    cgroup v2 CPU quota when one is set."""
    count = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else count
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            usable = min(usable, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return usable, count


def cpu_baseline(model, currents, options):
    """The CPU oracle (an FP64 port of the reference's host assembler + Jacobi PCG + Newton),
    one item per host thread on every usable host core, over a bounded sample of the same
    workload (one item per core)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as ol
    threads, nproc = host_cores()
    sample = currents[:threads]
    t0 = time.perf_counter()
    _, _, res = ol.newton_batch(model, sample, options, threads)
    wall = time.perf_counter() - t0
    ok = sum(1 for r in res if r["status"] == 0 and r["step"] < 1e-6 and r["eq"] <= 1e-8)
    return dict(value=ok / wall, unit="solves/s", cores=threads, kind="port",
                sample=f"the first {len(sample)} of the same 256 motor samples, one per thread "
                       f"on {threads} usable host cores (os.cpu_count() = {nproc}), "
                       f"{ok} passing, {wall:.1f} s wall")


def accuracy(forms, converged, B):
    """Per-form worst-item errors against the converged solutions (A relative L2, flux
    linkage relative to the item's largest |flux|, torque relative)."""
    out = {}
    ra, rf, rt = converged["a"], converged["flux"], converged["loads"][:, 2]
    for name, f in forms.items():
        a_err = ((f["a"] - ra).norm(dim=1) / ra.norm(dim=1).clamp_min(1e-300))
        f_err = ((f["flux"] - rf).abs().amax(dim=1) / rf.abs().amax(dim=1).clamp_min(1e-300))
        t_err = (f["loads"][:, 2] - rt).abs() / rt.abs().clamp_min(1e-300)
        out[name] = dict(max_A_rel_l2=a_err.max().item(), mean_A_rel_l2=a_err.mean().item(),
                         items_A_within_1e9=int((a_err <= 1e-9).sum().item()),
                         max_flux_rel=f_err.max().item(), max_torque_rel=t_err.max().item(),
                         items=B)
    return out


def artifact_parse(model, currents):
    """The setup cost reported beside the solve: the same mesh written as a reference-layout
    mesh artifact (magnetostatic v1: node constraints carry the sector's anti-periodic pairs)
    and parsed by the library's strict parser (femm_artifact_parse: strict JSON, canonical
    identity SHA-256, mm -> m and per-region labels, the model checks).  The best of three
    parses; the solve itself uses the generated model."""
    import tempfile

    from femm_gpu_amd import wire

    with tempfile.TemporaryDirectory() as tmp:
        path = Path(tmp) / "motor_sector.gpu.json"
        wire.write_artifact(path, model, currents, schema="gpu_femm_magnetostatic_mesh_v1")
        data = path.read_bytes()
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        parsed, _ = wire.parse_artifact(data)
        best = min(best, time.perf_counter() - t0)
    if not (np.array_equal(parsed.triangles, model.triangles)
            and np.array_equal(parsed.constraints, model.constraints)
            and np.allclose(parsed.node_xy, model.node_xy, rtol=0.0, atol=1e-15)):
        raise SystemExit("artifact round trip changed the mesh")
    return {"schema": "gpu_femm_magnetostatic_mesh_v1", "bytes": len(data), "parse_s": best}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed steps of the plain FP64 form (about 37 s each)")
    ap.add_argument("--items", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20251203)
    ap.add_argument("--train-samples", type=int, default=1024,
                    help="untimed warm-start training solves (disjoint seed)")
    ap.add_argument("--train-steps", type=int, default=24000,
                    help="Adam steps of the warm-start network fit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--angular-cells", type=int, default=480)
    ap.add_argument("--forms", default="plain_fp64,learned_mixed",
                    help="comma-separated subset of plain_fp64,learned_mixed,mixed_plain_start,"
                         "inexact_plain_start")
    ap.add_argument("--block-size", type=int, default=4,
                    help="PCG preconditioner of the mixed forms: 4 = block-Jacobi of 4 rows "
                         "(FP32, the chosen form), 1 = point Jacobi")
    ap.add_argument("--forcing", type=float, default=0.01,
                    help="learned form: inexact-Newton forcing of the safeguarded warm attempt "
                         "(linear tolerance = forcing x relative equation residual; 0 = exact)")
    ap.add_argument("--inner-tolerance", type=float, default=0.1,
                    help="mixed forms: FP64 residual replacement once the FP32 residual has "
                         "dropped by this factor since the last one")
    ap.add_argument("--split-newton", type=int, default=0,
                    help="learned form: warm-attempt iterations per item in the first launch "
                         "before the items still iterating resume as partitioned teams "
                         "(0 = one launch)")
    ap.add_argument("--probes", type=int, default=32, help="air-gap B probes per item")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first, as the known "
                         "reference dispatch for FETCH_SIZE / WRITE_SIZE scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from femm_gpu_amd import meshes
    from femm_gpu_amd.learned import WarmStart, train_settings
    from femm_gpu_amd.model import accepted_options
    from femm_gpu_amd.solver import RESULT_DTYPE, Plan, device_info, mixed_config, reference_config

    torch.cuda.set_device(0)
    model, _ = meshes.motor_sector(angular_cells=args.angular_cells)
    B = args.items
    currents = meshes.motor_currents(B, seed=args.seed)
    artifact = artifact_parse(model, currents[0])
    t0 = time.perf_counter()
    plan = Plan(model)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan.set_postprocess(**meshes.motor_postprocess(args.probes))
    post_setup_s = time.perf_counter() - t0
    currents_dev = torch.from_numpy(currents).to("cuda")
    options = accepted_options()
    wanted = [f.strip() for f in args.forms.split(",") if f.strip()]
    print(f"# device {device_info()}; mesh {model.name}: N={plan.N} T={plan.T} n={plan.n} "
          f"nnz={plan.nnz}; artifact parse {artifact['parse_s']:.3f} s "
          f"({artifact['bytes'] / 1e6:.1f} MB), plan setup {setup_s:.2f} s, "
          f"postprocess plan {post_setup_s:.2f} s",
          file=sys.stderr, flush=True)

    if args.calibrate_copy_mib > 0:
        count = args.calibrate_copy_mib * (1 << 20) // 4
        src = torch.ones(count, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# calibration copy: {args.calibrate_copy_mib} MiB read + "
              f"{args.calibrate_copy_mib} MiB written", file=sys.stderr, flush=True)
        del src, dst

    timer = Timed(torch)
    stream = timer.stream.cuda_stream
    a = torch.zeros((B, plan.N), dtype=torch.float64, device="cuda")
    flux = torch.zeros((B, plan.C), dtype=torch.float64, device="cuda")
    res = torch.zeros(B * RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    loads = torch.zeros((B, 3), dtype=torch.float64, device="cuda")
    radial = torch.zeros((B, max(1, args.probes)), dtype=torch.float64, device="cuda")

    def postprocess():
        plan.postprocess_batch_device(B, a.data_ptr(), loads.data_ptr(), radial.data_ptr(),
                                      stream)

    def collect(timing, mixed, block_size=1, steps=None, warmup=None):
        ms, (solve_ms, post_ms), wall = timing
        r = np.frombuffer(res.cpu().numpy().tobytes(), dtype=RESULT_DTYPE).copy()
        passing = (r["status"] == 0) & (r["relative_step"] < 1e-6) & \
                  (r["equation_residual"] <= 1e-8)
        bytes_total, pcg_iter_bytes = algorithmic_bytes(plan, r, B, mixed, block_size)
        newton = r["newton_iterations"]
        hist = {int(q): int(v) for q, v in zip(*np.unique(newton, return_counts=True))}
        return dict(ms=ms, solve_ms=solve_ms, postprocess_ms=post_ms, wall=wall,
                    steps=steps, warmup=warmup, passing=int(passing.sum()), newton_hist=hist,
                    solves_per_s=float(passing.sum()) / (ms * 1e-3) if ms > 0 else None,
                    newton_mean=float(r["newton_iterations"].mean()),
                    pcg_mean=float(r["linear_iterations"].mean()),
                    pcg_fp64_mean=float(r["fp64_iterations"].mean()),
                    trial_mean=float(r["trial_assemblies"].mean()),
                    pcg_max=int(r["linear_iterations"].max()),
                    max_equation_residual=float(r["equation_residual"].max()),
                    achieved_gbs=bytes_total / (solve_ms * 1e-3) / 1e9 if solve_ms > 0 else None,
                    pcg_iter_bytes=pcg_iter_bytes,
                    a=a.clone(), flux=flux.clone(), loads=loads.clone(), radial=radial.clone())

    def spmv_rate(fp32, reps=40):
        """The PCG's SpMV alone on the operators the last solve left (reps products per item):
        algorithmic bytes s_v nnz + (4 nnz + 4 n) / B + 2 s_x n per product."""
        dt = torch.float32 if fp32 else torch.float64
        x = torch.ones((B, plan.n), dtype=dt, device="cuda")
        y = torch.empty_like(x)
        prec = 1 if fp32 else 0
        plan.spmv_batch_device(B, prec, 2, x.data_ptr(), y.data_ptr(), stream)
        ms, _, _ = timer.run([lambda: plan.spmv_batch_device(B, prec, reps, x.data_ptr(),
                                                             y.data_ptr(), stream)], 1, 0)
        s = 4 if fp32 else 8
        per = s * plan.nnz + (4 * plan.nnz + 4 * plan.n) / B + 2 * s * plan.n
        gbs = per * B * reps / (ms * 1e-3) / 1e9
        return dict(gb_s=gbs, frac=gbs / HBM_PEAK_GBS, us_per_product=ms * 1e3 / reps,
                    bytes_per_item_product=per, items=B, reps=reps)

    def report(name, f):
        print(f"# {name}: {f['ms']:.1f} ms/step (solve {f['solve_ms']:.1f}, postprocess "
              f"{f['postprocess_ms']:.2f}), pass {f['passing']}/{B}, newton "
              f"{f['newton_mean']:.2f}, pcg/item {f['pcg_mean']:.0f} (max {f['pcg_max']}), "
              f"{f['achieved_gbs']:.0f} GB/s algorithmic; newton counts {f['newton_hist']}",
              file=sys.stderr, flush=True)

    def plain_solve(cfg, opts=options, safeguarded=False, warm=None):
        return lambda: plan.solve_batch_device(
            currents_dev.data_ptr(), B, a.data_ptr(), flux.data_ptr(), res.data_ptr(), stream,
            opts, cfg, warm_ptr=warm, safeguarded=safeguarded)

    forms = {}
    training = None
    ws = None
    acc = None
    if "plain_fp64" in wanted:
        cfg = reference_config()
        # Untimed: one Newton iteration of one PCG iteration allocates the batch workspace.
        plain_solve(cfg, accepted_options(max_newton_iterations=1, max_linear_iterations=1,
                                          recover_nonconvergence=0))()
        torch.cuda.synchronize()
        forms["plain_fp64"] = collect(timer.run([plain_solve(cfg), postprocess],
                                                args.plain_steps, 0), False,
                                      steps=args.plain_steps, warmup=0)
        report("plain_fp64", forms["plain_fp64"])
        forms["plain_fp64"]["spmv"] = spmv_rate(False)
        print(f"# spmv fp64: {forms['plain_fp64']['spmv']}", file=sys.stderr, flush=True)
    if "mixed_plain_start" in wanted:
        cfg = mixed_config(block_size=args.block_size)
        forms["mixed_plain_start"] = collect(timer.run([plain_solve(cfg), postprocess],
                                                       args.steps, args.warmup), True,
                                             args.block_size, args.steps, args.warmup)
        report("mixed_plain_start", forms["mixed_plain_start"])
    if "inexact_plain_start" in wanted:
        # The chosen form without the network: safeguarded inexact Newton from the plain start.
        cfg = mixed_config(block_size=args.block_size, inexact_forcing=args.forcing)
        forms["inexact_plain_start"] = collect(
            timer.run([plain_solve(cfg, safeguarded=True), postprocess], args.steps,
                      args.warmup), True, args.block_size, args.steps, args.warmup)
        report("inexact_plain_start", forms["inexact_plain_start"])
    if "learned_mixed" in wanted:
        # Untimed training: solver solutions of a disjoint seeded sample set.
        t0 = time.perf_counter()
        S = args.train_samples
        train_cur = torch.from_numpy(meshes.motor_currents(S, seed=TRAIN_SEED)).to("cuda")
        ta = torch.zeros((S, plan.N), dtype=torch.float64, device="cuda")
        tf = torch.zeros((S, plan.C), dtype=torch.float64, device="cuda")
        tr = torch.zeros(S * RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        plan.solve_batch_device(train_cur.data_ptr(), S, ta.data_ptr(), tf.data_ptr(),
                                tr.data_ptr(), stream, options,
                                mixed_config(block_size=args.block_size))
        torch.cuda.synchronize()
        solve_s = time.perf_counter() - t0
        ws = WarmStart(plan, inputs=3, modes=64, hidden=256, seed=12044)
        fit = ws.fit_device(S, train_cur.data_ptr(), ta.data_ptr(), stream,
                            train_settings(steps=args.train_steps, batch=128,
                                           learning_rate=2e-3))
        torch.cuda.synchronize()
        training = dict(samples=S, seed=TRAIN_SEED, steps=args.train_steps, solve_s=solve_s,
                        fit_s=time.perf_counter() - t0 - solve_s, report=fit)
        print(f"# training: {S} solves {solve_s:.1f} s, fit {training['fit_s']:.1f} s, "
              f"POD energy {fit['pod_energy_fraction']:.8f}, loss {fit['first_loss']:.3e}"
              f" -> {fit['final_loss']:.3e}", file=sys.stderr, flush=True)
        del ta, tf, tr, train_cur
        cfg = mixed_config(block_size=args.block_size, split_newton_iterations=args.split_newton,
                           inexact_forcing=args.forcing,
                           inner_relative_tolerance=args.inner_tolerance)

        def learned_solve():
            ws.solve_batch_device(B, currents_dev.data_ptr(), currents_dev.data_ptr(),
                                  a.data_ptr(), flux.data_ptr(), res.data_ptr(), stream, options,
                                  cfg, WARM_BUDGET)

        cfg_accept = cfg.acceptance_linear_tolerance
        forms["learned_mixed"] = collect(timer.run([learned_solve, postprocess], args.steps,
                                                   args.warmup), True, args.block_size,
                                         args.steps, args.warmup)
        report("learned_mixed", forms["learned_mixed"])
        # The same solve through the host-pointer entry point: inputs copied in, A / flux /
        # results copied out (PCIe-inclusive; not the headline value).
        ws.solve_batch(currents, currents, options, cfg, WARM_BUDGET)   # untimed warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        host = ws.solve_batch(currents, currents, options, cfg, WARM_BUDGET)
        host_s = time.perf_counter() - t0
        host_ok = int(((host["results"]["status"] == 0)
                       & (host["results"]["relative_step"] < 1e-6)
                       & (host["results"]["equation_residual"] <= 1e-8)).sum())
        forms["learned_mixed"]["host_buffers"] = dict(
            ms=host_s * 1e3, solves_per_s=host_ok / host_s, items_passing=host_ok,
            bytes_out=int(host["a"].nbytes + host["flux"].nbytes + host["results"].nbytes),
            postprocess="not included")
        print(f"# learned_mixed through host buffers: {host_s * 1e3:.1f} ms, pass {host_ok}/{B}",
              file=sys.stderr, flush=True)
        del host
        forms["learned_mixed"]["spmv"] = spmv_rate(True)
        print(f"# spmv fp32: {forms['learned_mixed']['spmv']}", file=sys.stderr, flush=True)

    # Converged solutions (untimed): the reference's safeguarded Newton in FP64 (x0 = 0 in
    # every linear solve) continued past the gates, from the plain FP64 form's states (a path
    # independent of the chosen form's) when that form ran.
    if forms:
        last = forms.get("plain_fp64") or forms[list(forms)[-1]]
        warm = last["a"].clone()
        tight = accepted_options(relative_tolerance=1e-11, linear_relative_tolerance=1e-12,
                                 recover_nonconvergence=0)
        t0 = time.perf_counter()
        plain_solve(reference_config(), tight, True, warm.data_ptr())()
        postprocess()
        torch.cuda.synchronize()
        conv = collect((0.0, (0.0, 0.0), 0.0), False)
        conv_s = time.perf_counter() - t0
        print(f"# converged reference: {conv['passing']}/{B} within the gates, max equation "
              f"residual {conv['max_equation_residual']:.2e}, {conv_s:.1f} s",
              file=sys.stderr, flush=True)
        acc = accuracy(forms, conv, B)
        acc["converged_reference"] = dict(
            method="FP64 safeguarded Newton from the plain FP64 form's states, equation "
                   "residual <= 1e-11, 1e-12 linear solves from x0 = 0", max_equation_residual=conv["max_equation_residual"],
            items_meeting_tight_gates=int(conv["passing"]))
        print(f"# accuracy vs converged: {acc}", file=sys.stderr, flush=True)
        del warm

    if "plain_fp64" not in forms or "learned_mixed" not in forms:
        print(json.dumps({"profiling_run": True, "accuracy": acc, "forms": {
            k: {f: v for f, v in d.items() if f not in ("a", "flux", "loads", "radial")}
            for k, d in forms.items()}}), flush=True)
        if ws is not None:
            ws.close()
        plan.close()
        return
    plain, chosen = forms["plain_fp64"], forms["learned_mixed"]
    diff = chosen["a"] - plain["a"]
    rel_l2 = (diff.norm(dim=1) / plain["a"].norm(dim=1).clamp_min(1e-300)).max().item()
    value = chosen["solves_per_s"]
    achieved = chosen["achieved_gbs"]

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(model, currents, options)
        print(f"# cpu baseline {cpu}", file=sys.stderr, flush=True)


    def summary(f):
        return {"solves_per_s": f["solves_per_s"], "ms_per_step": f["ms"],
                "solve_ms": f["solve_ms"], "postprocess_ms": f["postprocess_ms"],
                "steps": f["steps"], "warmup": f["warmup"], "items_passing": f["passing"],
                "newton_mean": f["newton_mean"], "pcg_iterations_mean": f["pcg_mean"],
                "pcg_iterations_max": f["pcg_max"],
                "pcg_fp64_iterations_mean": f["pcg_fp64_mean"],
                "trial_assemblies_mean": f["trial_mean"], "achieved_GBps": f["achieved_gbs"]}

    line = {
        "metric": "nonlinear solves/sec @1e-8 resid, 60k nodes batch 256: plain FP64 vs warm+mixed",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 PCG iterations (values, SpMV, block-Jacobi) with f64 residual "
                 "replacement, f64 assembly / Newton / residuals / flux / torque; bf16 MFMA "
                 "warm-start network",
        "data": "synthetic: seeded 12-slot motor sector mesh and seeded phase currents",
        "config": {"workload": f"motor_sector 12-slot/4-pole anti-periodic, N={plan.N} nodes, "
                               f"T={plan.T} triangles, batch {B}, U(-8e6,8e6) A/m^2 phase "
                               f"currents; step = solve + flux + weighted-stress torque + "
                               f"{args.probes} air-gap B probes per item",
                   "form": "learned warm start (POD 64 + linear + MLP 3-256-256-64 bf16) -> "
                           "safeguarded Newton"
                           + (f" (inexact, forcing {args.forcing:g})" if args.forcing > 0 else "")
                           + ", FP32 PCG "
                           + (f"(block-Jacobi, {args.block_size}-row blocks) "
                              if args.block_size > 1 else "(point Jacobi) ")
                           + "with FP64 reliable updates from the current state, plain-start "
                           f"fallback after {WARM_BUDGET} iterations",
                   "chosen": summary(chosen),
                   "plain_fp64": summary(plain),
                   "speedup_vs_plain_fp64": value / max(plain["solves_per_s"], 1e-30),
                   "accuracy_vs_converged": acc,
                   "max_A_rel_l2_chosen_vs_plain_fp64": rel_l2,
                   "host_buffers": chosen["host_buffers"],
                   "artifact": artifact,
                   "plan_setup_s": setup_s, "postprocess_plan_s": post_setup_s,
                   "training": training},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(
                         f"learned_mixed block {args.block_size} acceptance "
                         f"{cfg_accept:g}", f"{model.name} batch {B}"),
                     "kernel": f"SolveKernel<{3 if args.block_size > 1 else 1}, 1> (mixed "
                               "form, safeguarded attempt from the learned start; the plain-"
                               "start redo launch of rejected items is timed with it; the "
                               "postprocess launch is excluded)",
                     "pcg_iteration_bytes_per_item": chosen["pcg_iter_bytes"],
                     "spmv_alone_fp32": chosen["spmv"],
                     "spmv_alone_fp64": plain["spmv"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ws.close()
    plan.close()


if __name__ == "__main__":
    main()
