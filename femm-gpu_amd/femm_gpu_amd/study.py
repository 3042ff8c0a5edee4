"""Learned warm-start study on the 60k-node motor with per-sample materials.

    python -m femm_gpu_amd.study --train 4096 --test 1024 --out results.json

Per sample (11 inputs, :func:`meshes.study_samples`): three phase current densities, two
magnet coercivities, the B-H saturation scale, four coil-region fill factors and a rotor
position index; every sample is its own material variant on one mesh
(:func:`meshes.study_family`).  Training solves (mixed precision, plain start) give the
POD basis (64 modes) and the 11-256-256-64 BF16 MLP; the held-out samples are then solved
twice at the same 1e-6 step / 1e-8 equation-residual gates: the plain FP64 reference form
from the plain start, and the learned start with the mixed-precision form.  Reported per
form: items passing, Newton and PCG iteration means, solves/s, and the A / flux difference
between the two forms.  Work is chunked so every GPU step prints within a minute.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _log(*args):
    print(*args, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--test", type=int, default=1024)
    ap.add_argument("--chunk", type=int, default=256, help="held-out items per launch")
    ap.add_argument("--train-chunk", type=int, default=512, help="training items per launch")
    ap.add_argument("--angular-cells", type=int, default=480)
    ap.add_argument("--train-seed", type=int, default=5001)
    ap.add_argument("--test-seed", type=int, default=9001)
    ap.add_argument("--modes", type=int, default=64)
    ap.add_argument("--steps", type=int, default=24000)
    ap.add_argument("--warm-budget", type=int, default=24,
                    help="safeguarded iterations of the warm attempt before the plain-start "
                         "fallback (the inexact attempt needs more, cheaper iterations)")
    ap.add_argument("--forcing", type=float, default=0.01,
                    help="inexact-Newton forcing of the warm attempt (0 = exact linear solves)")
    ap.add_argument("--block-size", type=int, default=4,
                    help="mixed-form PCG preconditioner: 4 = FP32 block-Jacobi, 1 = point")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import torch

    from . import meshes
    from .learned import WarmStart, train_settings
    from .model import accepted_options
    from .solver import RESULT_DTYPE, Plan, mixed_config, reference_config

    torch.cuda.set_device(0)
    fam = meshes.study_family(angular_cells=args.angular_cells)
    x_train = meshes.study_samples(args.train, seed=args.train_seed)
    x_test = meshes.study_samples(args.test, seed=args.test_seed)
    plan = Plan(meshes.study_model(fam, x_train[0]))
    opts = accepted_options()
    stream = torch.cuda.current_stream()
    N, C = plan.N, plan.C
    _log(f"# study mesh N={N} T={plan.T} n={plan.n} nnz={plan.nnz}; train {args.train}, "
         f"test {args.test}, chunk {args.chunk}")

    def run_chunks(x, solve, label, chunk):
        """solve(index_dev, cur_dev, count, a_dev, flux_dev, res_dev) per chunk of variants."""
        total = len(x)
        a_all = torch.zeros((total, N), dtype=torch.float64, device="cuda")
        flux_all = np.zeros((total, C))
        res_all = np.zeros(total, RESULT_DTYPE)
        gpu_ms = 0.0
        for lo in range(0, total, chunk):
            hi = min(total, lo + chunk)
            t0 = time.perf_counter()
            plan.set_variants([meshes.study_model(fam, s) for s in x[lo:hi]])
            setup = time.perf_counter() - t0
            B = hi - lo
            index = torch.arange(B, dtype=torch.int32, device="cuda")
            cur = torch.from_numpy(np.ascontiguousarray(x[lo:hi, :3])).to("cuda")
            flux = torch.zeros((B, C), dtype=torch.float64, device="cuda")
            res = torch.zeros(B * RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            solve(index, cur, B, a_all[lo:hi], flux, res, lo)
            ev1.record(stream)
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1)
            gpu_ms += ms
            r = np.frombuffer(res.cpu().numpy().tobytes(), dtype=RESULT_DTYPE)
            res_all[lo:hi] = r
            flux_all[lo:hi] = flux.cpu().numpy()
            _log(f"#   {label} {lo}..{hi}: {ms / 1e3:.1f} s GPU, variants {setup:.1f} s host, "
                 f"pass {int(_passing(r).sum())}/{B}, newton {r['newton_iterations'].mean():.2f}")
        return a_all, flux_all, res_all, gpu_ms

    # Training solves (mixed precision, plain start).
    t0 = time.perf_counter()
    a_train, _, res_train, train_ms = run_chunks(
        x_train, lambda idx, cur, B, a, flux, res, lo: plan.solve_batch_variants_device(
            idx.data_ptr(), cur.data_ptr(), B, a.data_ptr(), flux.data_ptr(), res.data_ptr(),
            stream.cuda_stream, opts, mixed_config(block_size=args.block_size)), "train",
        args.train_chunk)
    ok = _passing(res_train)
    feats = torch.from_numpy(np.ascontiguousarray(x_train[ok])).to("cuda")
    states = a_train[torch.from_numpy(np.flatnonzero(ok)).to("cuda")].contiguous()
    del a_train
    ws = WarmStart(plan, inputs=11, modes=args.modes, hidden=256, seed=12044)
    t1 = time.perf_counter()
    fit = ws.fit_device(int(ok.sum()), feats.data_ptr(), states.data_ptr(), stream.cuda_stream,
                        train_settings(steps=args.steps, batch=256, learning_rate=2e-3))
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t1
    _log(f"# fit {fit_s:.1f} s: {fit}")
    del feats, states

    # Held-out: plain FP64 reference form vs learned start + mixed form.
    a_plain, flux_plain, res_plain, plain_ms = run_chunks(
        x_test, lambda idx, cur, B, a, flux, res, lo: plan.solve_batch_variants_device(
            idx.data_ptr(), cur.data_ptr(), B, a.data_ptr(), flux.data_ptr(), res.data_ptr(),
            stream.cuda_stream, opts, reference_config()), "plain", args.chunk)
    feats_test = torch.from_numpy(np.ascontiguousarray(x_test)).to("cuda")
    a_learn, flux_learn, res_learn, learn_ms = run_chunks(
        x_test, lambda idx, cur, B, a, flux, res, lo: ws.solve_batch_device(
            B, feats_test[lo:lo + B].data_ptr(), cur.data_ptr(), a.data_ptr(), flux.data_ptr(),
            res.data_ptr(), stream.cuda_stream, opts,
            mixed_config(block_size=args.block_size, inexact_forcing=args.forcing),
            args.warm_budget,
            variant_ptr=idx.data_ptr()), "learned", args.chunk)
    da = ((a_learn - a_plain).norm(dim=1) / a_plain.norm(dim=1).clamp_min(1e-300)).cpu().numpy()
    dflux = np.abs(flux_learn - flux_plain).max(1) / np.maximum(np.abs(flux_plain).max(1), 1e-300)

    def summary(res, ms):
        p = _passing(res)
        return dict(items=len(res), passing=int(p.sum()), solves_per_s=float(p.sum()) / (ms / 1e3),
                    gpu_s=ms / 1e3, newton_mean=float(res["newton_iterations"].mean()),
                    pcg_mean=float(res["linear_iterations"].mean()))

    out = dict(mesh=dict(N=N, T=plan.T, n=plan.n, nnz=plan.nnz),
               train=dict(samples=args.train, seed=args.train_seed, passing=int(ok.sum()),
                          gpu_s=train_ms / 1e3, wall_s=time.perf_counter() - t0 - fit_s,
                          fit_s=fit_s, report=fit),
               test_seed=args.test_seed, warm_budget=args.warm_budget,
               block_size=args.block_size, forcing=args.forcing,
               plain_fp64=summary(res_plain, plain_ms),
               learned_mixed=summary(res_learn, learn_ms),
               max_A_rel_l2=float(da.max()), mean_A_rel_l2=float(da.mean()),
               max_flux_rel=float(dflux.max()),
               warm_accepted=int((res_learn["newton_iterations"] <= args.warm_budget).sum()))
    out["speedup"] = out["learned_mixed"]["solves_per_s"] / max(out["plain_fp64"]["solves_per_s"],
                                                                1e-30)
    text = json.dumps(out, indent=1)
    print(text, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)


def _passing(r):
    return (r["status"] == 0) & (r["relative_step"] < 1e-6) & (r["equation_residual"] <= 1e-8)


if __name__ == "__main__":
    main()
