"""Uneven meshes in one launch: 64 independent square-coil / iron-ring meshes.

    python -m femm_gpu_amd.group_bench --meshes 64 --out group.json

Meshes come from meshes.square_coil_ring (0.2 m air box, 2 triangles per cell, jittered
interior nodes, J = 1e6 A/m^2 coil, mu_r = 1000 ring) with cells per side drawn
log-uniformly from 32..512 (seeded), so node counts run from ~1k to ~260k; one item per
mesh, FP64 reference form, 1e-6 step / 1e-8 equation-residual gates.  Forms timed:

* per_mesh: one launch per mesh, each with the partition count femm_plan_autotune chose
  for a single item (meshes solved one after another on the stream);
* group: one cooperative launch over all meshes with the group autotuner's partitions
  (work priced as sqrt(free DOFs));
* group_tuned: the same after re-tuning with each mesh's measured PCG iterations as work;
* group_inexact / group_inexact_mixed: the tuned launch running the safeguarded Newton with
  inexact linear solves (forcing 0.01, femm_newton_group_safeguarded) in the FP64 and the
  mixed form;
* group_graph: group_tuned captured once in the group's HIP graph (femm_group_capture) and
  replayed (femm_group_replay).

Reported per form: GPU time, solves/s, items passing, and the largest A / flux difference
to the per-mesh form over the meshes that pass the gates in both.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _log(*args):
    print(*args, file=sys.stderr, flush=True)


def _passing(r):
    return (r["status"] == 0) & (r["relative_step"] < 1e-6) & (r["equation_residual"] <= 1e-8)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=64)
    ap.add_argument("--min-cells", type=int, default=32)
    ap.add_argument("--max-cells", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--max-linear", type=int, default=16384,
                    help="PCG iteration cap per Newton step (16,384: ProductionSampleOptions)")
    ap.add_argument("--block-size", type=int, default=1,
                    help="1 = point Jacobi (the reference default), 4 = block-Jacobi")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import torch

    from . import meshes
    from .model import accepted_options
    from .solver import RESULT_DTYPE, Group, Plan, mixed_config, reference_config

    torch.cuda.set_device(0)
    rng = np.random.default_rng(args.seed)
    cells = np.exp(rng.uniform(np.log(args.min_cells), np.log(args.max_cells), args.meshes))
    cells = np.round(cells).astype(int)
    cells[np.argmin(cells)], cells[np.argmax(cells)] = args.min_cells, args.max_cells
    t0 = time.perf_counter()
    plans, currents = [], []
    for i, c in enumerate(cells):
        model, cur = meshes.square_coil_ring(cells=int(c), seed=args.seed + i)
        plans.append(Plan(model))
        currents.append(cur)
    nodes = [p.N for p in plans]
    _log(f"# {len(plans)} meshes, cells {cells.min()}..{cells.max()}, nodes {min(nodes)}.."
         f"{max(nodes)} (total {sum(nodes)}), setup {time.perf_counter() - t0:.1f} s")
    opts = accepted_options(max_linear_iterations=args.max_linear)
    cfg = reference_config(block_size=args.block_size)
    stream = torch.cuda.current_stream()
    bufs = []
    for p, cur in zip(plans, currents):
        bufs.append(dict(cur=torch.from_numpy(np.ascontiguousarray(cur[None, :])).cuda(),
                         a=torch.zeros((1, p.N), dtype=torch.float64, device="cuda"),
                         flux=torch.zeros((1, p.C), dtype=torch.float64, device="cuda"),
                         res=torch.zeros(RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")))

    def read(b):
        return (b["a"].cpu().numpy()[0], b["flux"].cpu().numpy()[0],
                np.frombuffer(b["res"].cpu().numpy().tobytes(), dtype=RESULT_DTYPE)[0])

    def timed(fn):
        best = None
        for _ in range(args.repeats):
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        return best

    report = dict(meshes=len(plans), cells=[int(c) for c in cells], nodes=nodes, forms={})

    # per_mesh
    t = time.perf_counter()
    parts_single = [p.autotune(1, reference_config(partitions=0, block_size=args.block_size))
                    for p in plans]
    _log(f"# per-mesh autotune {time.perf_counter() - t:.1f} s: partitions {parts_single}")

    def per_mesh():
        for p, b, k in zip(plans, bufs, parts_single):
            p.solve_batch_device(b["cur"].data_ptr(), 1, b["a"].data_ptr(), b["flux"].data_ptr(),
                                 b["res"].data_ptr(), stream.cuda_stream, opts,
                                 reference_config(partitions=k, block_size=args.block_size))

    ms = timed(per_mesh)
    base = [read(b) for b in bufs]
    res = np.array([r for _, _, r in base], dtype=RESULT_DTYPE)
    report["forms"]["per_mesh"] = dict(ms=ms, solves_per_s=len(plans) / (ms / 1e3),
                                       passing=int(_passing(res).sum()),
                                       partitions=parts_single,
                                       pcg_iterations=[int(v) for v in res["linear_iterations"]],
                                       newton=[int(v) for v in res["newton_iterations"]])
    _log(f"# per_mesh: {ms:.1f} ms, pass {int(_passing(res).sum())}/{len(plans)}")

    group = Group([dict(plan=p, items=1, currents=b["cur"].data_ptr(), a=b["a"].data_ptr(),
                        flux=b["flux"].data_ptr(), results=b["res"].data_ptr())
                   for p, b in zip(plans, bufs)])

    def run_group():
        group.solve(stream.cuda_stream, opts, cfg)

    def record(name, ms, extra):
        outs = [read(b) for b in bufs]
        res = np.array([r for _, _, r in outs], dtype=RESULT_DTYPE)
        # Differences over the meshes that pass the gates in both forms.
        both = [(o, b0) for o, b0 in zip(outs, base)
                if _passing(np.array([o[2]], dtype=RESULT_DTYPE))[0]
                and _passing(np.array([b0[2]], dtype=RESULT_DTYPE))[0]]
        da = max([float(np.linalg.norm(a - a0) / max(np.linalg.norm(a0), 1e-300))
                  for (a, _, _), (a0, _, _) in both] or [0.0])
        df = max([float(np.abs(f - f0).max() / max(np.abs(f0).max(), 1e-300))
                  for (_, f, _), (_, f0, _) in both] or [0.0])
        report["forms"][name] = dict(ms=ms, solves_per_s=len(plans) / (ms / 1e3),
                                     passing=int(_passing(res).sum()), max_A_rel_l2_vs_per_mesh=da,
                                     max_flux_rel_vs_per_mesh=df, **extra)
        _log(f"# {name}: {ms:.1f} ms, pass {int(_passing(res).sum())}/{len(plans)}, dA {da:.2e}, "
             f"dflux {df:.2e}, {extra.get('blocks', '')} blocks")
        return res

    t = time.perf_counter()
    parts, blocks = group.autotune(cfg)
    tune_s = time.perf_counter() - t
    ms = timed(run_group)
    res = record("group", ms, dict(partitions=parts, blocks=blocks, autotune_s=tune_s))

    work = np.maximum(res["linear_iterations"].astype(np.float64), 1.0)
    t = time.perf_counter()
    parts, blocks = group.autotune(cfg, work=work)
    tune_s = time.perf_counter() - t
    ms = timed(run_group)
    record("group_tuned", ms, dict(partitions=parts, blocks=blocks, autotune_s=tune_s))

    # The tuned launch with the safeguarded Newton and inexact linear solves (forcing 0.01),
    # in the FP64 form and in the mixed form.
    for name, icfg in (("group_inexact", reference_config(block_size=args.block_size,
                                                          inexact_forcing=0.01)),
                       ("group_inexact_mixed", mixed_config(block_size=args.block_size,
                                                            inexact_forcing=0.01))):
        ms = timed(lambda: group.solve(stream.cuda_stream, opts, icfg, safeguarded=True))
        record(name, ms, dict(partitions=parts, blocks=blocks))

    # The tuned group launch captured in the library's HIP graph (femm_group_capture) and
    # replayed (femm_group_replay).
    group.capture(opts, cfg)
    ms = timed(lambda: group.replay(stream.cuda_stream))
    record("group_graph", ms, dict(partitions=parts, blocks=blocks))
    per = report["forms"]["per_mesh"]["ms"]
    for k in ("group", "group_tuned", "group_inexact", "group_inexact_mixed", "group_graph"):
        if "ms" in report["forms"].get(k, {}):
            report["forms"][k]["speedup_vs_per_mesh"] = per / report["forms"][k]["ms"]
    group.close()
    for p in plans:
        p.close()
    text = json.dumps(report, indent=1)
    print(text, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
