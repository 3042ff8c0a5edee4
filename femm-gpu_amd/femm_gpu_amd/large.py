"""Single-RHS solves of large models with the partitioned solve.

    python -m femm_gpu_amd.large --out large.json [--cases c_core_62k,c_core_1m,solenoid_rz_249k]

Cases: the C-core with a 2 mm gap (NdFeB block, two coils at +-5e5 A/m^2, B-H iron) at
0.5 mm iron edges (62,237 nodes) and 0.1 mm (1,163,107 nodes), and the r-z solenoid with a
B-H plunger (249,865 nodes, 7-point quadrature, Br / Bz per element).  Each case is solved
in the FP64 reference form (point Jacobi PCG), in the mixed form (FP32 PCG iterations with
FP64 reliable updates to the same 1e-8 equation residual) and in the mixed form with
block-Jacobi (block size 4, FP32 planes), each with the autotuned number of workgroups, and
the last form once more as the safeguarded Newton with inexact linear solves (forcing 0.01,
femm_newton_batch_safeguarded) from the plain start;
reported per form: wall time, Newton and PCG counts, gates, and each mixed form's A relative
L2 / flux / element-B difference to the FP64 form.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _log(*args):
    print(*args, file=sys.stderr, flush=True)


def _cases():
    from . import meshes
    return {
        "c_core_62k": lambda: meshes.c_core(),
        "c_core_1m": lambda: meshes.c_core(edge_iron=1e-4),
        "solenoid_rz_249k": lambda: meshes.solenoid_rz(radial_cells=352, axial_cells=704),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", default="c_core_62k,c_core_1m,solenoid_rz_249k")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import torch  # noqa: F401  (initialises the device the library runs on)

    from .model import accepted_options
    from .solver import Plan, mixed_config, reference_config

    cases = _cases()
    report = {}
    for name in [c.strip() for c in args.cases.split(",") if c.strip()]:
        t0 = time.perf_counter()
        model, currents = cases[name]()
        plan = Plan(model)
        _log(f"# {name}: N={plan.N} T={plan.T} n={plan.n} nnz={plan.nnz} "
             f"(mesh + plan {time.perf_counter() - t0:.1f} s)")
        opts = accepted_options(max_newton_iterations=64, max_linear_iterations=200000)
        forms = {}
        for form, cfg, bs, forcing in (("fp64", reference_config, 1, 0.0),
                                       ("mixed", mixed_config, 1, 0.0),
                                       ("mixed_block4", mixed_config, 4, 0.0),
                                       ("mixed_block4_inexact", mixed_config, 4, 0.01)):
            parts = plan.autotune(1, cfg(partitions=0, block_size=bs))
            t = time.perf_counter()
            out = plan.solve_batch(np.asarray(currents)[None, :], opts,
                                   cfg(partitions=0, block_size=bs, inexact_forcing=forcing),
                                   fields=True, safeguarded=forcing > 0)
            wall = time.perf_counter() - t
            r = out["results"][0]
            forms[form] = dict(out=out, summary=dict(
                partitions=parts, wall_s=wall, status=int(r["status"]),
                newton=int(r["newton_iterations"]), pcg=int(r["linear_iterations"]),
                replacements=int(r["refinements"]), relative_step=float(r["relative_step"]),
                equation_residual=float(r["equation_residual"]),
                flux=[float(v) for v in out["flux"][0]]))
            _log(f"#   {form}: {wall:.2f} s with {parts} workgroups, status {r['status']}, newton "
                 f"{r['newton_iterations']}, pcg {r['linear_iterations']}, eq "
                 f"{r['equation_residual']:.2e}")
        a64, f64 = forms["fp64"]["out"]["a"][0], forms["fp64"]["out"]["flux"][0]
        b64 = np.hypot(forms["fp64"]["out"]["bx"][0], forms["fp64"]["out"]["by"][0])
        report[name] = dict(N=plan.N, T=plan.T, n=plan.n, nnz=plan.nnz)
        for form in forms:
            report[name][form] = forms[form]["summary"]
            if form == "fp64":
                continue
            am, fm = forms[form]["out"]["a"][0], forms[form]["out"]["flux"][0]
            bm = np.hypot(forms[form]["out"]["bx"][0], forms[form]["out"]["by"][0])
            report[name][form]["vs_fp64"] = dict(
                a_rel_l2=float(np.linalg.norm(am - a64) / np.linalg.norm(a64)),
                flux_rel=float(np.max(np.abs(fm - f64)) / max(np.max(np.abs(f64)), 1e-300)),
                b_max_abs_T=float(np.max(np.abs(bm - b64))),
                speedup=forms["fp64"]["summary"]["wall_s"] / forms[form]["summary"]["wall_s"])
            _log(f"#   {form} vs fp64: {report[name][form]['vs_fp64']}")
        plan.close()
    text = json.dumps(report, indent=1)
    print(text, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
