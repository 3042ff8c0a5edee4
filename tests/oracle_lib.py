"""ctypes access to the CPU oracle and to the reference host-source bridge.

Test infrastructure: the oracle is the checker, never the thing measured.
"""

import ctypes
import os
from pathlib import Path

import numpy as np

from femm_gpu_amd.model import CModel, CNewtonOptions, CMaterial, CConstraint

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "_build" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref_bridge.so"

P = ctypes.POINTER
c_d, c_i32, c_i8 = ctypes.c_double, ctypes.c_int32, ctypes.c_int8


class OracleResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("newton_iterations", ctypes.c_int32),
                ("linear_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("relative_step", ctypes.c_double), ("equation_residual", ctypes.c_double)]


def _dp(a):
    return a.ctypes.data_as(P(c_d))


def _ip(a):
    return a.ctypes.data_as(P(c_i32))


def _bp(a):
    return a.ctypes.data_as(P(c_i8))


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        _oracle = ctypes.CDLL(str(ORACLE_SO))
    return _oracle


def ref_available() -> bool:
    return REF_SO.exists()


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(str(REF_SO), mode=os.RTLD_LAZY)
    return _ref


def bh_spline(b, h, lib="oracle"):
    l = oracle() if lib == "oracle" else ref()
    fn = l.oracle_bh_spline if lib == "oracle" else l.ref_bh_spline
    n = len(b)
    b = np.ascontiguousarray(b, np.float64)
    h = np.ascontiguousarray(h, np.float64)
    ob, oh, os_ = np.zeros(n), np.zeros(n), np.zeros(n)
    cold = c_d(0)
    passes = ctypes.c_int(0)
    ok = fn(_dp(b), _dp(h), n, _dp(ob), _dp(oh), _dp(os_), ctypes.byref(cold), ctypes.byref(passes))
    return bool(ok), ob, oh, os_, cold.value, passes.value


def bh_eval(b, h, s, field, lib="oracle"):
    l = oracle() if lib == "oracle" else ref()
    fn = l.oracle_bh_eval if lib == "oracle" else l.ref_bh_eval
    out = np.zeros(4)
    b, h, s = (np.ascontiguousarray(x, np.float64) for x in (b, h, s))
    fn(_dp(b), _dp(h), _dp(s), len(b), c_d(field), _dp(out))
    return out


def validate(model, lib="oracle"):
    if lib == "oracle":
        return oracle().oracle_validate(ctypes.byref(model.c_struct()))
    return ref().ref_validate(ctypes.byref(model.c_struct()))


def assemble(model, a_old, currents, newton, lib="oracle"):
    """Reduced Newton system as a dict of numpy arrays, or raises on status."""
    l = oracle() if lib == "oracle" else ref()
    pre = "oracle_" if lib == "oracle" else "ref_"
    handle = ctypes.c_void_p()
    a_old = np.ascontiguousarray(a_old, np.float64)
    currents = np.ascontiguousarray(currents, np.float64)
    st = getattr(l, pre + "assemble")(ctypes.byref(model.c_struct()), _dp(a_old), _dp(currents),
                                      int(newton), ctypes.byref(handle))
    if st != 0:
        return st, None
    n, nnz = c_i32(), c_i32()
    getattr(l, pre + "assembly_sizes")(handle, ctypes.byref(n), ctypes.byref(nnz))
    n, nnz, N = n.value, nnz.value, model.node_count
    out = dict(rows=np.zeros(n + 1, np.int32), cols=np.zeros(nnz, np.int32),
               values=np.zeros(nnz), diag=np.zeros(n), rhs_per_amp=np.zeros(n),
               rhs_offset=np.zeros(n), free_nodes=np.zeros(n, np.int32),
               node_to_free=np.zeros(N, np.int32), node_sign=np.zeros(N, np.int8),
               boundary=np.zeros(N))
    getattr(l, pre + "assembly_copy")(handle, _ip(out["rows"]), _ip(out["cols"]),
                                      _dp(out["values"]), _dp(out["diag"]),
                                      _dp(out["rhs_per_amp"]), _dp(out["rhs_offset"]),
                                      _ip(out["free_nodes"]), _ip(out["node_to_free"]),
                                      _bp(out["node_sign"]), _dp(out["boundary"]))
    getattr(l, pre + "assembly_destroy")(handle)
    return 0, out


def prepare_symbolic(model, lib="oracle"):
    l = oracle() if lib == "oracle" else ref()
    pre = "oracle_" if lib == "oracle" else "ref_"
    handle = ctypes.c_void_p()
    st = getattr(l, pre + "prepare_symbolic")(ctypes.byref(model.c_struct()), ctypes.byref(handle))
    if st != 0:
        return st, None
    n, nnz, me, re = c_i32(), c_i32(), c_i32(), c_i32()
    getattr(l, pre + "symbolic_sizes")(handle, ctypes.byref(n), ctypes.byref(nnz),
                                       ctypes.byref(me), ctypes.byref(re))
    n, nnz, me, re, N = n.value, nnz.value, me.value, re.value, model.node_count
    out = dict(rows=np.zeros(n + 1, np.int32), cols=np.zeros(nnz, np.int32),
               diag_slots=np.zeros(n, np.int32), mat_offsets=np.zeros(nnz + 1, np.int32),
               mat_entries=np.zeros((me, 4), np.int32), rhs_offsets=np.zeros(n + 1, np.int32),
               rhs_entries=np.zeros((re, 3), np.int32), pm_rhs=np.zeros(n),
               circuit_rhs=np.zeros(max(1, model.circuit_count) * n),
               free_nodes=np.zeros(n, np.int32), node_to_free=np.zeros(N, np.int32),
               node_sign=np.zeros(N, np.int8), boundary=np.zeros(N))
    getattr(l, pre + "symbolic_copy")(
        handle, _ip(out["rows"]), _ip(out["cols"]), _ip(out["diag_slots"]),
        _ip(out["mat_offsets"]), _ip(out["mat_entries"]), _ip(out["rhs_offsets"]),
        _ip(out["rhs_entries"]), _dp(out["pm_rhs"]), _dp(out["circuit_rhs"]),
        _ip(out["free_nodes"]), _ip(out["node_to_free"]), _bp(out["node_sign"]),
        _dp(out["boundary"]))
    if lib == "oracle":
        out["age_values"], out["age_rhs"] = np.zeros(nnz), np.zeros(n)
        l.oracle_symbolic_age(handle, _dp(out["age_values"]), _dp(out["age_rhs"]))
    getattr(l, pre + "symbolic_destroy")(handle)
    out["circuit_rhs"] = out["circuit_rhs"][: model.circuit_count * n].reshape(
        model.circuit_count, n)
    return 0, out


def age_matrix(K, Ki, ci, co):
    """The oracle's 10 x 10 air-gap element cell matrix, or None for invalid inputs."""
    out = np.zeros(100)
    ok = oracle().oracle_age_matrix(ctypes.c_double(K), ctypes.c_double(Ki), ctypes.c_double(ci),
                                    ctypes.c_double(co), _dp(out))
    return out.reshape(10, 10) if ok else None


def pcg(rows, cols, values, diag, rhs, tol, maxit):
    n = len(rhs)
    x = np.zeros(n)
    info = np.zeros(3)
    rows, cols = (np.ascontiguousarray(a, np.int32) for a in (rows, cols))
    values, diag, rhs = (np.ascontiguousarray(a, np.float64) for a in (values, diag, rhs))
    oracle().oracle_pcg(n, _ip(rows), _ip(cols), _dp(values), _dp(diag), _dp(rhs), c_d(tol),
                        int(maxit), _dp(x), _dp(info))
    return x, int(info[0]), int(info[1]), info[2]


def pcg_block(rows, cols, values, diag, rhs, tol, maxit, block_size):
    n = len(rhs)
    x = np.zeros(n)
    info = np.zeros(3)
    rows, cols = (np.ascontiguousarray(a, np.int32) for a in (rows, cols))
    values, diag, rhs = (np.ascontiguousarray(a, np.float64) for a in (values, diag, rhs))
    oracle().oracle_pcg_block(n, _ip(rows), _ip(cols), _dp(values), _dp(diag), _dp(rhs),
                              c_d(tol), int(maxit), int(block_size), _dp(x), _dp(info))
    return x, int(info[0]), int(info[1]), info[2]


def block_jacobi(rows, cols, values, diag, block_size):
    """(planes[block_size][n], fallback[blocks]) of the oracle's block-Jacobi inverse."""
    n = len(diag)
    rows, cols = (np.ascontiguousarray(a, np.int32) for a in (rows, cols))
    values, diag = (np.ascontiguousarray(a, np.float64) for a in (values, diag))
    planes = np.zeros((block_size, n))
    fallback = np.zeros((n + block_size - 1) // block_size, np.int8)
    oracle().oracle_block_jacobi(n, _ip(rows), _ip(cols), _dp(values), _dp(diag), int(block_size),
                                 _dp(planes), fallback.ctypes.data_as(ctypes.c_void_p))
    return planes, fallback


class block_size:
    """Context: later oracle Newton solves use block-Jacobi of this block size."""

    def __init__(self, size):
        self.size = size

    def __enter__(self):
        oracle().oracle_set_block_size(int(self.size))

    def __exit__(self, *exc):
        oracle().oracle_set_block_size(1)


def newton_solve(model, currents, options, warm_start=None):
    N, T, C = model.node_count, model.triangle_count, model.circuit_count
    a, bx, by, flux = np.zeros(N), np.zeros(T), np.zeros(T), np.zeros(max(1, C))
    res = OracleResult()
    currents = np.ascontiguousarray(currents, np.float64)
    warm = None if warm_start is None else np.ascontiguousarray(warm_start, np.float64)
    st = oracle().oracle_newton_solve(ctypes.byref(model.c_struct()), _dp(currents),
                                      ctypes.byref(options), None if warm is None else _dp(warm),
                                      _dp(a), _dp(bx), _dp(by), _dp(flux), ctypes.byref(res))
    return dict(status=st, a=a, bx=bx, by=by, flux=flux[:C], newton=res.newton_iterations,
                linear=res.linear_iterations, step=res.relative_step,
                eq=res.equation_residual)


def safeguarded_solve(model, currents, options, warm_start=None, forcing=0.0):
    """SolveSafeguarded alone (optionally from a warm start); forcing > 0 makes the linear
    solves inexact (tolerance forcing x relative equation residual, exact before acceptance)."""
    a = np.zeros(model.node_count)
    res = OracleResult()
    currents = np.ascontiguousarray(currents, np.float64)
    warm = None if warm_start is None else np.ascontiguousarray(warm_start, np.float64)
    st = oracle().oracle_safeguarded_solve_forcing(ctypes.byref(model.c_struct()), _dp(currents),
                                                   ctypes.byref(options), c_d(forcing),
                                                   None if warm is None else _dp(warm), _dp(a),
                                                   ctypes.byref(res))
    return dict(status=st, a=a, newton=res.newton_iterations, linear=res.linear_iterations,
                step=res.relative_step, eq=res.equation_residual)


# Past the acceptance gates: the reference's SolveSafeguarded policy in FP64 with the
# equation-residual gate at 1e-11 (step gate 1e-9) and every linear solve to 1e-12.
CONVERGED_OPTIONS = dict(relative_tolerance=1e-11, linear_relative_tolerance=1e-12,
                         recover_nonconvergence=0)


def converged_solve(model, currents, start=None):
    """The oracle's solution continued past the gates: the reference driver's accepted state
    (or `start`), then safeguarded Newton with CONVERGED_OPTIONS; adds bx / by / flux."""
    from femm_gpu_amd.model import accepted_options
    if start is None:
        start = newton_solve(model, currents, accepted_options())["a"]
    r = safeguarded_solve(model, currents, accepted_options(**CONVERGED_OPTIONS), warm_start=start)
    st, bx, by, flux = finalize(model, r["a"])
    r.update(bx=bx, by=by, flux=flux, finalize_status=st)
    return r


def converged_batch(model, currents, threads=4):
    """converged_solve for every row of currents, `threads` items at a time."""
    from concurrent.futures import ThreadPoolExecutor
    currents = np.ascontiguousarray(currents, np.float64).reshape(-1, model.circuit_count)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as pool:
        return list(pool.map(lambda c: converged_solve(model, c), list(currents)))


def tangent(model, base_a, base_currents, directions, options):
    """SolveTangentMultiRhs restated: da[D][N], dflux[D][C], per-direction status / iterations."""
    N, C = model.node_count, model.circuit_count
    directions = np.ascontiguousarray(directions, np.float64).reshape(-1, C)
    D = directions.shape[0]
    da, dflux = np.zeros((D, N)), np.zeros((D, max(1, C)))
    res = (OracleResult * D)()
    base_a = np.ascontiguousarray(base_a, np.float64)
    base_currents = np.ascontiguousarray(base_currents, np.float64)
    st = oracle().oracle_tangent_multi_rhs(ctypes.byref(model.c_struct()), _dp(base_a),
                                           _dp(base_currents), D, _dp(directions),
                                           ctypes.byref(options), _dp(da), _dp(dflux), res)
    return dict(status=st, da=da, dflux=dflux[:, :C],
                results=[dict(status=r.status, linear=r.linear_iterations) for r in res])


def newton_batch(model, currents, options, threads):
    currents = np.ascontiguousarray(currents, np.float64)
    B = currents.shape[0]
    N, C = model.node_count, model.circuit_count
    a, flux = np.zeros((B, N)), np.zeros((B, C))
    results = (OracleResult * B)()
    oracle().oracle_newton_batch(ctypes.byref(model.c_struct()), B, _dp(currents),
                                 ctypes.byref(options), int(threads), _dp(a), _dp(flux), results)
    return a, flux, [dict(status=r.status, newton=r.newton_iterations, linear=r.linear_iterations,
                          step=r.relative_step, eq=r.equation_residual) for r in results]


def equation_residual(model, a, currents):
    oracle().oracle_equation_residual.restype = c_d
    a = np.ascontiguousarray(a, np.float64)
    currents = np.ascontiguousarray(currents, np.float64)
    return oracle().oracle_equation_residual(ctypes.byref(model.c_struct()), _dp(a),
                                             _dp(currents))


def finalize(model, a):
    T, C = model.triangle_count, model.circuit_count
    bx, by, flux = np.zeros(T), np.zeros(T), np.zeros(max(1, C))
    a = np.ascontiguousarray(a, np.float64)
    st = oracle().oracle_finalize(ctypes.byref(model.c_struct()), _dp(a), _dp(bx), _dp(by),
                                  _dp(flux))
    return st, bx, by, flux[:C]


def weighted_stress(model, a, bx, by, selected_label=1, air_label=3, selected_group=-1,
                    air_group=-1, radius_m=0.0, angles_rad=()):
    angles = np.ascontiguousarray(angles_rad, np.float64)
    out, br, mask = np.zeros(3), np.zeros(max(1, len(angles))), np.zeros(model.node_count)
    a, bx, by = (np.ascontiguousarray(v, np.float64) for v in (a, bx, by))
    st = oracle().oracle_weighted_stress(ctypes.byref(model.c_struct()), _dp(a), _dp(bx), _dp(by),
                                         selected_label, air_label, selected_group, air_group,
                                         c_d(radius_m), _dp(angles), len(angles), _dp(out),
                                         _dp(br), _dp(mask))
    return st, out, br[: len(angles)], mask


def ref_parse_artifact(path):
    """Model arrays from the reference artifact parser (for loader parity)."""
    text = Path(path).read_bytes()
    err = ctypes.create_string_buffer(512)
    l = ref()
    l.ref_parse_artifact.restype = ctypes.c_void_p
    handle = l.ref_parse_artifact(text, len(text), err, 512)
    if not handle:
        raise ValueError(err.value.decode())
    handle = ctypes.c_void_p(handle)
    sizes = np.zeros(6, np.int32)
    l.ref_artifact_sizes(handle, _ip(sizes))
    N, T, M, K, D, S = (int(s) for s in sizes)
    xy, tri, mat = np.zeros(2 * N), np.zeros(3 * T, np.int32), np.zeros(T, np.int32)
    mats = (CMaterial * M)()
    dn = np.zeros(D, np.int32)
    currents = np.zeros(64)
    cons = (CConstraint * max(1, S))()
    depth, problem, circuits = c_d(), c_i32(), c_i32()
    l.ref_artifact_copy(handle, _dp(xy), _ip(tri), _ip(mat), mats, _ip(dn), _dp(currents), cons,
                        ctypes.byref(depth), ctypes.byref(problem), ctypes.byref(circuits))
    ages = None
    total = c_i32()
    count = l.ref_artifact_air_gaps(handle, None, None, ctypes.byref(total))
    if count >= 0:
        from femm_gpu_amd.model import CAgePoint
        params = np.zeros(max(1, 8 * count))
        pts = (CAgePoint * max(1, total.value))()
        l.ref_artifact_air_gaps(handle, _dp(params), pts, ctypes.byref(total))
        recs = [pts[k] for k in range(total.value)]
        ages = dict(params=params[: 8 * count].reshape(count, 8),
                    nodes=np.array([list(r.node) for r in recs], np.int32).reshape(-1, 4),
                    weights=np.array([list(r.weight) for r in recs]).reshape(-1, 4))
    l.ref_artifact_destroy(handle)
    out = dict(node_xy=xy.reshape(N, 2), triangles=tri.reshape(T, 3), triangle_material=mat,
                materials=[(m.reluctivity_zero, m.alpha, m.source_j_per_a, m.hc,
                            m.magnetization_deg, m.bh_curve, m.circuit, m.group) for m in mats],
                dirichlet_nodes=dn, currents=currents[: circuits.value],
                constraints=[(c.node_a, c.node_b, c.sign) for c in list(cons)[:S]],
                depth_m=depth.value, problem_type=problem.value, curve_count=K)
    out["air_gaps"] = ages   # None without a sliding band
    return out


def ref_sliding_band(model, angle_deg):
    """The reference's ApplySlidingBandAngle on `model` posed at 0 deg: (inner shifts per
    element, records [total, 4] nodes, [total, 4] weights)."""
    from femm_gpu_amd.model import CAgePoint
    total = sum(g.cells + 1 for g in model.air_gaps)
    shifts = np.zeros(max(1, len(model.air_gaps)))
    pts = (CAgePoint * max(1, total))()
    ok = ref().ref_sliding_band(ctypes.byref(model.c_struct()), ctypes.c_double(angle_deg),
                                _dp(shifts), pts)
    if not ok:
        return None
    nodes = np.array([list(pts[k].node) for k in range(total)], np.int32)
    weights = np.array([list(pts[k].weight) for k in range(total)])
    return shifts[: len(model.air_gaps)], nodes, weights


def ref_age_postprocess(model, a, angles_rad):
    """The reference's ComputeAirGapElementPostprocess: (status, [Fx, Fy, torque], radial B)."""
    a = np.ascontiguousarray(a, np.float64)
    angles = np.ascontiguousarray(angles_rad, np.float64)
    out, radial = np.zeros(3), np.zeros(max(1, len(angles)))
    st = ref().ref_age_postprocess(ctypes.byref(model.c_struct()), _dp(a), _dp(angles),
                                   len(angles), _dp(out), _dp(radial))
    return st, out, radial[: len(angles)]


def age_postprocess(model, a, angles_rad):
    """The oracle's air-gap element postprocess: (status, [Fx, Fy, torque], radial B)."""
    a = np.ascontiguousarray(a, np.float64)
    angles = np.ascontiguousarray(angles_rad, np.float64)
    out, radial = np.zeros(3), np.zeros(max(1, len(angles)))
    st = oracle().oracle_age_postprocess(ctypes.byref(model.c_struct()), _dp(a), _dp(angles),
                                         len(angles), _dp(out), _dp(radial))
    return st, out, radial[: len(angles)]
