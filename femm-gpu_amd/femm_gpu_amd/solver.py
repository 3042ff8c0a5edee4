"""ctypes binding of the C ABI (include/femm_gpu_amd.h).

Mirrors the reference's C++ solver interface for this path
(``NonlinearP1Solver`` in gpu_solver/src/solvers/nonlinear_solver.cuh:32-85):
``Plan(model)`` is Initialize + device plan, ``Plan.solve_batch`` is SolveBatch
(+ FinalizeResult), ``Plan.assemble_batch`` / ``Plan.pcg_batch`` expose the device
assembly and PCG of DeviceNonlinearAssemblyPlan / GpuCsrSolver for tests.

The native library is required: if it is missing or was built without a GPU
target, loading fails loudly.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import sys
from pathlib import Path

import numpy as np

from .model import CModel, CNewtonOptions, Model, production_options, status_name

LIB_PATH = Path(__file__).resolve().parents[1] / "lib" / "libfemm_gpu_amd.so"

P = ctypes.POINTER
c_d, c_i32 = ctypes.c_double, ctypes.c_int32

FP64 = 0
MIXED = 1


class SolveConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("pcg_warm_start", ctypes.c_int32),
                ("verify_equation_residual", ctypes.c_int32), ("max_refinements", ctypes.c_int32),
                ("inner_relative_tolerance", ctypes.c_double), ("partitions", ctypes.c_int32),
                ("block_size", ctypes.c_int32), ("split_newton_iterations", ctypes.c_int32),
                ("inexact_forcing", ctypes.c_double),
                ("acceptance_linear_tolerance", ctypes.c_double)]


class ItemResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("newton_iterations", ctypes.c_int32),
                ("linear_iterations", ctypes.c_int32), ("refinements", ctypes.c_int32),
                ("relative_step", ctypes.c_double), ("equation_residual", ctypes.c_double),
                ("last_linear_tolerance", ctypes.c_double), ("linear_residual", ctypes.c_double),
                ("fp64_iterations", ctypes.c_int32), ("trial_assemblies", ctypes.c_int32)]


class PostprocessOptions(ctypes.Structure):
    _fields_ = [("selected_label", ctypes.c_int32), ("air_label", ctypes.c_int32),
                ("selected_group", ctypes.c_int32), ("air_group", ctypes.c_int32),
                ("airgap_radius_m", ctypes.c_double), ("angles_rad", P(ctypes.c_double)),
                ("angle_count", ctypes.c_int32), ("max_mask_iterations", ctypes.c_int32),
                ("mask_relative_tolerance", ctypes.c_double)]


RESULT_DTYPE = np.dtype([("status", np.int32), ("newton_iterations", np.int32),
                         ("linear_iterations", np.int32), ("refinements", np.int32),
                         ("relative_step", np.float64), ("equation_residual", np.float64),
                         ("last_linear_tolerance", np.float64), ("linear_residual", np.float64),
                         ("fp64_iterations", np.int32), ("trial_assemblies", np.int32)])


def reference_config(**kw) -> SolveConfig:
    """The reference algorithm: FP64 Jacobi PCG from x0 = 0 in every Newton correction."""
    values = dict(precision=FP64, pcg_warm_start=0, verify_equation_residual=1,
                  max_refinements=64, inner_relative_tolerance=0.1, partitions=1, block_size=1)
    values.update(kw)
    return SolveConfig(**values)


def mixed_config(**kw) -> SolveConfig:
    """FP32 Jacobi PCG with FP64 reliable updates (true residual replaced whenever the FP32
    residual has dropped 10x), each Newton correction starting from the current state.  In
    the safeguarded policy (learned solves, femm_newton_batch_safeguarded) the step that meets
    both gates is solved again in FP64 PCG to 1e-12 and returned at that candidate, so the
    state is fixed to about 1e-12 times the operator's conditioning (a re-solve that does not
    converge falls back to the exact tolerance); the reference driver ignores it.  The warm
    starts carry error in the soft (low-reluctivity) modes that a 1e-11 residual does not
    see on every model: the FP64 re-solve has to iterate past it."""
    values = dict(precision=MIXED, pcg_warm_start=1, verify_equation_residual=1,
                  max_refinements=64, inner_relative_tolerance=0.1, partitions=1, block_size=1,
                  acceptance_linear_tolerance=1e-12)
    values.update(kw)
    return SolveConfig(**values)


_lib = None


def library():
    """Load libfemm_gpu_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.femm_status_name.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(P(c_d))


def _ip(a):
    return a.ctypes.data_as(P(c_i32))


class FemmError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {status_name(code)} ({code})")
        self.code = code


def _check(code, what):
    if code != 0:
        raise FemmError(code, what)


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    library().femm_device_info(buf, 256)
    return buf.value.decode()


class Plan:
    """One geometry uploaded to the GPU (reference NonlinearP1Solver::Initialize)."""

    def __init__(self, model: Model):
        self.model = model
        self._handle = ctypes.c_void_p()
        self._lib = library()
        _check(self._lib.femm_plan_create(ctypes.byref(model.c_struct()),
                                          ctypes.byref(self._handle)), "femm_plan_create")
        sizes = np.zeros(6, np.int32)
        _check(self._lib.femm_plan_sizes(self._handle, _ip(sizes)), "femm_plan_sizes")
        self.N, self.T, self.n, self.nnz, self.C, self.Q = (int(s) for s in sizes)

    def close(self):
        if self._handle:
            self._lib.femm_plan_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        # At interpreter shutdown the HIP runtime may already be torn down: leave the
        # handle to process exit instead of calling into the device from a finalizer.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._handle

    def symbolic(self):
        out = dict(rows=np.zeros(self.n + 1, np.int32), cols=np.zeros(self.nnz, np.int32),
                   diag_slots=np.zeros(self.n, np.int32), free_nodes=np.zeros(self.n, np.int32),
                   node_to_free=np.zeros(self.N, np.int32), node_sign=np.zeros(self.N, np.int8),
                   boundary=np.zeros(self.N))
        _check(self._lib.femm_plan_copy_symbolic(
            self._handle, _ip(out["rows"]), _ip(out["cols"]), _ip(out["diag_slots"]),
            _ip(out["free_nodes"]), _ip(out["node_to_free"]),
            out["node_sign"].ctypes.data_as(P(ctypes.c_int8)), _dp(out["boundary"])),
            "femm_plan_copy_symbolic")
        out["age_values"], out["age_rhs"] = np.zeros(self.nnz), np.zeros(self.n)
        _check(self._lib.femm_plan_copy_air_gap_terms(self._handle, _dp(out["age_values"]),
                                                      _dp(out["age_rhs"])),
               "femm_plan_copy_air_gap_terms")
        return out

    def assemble_batch(self, states, currents, newton):
        states = np.ascontiguousarray(states, np.float64).reshape(-1, self.N)
        B = states.shape[0]
        currents = np.ascontiguousarray(currents, np.float64).reshape(B, self.C)
        values, rhs, diag = np.zeros((B, self.nnz)), np.zeros((B, self.n)), np.zeros((B, self.n))
        code = self._lib.femm_assemble_batch(self._handle, B, _dp(states), _dp(currents),
                                             int(newton), _dp(values), _dp(rhs), _dp(diag))
        return code, values, rhs, diag

    def pcg_batch(self, values, rhs, tol, maxit, block_size=1):
        """Batched PCG (femm_pcg_batch; block_size > 1: femm_pcg_batch_block)."""
        values = np.ascontiguousarray(values, np.float64).reshape(-1, self.nnz)
        B = values.shape[0]
        rhs = np.ascontiguousarray(rhs, np.float64).reshape(B, self.n)
        x, info = np.zeros((B, self.n)), np.zeros((B, 3))
        if block_size == 1:
            _check(self._lib.femm_pcg_batch(self._handle, B, _dp(values), _dp(rhs), c_d(tol),
                                            int(maxit), _dp(x), _dp(info)), "femm_pcg_batch")
        else:
            _check(self._lib.femm_pcg_batch_block(self._handle, B, _dp(values), _dp(rhs),
                                                  c_d(tol), int(maxit), int(block_size), _dp(x),
                                                  _dp(info)), "femm_pcg_batch_block")
        return x, info

    def solve_batch(self, currents, options: CNewtonOptions = None, config: SolveConfig = None,
                    warm_starts=None, fields=False, safeguarded=False):
        """Batched Newton solve (SolveBatch + FinalizeResult; with safeguarded=True the
        batched SolveSafeguarded, femm_newton_batch_safeguarded).  Returns a dict of arrays."""
        currents = np.ascontiguousarray(currents, np.float64).reshape(-1, self.C)
        B = currents.shape[0]
        options = options or production_options()
        config = config or reference_config()
        a, flux = np.zeros((B, self.N)), np.zeros((B, self.C))
        bx = np.zeros((B, self.T)) if fields else None
        by = np.zeros((B, self.T)) if fields else None
        results = np.zeros(B, RESULT_DTYPE)
        warm = None if warm_starts is None else np.ascontiguousarray(warm_starts, np.float64)
        name = "femm_newton_batch_safeguarded" if safeguarded else "femm_newton_batch"
        _check(getattr(self._lib, name)(
            self._handle, B, _dp(currents), ctypes.byref(options), ctypes.byref(config),
            None if warm is None else _dp(warm), _dp(a), None if bx is None else _dp(bx),
            None if by is None else _dp(by), _dp(flux),
            results.ctypes.data_as(ctypes.c_void_p)), name)
        out = dict(a=a, flux=flux, results=results)
        if fields:
            out.update(bx=bx, by=by)
        return out

    def solve_batch_device(self, currents_ptr, items, a_ptr, flux_ptr, results_ptr, stream_ptr,
                           options: CNewtonOptions = None, config: SolveConfig = None,
                           warm_ptr=None, safeguarded=False):
        """Device-pointer entry (femm_newton_batch[_safeguarded]_device); nothing is
        synchronised."""
        options = options or production_options()
        config = config or reference_config()
        name = "femm_newton_batch_safeguarded_device" if safeguarded else "femm_newton_batch_device"
        _check(getattr(self._lib, name)(
            self._handle, int(items), ctypes.c_void_p(currents_ptr), ctypes.byref(options),
            ctypes.byref(config), None if warm_ptr is None else ctypes.c_void_p(warm_ptr),
            ctypes.c_void_p(a_ptr), ctypes.c_void_p(flux_ptr), ctypes.c_void_p(results_ptr),
            ctypes.c_void_p(stream_ptr)), name)

    def solve_batch_device_noalloc(self, currents_ptr, items, a_ptr, flux_ptr, results_ptr,
                                   stream_ptr, options: CNewtonOptions = None,
                                   config: SolveConfig = None, warm_ptr=None):
        """femm_newton_batch_device_noalloc: the device-pointer solve with its workspace in
        the library's fixed device array (no device allocation; the plan must already be
        resident).  Returns the status code instead of raising, so callers can see the
        array's size limit (FEMM_GPU_ALLOCATION_FAILED)."""
        options = options or production_options()
        config = config or reference_config()
        return int(self._lib.femm_newton_batch_device_noalloc(
            self._handle, int(items), ctypes.c_void_p(currents_ptr), ctypes.byref(options),
            ctypes.byref(config), None if warm_ptr is None else ctypes.c_void_p(warm_ptr),
            ctypes.c_void_p(a_ptr), ctypes.c_void_p(flux_ptr), ctypes.c_void_p(results_ptr),
            ctypes.c_void_p(stream_ptr)))

    def spmv_batch_device(self, items, precision, reps, x_ptr, y_ptr, stream_ptr):
        """reps products per item on the workspace operator (femm_spmv_batch_device)."""
        _check(self._lib.femm_spmv_batch_device(self._handle, int(items), int(precision),
                                                int(reps), ctypes.c_void_p(x_ptr),
                                                ctypes.c_void_p(y_ptr),
                                                ctypes.c_void_p(stream_ptr)),
               "femm_spmv_batch_device")

    def autotune(self, items, config: SolveConfig = None):
        """Time partition counts for batches of `items` items (femm_plan_autotune); later
        solves with config.partitions = 0 use the fastest.  Returns it."""
        config = config or reference_config()
        out = ctypes.c_int32(0)
        _check(self._lib.femm_plan_autotune(self._handle, int(items), ctypes.byref(config),
                                            ctypes.byref(out)), "femm_plan_autotune")
        return int(out.value)

    def set_variants(self, models):
        """Per-item material variants on this plan's mesh (femm_plan_set_variants)."""
        models = list(models)
        structs = (CModel * max(1, len(models)))()
        for i, m in enumerate(models):
            structs[i] = m.c_struct()
        self._variant_models = models          # keep the arrays alive while the call runs
        _check(self._lib.femm_plan_set_variants(self._handle, len(models), structs),
               "femm_plan_set_variants")
        self.variant_count = len(models)
        self._variant_models = None

    def solve_batch_variants(self, variant_of_item, currents, options: CNewtonOptions = None,
                             config: SolveConfig = None, warm_starts=None):
        """Batched Newton with item b on variant variant_of_item[b]."""
        currents = np.ascontiguousarray(currents, np.float64).reshape(-1, self.C)
        B = currents.shape[0]
        index = np.ascontiguousarray(variant_of_item, np.int32).reshape(B)
        options = options or production_options()
        config = config or reference_config()
        a, flux = np.zeros((B, self.N)), np.zeros((B, self.C))
        results = np.zeros(B, RESULT_DTYPE)
        warm = None if warm_starts is None else np.ascontiguousarray(warm_starts, np.float64)
        _check(self._lib.femm_newton_batch_variants(
            self._handle, B, _ip(index), _dp(currents), ctypes.byref(options),
            ctypes.byref(config), None if warm is None else _dp(warm), _dp(a), _dp(flux),
            results.ctypes.data_as(ctypes.c_void_p)), "femm_newton_batch_variants")
        return dict(a=a, flux=flux, results=results)

    def solve_batch_variants_device(self, variant_ptr, currents_ptr, items, a_ptr, flux_ptr,
                                    results_ptr, stream_ptr, options: CNewtonOptions = None,
                                    config: SolveConfig = None, warm_ptr=None):
        options = options or production_options()
        config = config or reference_config()
        _check(self._lib.femm_newton_batch_variants_device(
            self._handle, int(items), ctypes.c_void_p(variant_ptr), ctypes.c_void_p(currents_ptr),
            ctypes.byref(options), ctypes.byref(config),
            None if warm_ptr is None else ctypes.c_void_p(warm_ptr), ctypes.c_void_p(a_ptr),
            ctypes.c_void_p(flux_ptr), ctypes.c_void_p(results_ptr), ctypes.c_void_p(stream_ptr)),
            "femm_newton_batch_variants_device")

    def set_postprocess(self, selected_label=-1, air_label=-1, selected_group=-1, air_group=-1,
                        radius_m=0.0, angles_rad=(), max_mask_iterations=256,
                        mask_relative_tolerance=1e-12):
        """Weighted-stress mask and probe plan (BuildFrozenPostprocessPlan); on a model with
        air-gap elements the elements' own postprocess (Maxwell stress on the gap's mean
        circle, Fourier radial B at `angles_rad`), which ignores the selections and radius."""
        self._angles = np.ascontiguousarray(angles_rad, np.float64)
        opts = PostprocessOptions(selected_label, air_label, selected_group, air_group,
                                  radius_m, _dp(self._angles), len(self._angles),
                                  max_mask_iterations, mask_relative_tolerance)
        _check(self._lib.femm_plan_set_postprocess(self._handle, ctypes.byref(opts)),
               "femm_plan_set_postprocess")
        self.probe_count = len(self._angles)

    def postprocess_mask(self):
        mask = np.zeros(self.N)
        _check(self._lib.femm_postprocess_copy_mask(self._handle, _dp(mask)),
               "femm_postprocess_copy_mask")
        return mask

    def postprocess_batch(self, states):
        """(loads[B][3] = Fx, Fy, torque; radial_b[B][probes]) for nodal states."""
        states = np.ascontiguousarray(states, np.float64).reshape(-1, self.N)
        B = states.shape[0]
        loads, radial = np.zeros((B, 3)), np.zeros((B, max(1, self.probe_count)))
        _check(self._lib.femm_postprocess_batch(self._handle, B, _dp(states), _dp(loads),
                                                _dp(radial)), "femm_postprocess_batch")
        return loads, radial[:, : self.probe_count]

    def postprocess_batch_device(self, items, states_ptr, loads_ptr, radial_ptr, stream_ptr):
        _check(self._lib.femm_postprocess_batch_device(
            self._handle, int(items), ctypes.c_void_p(states_ptr), ctypes.c_void_p(loads_ptr),
            ctypes.c_void_p(radial_ptr), ctypes.c_void_p(stream_ptr)),
            "femm_postprocess_batch_device")

    def tangent_multi_rhs(self, base_a, base_currents, directions, options: CNewtonOptions = None,
                          fields=False):
        """Shared-operator tangent solves dA/dI along circuit-current directions at a
        converged state (femm_tangent_multi_rhs; SolveTangentMultiRhs)."""
        base_a = np.ascontiguousarray(base_a, np.float64).reshape(self.N)
        base_currents = np.ascontiguousarray(base_currents, np.float64).reshape(self.C)
        directions = np.ascontiguousarray(directions, np.float64).reshape(-1, self.C)
        D = directions.shape[0]
        options = options or production_options()
        da, dflux = np.zeros((D, self.N)), np.zeros((D, self.C))
        dbx = np.zeros((D, self.T)) if fields else None
        dby = np.zeros((D, self.T)) if fields else None
        results = np.zeros(D, RESULT_DTYPE)
        _check(self._lib.femm_tangent_multi_rhs(
            self._handle, _dp(base_a), _dp(base_currents), D, _dp(directions),
            ctypes.byref(options), _dp(da), None if dbx is None else _dp(dbx),
            None if dby is None else _dp(dby), _dp(dflux),
            results.ctypes.data_as(ctypes.c_void_p)), "femm_tangent_multi_rhs")
        out = dict(da=da, dflux=dflux, results=results)
        if fields:
            out.update(dbx=dbx, dby=dby)
        return out

    def finalize_batch(self, states):
        states = np.ascontiguousarray(states, np.float64).reshape(-1, self.N)
        B = states.shape[0]
        bx, by, flux = np.zeros((B, self.T)), np.zeros((B, self.T)), np.zeros((B, self.C))
        _check(self._lib.femm_finalize_batch(self._handle, B, _dp(states), _dp(bx), _dp(by),
                                             _dp(flux)), "femm_finalize_batch")
        return bx, by, flux


class GroupEntry(ctypes.Structure):
    _fields_ = [("plan", ctypes.c_void_p), ("items", ctypes.c_int32),
                ("partitions", ctypes.c_int32), ("currents", ctypes.c_void_p),
                ("a", ctypes.c_void_p), ("flux", ctypes.c_void_p), ("results", ctypes.c_void_p)]


class Group:
    """Plans of uneven size solved together in one cooperative launch (femm_group_*).

    entries: sequence of dicts with plan, items, currents / a / flux / results (device
    pointers of the sizes Plan.solve_batch_device takes) and optionally partitions (1).
    The plans must stay open while the group is used."""

    def __init__(self, entries):
        self._lib = library()
        self.plans = [e["plan"] for e in entries]
        arr = (GroupEntry * len(entries))()
        for i, e in enumerate(entries):
            arr[i] = GroupEntry(e["plan"].handle, int(e["items"]), int(e.get("partitions", 1)),
                                e["currents"], e["a"], e["flux"], e["results"])
        self.count = len(entries)
        self.partitions = [int(e.get("partitions", 1)) for e in entries]
        self._handle = ctypes.c_void_p()
        _check(self._lib.femm_group_create(self.count, arr, ctypes.byref(self._handle)),
               "femm_group_create")

    def close(self):
        if self._handle:
            self._lib.femm_group_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        # At interpreter shutdown the HIP runtime may already be torn down: leave the
        # handle to process exit instead of calling into the device from a finalizer.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def set_partitions(self, partitions):
        p = np.ascontiguousarray(partitions, np.int32).reshape(self.count)
        _check(self._lib.femm_group_set_partitions(self._handle, _ip(p)),
               "femm_group_set_partitions")
        self.partitions = [int(v) for v in p]

    def autotune(self, config: SolveConfig = None, work=None):
        """Per-entry partitions from the group autotuner; returns (partitions, blocks)."""
        config = config or reference_config()
        p = np.zeros(self.count, np.int32)
        blocks = ctypes.c_int32(0)
        w = None if work is None else np.ascontiguousarray(work, np.float64).reshape(self.count)
        _check(self._lib.femm_group_autotune(self._handle, ctypes.byref(config),
                                             None if w is None else _dp(w), _ip(p),
                                             ctypes.byref(blocks)), "femm_group_autotune")
        self.partitions = [int(v) for v in p]
        return self.partitions, int(blocks.value)

    def solve(self, stream_ptr, options: CNewtonOptions = None, config: SolveConfig = None,
              safeguarded=False):
        """One launch over every entry (femm_newton_group, or with safeguarded=True
        femm_newton_group_safeguarded); nothing is synchronised."""
        options = options or production_options()
        config = config or reference_config()
        name = "femm_newton_group_safeguarded" if safeguarded else "femm_newton_group"
        _check(getattr(self._lib, name)(self._handle, ctypes.byref(options),
                                        ctypes.byref(config), ctypes.c_void_p(stream_ptr)), name)

    def capture(self, options: CNewtonOptions = None, config: SolveConfig = None,
                safeguarded=False):
        """Capture the launches of one solve(...) call into the group's HIP graph
        (femm_group_capture); synchronises the device."""
        options = options or production_options()
        config = config or reference_config()
        _check(self._lib.femm_group_capture(self._handle, ctypes.byref(options),
                                            ctypes.byref(config), int(bool(safeguarded))),
               "femm_group_capture")

    def replay(self, stream_ptr):
        """One replay of the captured graph on the stream (femm_group_replay); nothing is
        synchronised."""
        _check(self._lib.femm_group_replay(self._handle, ctypes.c_void_p(stream_ptr)),
               "femm_group_replay")
