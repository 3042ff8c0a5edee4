"""ctypes binding of the learned warm start (femm_mlp_* / femm_warm_start_* in
include/femm_gpu_amd.h).

The reference has a warm-start hook (``NonlinearP1Solver::Solve(..., warm_start)``,
gpu_solver/src/solvers/nonlinear_solver.cu:667-698) and no learned part; ``WarmStart``
produces starts for that hook: POD basis of solver states + FP64 linear path + BF16 MFMA
MLP, trained on the GPU.  ``Mlp`` exposes the network alone (gradient checks, training).
"""

from __future__ import annotations

import ctypes
import sys

import numpy as np

from .model import CNewtonOptions, accepted_options
from .solver import RESULT_DTYPE, FemmError, Plan, SolveConfig, _check, library, mixed_config

P = ctypes.POINTER
c_d, c_f, c_i32, c_u64 = ctypes.c_double, ctypes.c_float, ctypes.c_int32, ctypes.c_uint64


def _dp(a):
    return a.ctypes.data_as(P(c_d))


def _fp(a):
    return a.ctypes.data_as(P(c_f))


class WarmStartTrain(ctypes.Structure):
    _fields_ = [("steps", c_i32), ("batch", c_i32), ("learning_rate", c_f),
                ("pod_iterations", c_i32), ("linear_skip", c_i32), ("reserved", c_i32),
                ("seed", c_u64)]


class WarmStartReport(ctypes.Structure):
    _fields_ = [("samples", c_i32), ("pod_iterations", c_i32),
                ("pod_energy_fraction", c_d), ("pod_leading_sigma", c_d), ("pod_last_sigma", c_d),
                ("linear_residual_fraction", c_d), ("coefficient_scale", c_d),
                ("first_loss", c_f), ("final_loss", c_f)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def train_settings(steps=4000, batch=128, learning_rate=2e-3, pod_iterations=60, linear_skip=1,
                   seed=12044) -> WarmStartTrain:
    return WarmStartTrain(steps, batch, learning_rate, pod_iterations, linear_skip, 0, seed)


class Mlp:
    """inputs -> hidden -> hidden -> outputs (SiLU), BF16 MFMA products, FP32 master weights."""

    def __init__(self, inputs, hidden, outputs, seed=1):
        self._lib = library()
        self._lib.femm_mlp_destroy.restype = None
        self._handle = ctypes.c_void_p()
        _check(self._lib.femm_mlp_create(int(inputs), int(hidden), int(outputs), c_u64(seed),
                                         ctypes.byref(self._handle)), "femm_mlp_create")
        lay = np.zeros(10, np.int64)
        _check(self._lib.femm_mlp_layout(self._handle, lay.ctypes.data_as(P(ctypes.c_int64))),
               "femm_mlp_layout")
        (self.inputs, self.hidden, self.outputs, self.in_p, self.out_p, self.count,
         self.off_b1, self.off_w2, self.off_w3, self.off_b3) = (int(v) for v in lay)

    def close(self):
        if self._handle:
            self._lib.femm_mlp_destroy(self._handle)
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

    def unpack(self, flat):
        """Split a flat parameter array into (W1, b1, W2, b2, W3, b3) views."""
        h = self.hidden
        off_b2 = self.off_w2 + h * h
        return (flat[: self.off_b1].reshape(self.in_p, h), flat[self.off_b1: self.off_w2],
                flat[self.off_w2: off_b2].reshape(h, h), flat[off_b2: self.off_w3],
                flat[self.off_w3: self.off_b3].reshape(h, self.out_p),
                flat[self.off_b3: self.count])

    def parameters(self):
        out = np.zeros(self.count, np.float32)
        _check(self._lib.femm_mlp_get_parameters(self._handle, _fp(out)), "femm_mlp_get_parameters")
        return out

    def set_parameters(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        _check(self._lib.femm_mlp_set_parameters(self._handle, _fp(flat)), "femm_mlp_set_parameters")

    def set_standardization(self, mean, std):
        mean = np.ascontiguousarray(mean, np.float64)
        std = np.ascontiguousarray(std, np.float64)
        _check(self._lib.femm_mlp_set_standardization(self._handle, _dp(mean), _dp(std)),
               "femm_mlp_set_standardization")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.inputs)
        y = np.zeros((x.shape[0], self.outputs), np.float32)
        _check(self._lib.femm_mlp_forward(self._handle, x.shape[0], _dp(x), _fp(y)),
               "femm_mlp_forward")
        return y

    def gradients(self, x, target):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.inputs)
        target = np.ascontiguousarray(target, np.float32).reshape(x.shape[0], self.outputs)
        loss = c_f()
        grads = np.zeros(self.count, np.float32)
        _check(self._lib.femm_mlp_gradients(self._handle, x.shape[0], _dp(x), _fp(target),
                                            ctypes.byref(loss), _fp(grads)), "femm_mlp_gradients")
        return float(loss.value), grads

    def train(self, x, target, steps, batch, learning_rate, seed=1):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.inputs)
        target = np.ascontiguousarray(target, np.float32).reshape(x.shape[0], self.outputs)
        hist = np.zeros(max(1, steps), np.float32)
        _check(self._lib.femm_mlp_train(self._handle, x.shape[0], _dp(x), _fp(target), int(steps),
                                        int(batch), c_f(learning_rate), c_u64(seed), _fp(hist)),
               "femm_mlp_train")
        return hist[:steps]


class WarmStart:
    """POD + linear + MLP warm-start model for one plan (the plan must outlive it)."""

    def __init__(self, plan: Plan, inputs, modes=64, hidden=256, seed=1):
        self.plan = plan
        self.inputs, self.modes = int(inputs), int(modes)
        self._lib = library()
        self._lib.femm_warm_start_destroy.restype = None
        self._handle = ctypes.c_void_p()
        _check(self._lib.femm_warm_start_create(plan.handle, self.inputs, self.modes, int(hidden),
                                                c_u64(seed), ctypes.byref(self._handle)),
               "femm_warm_start_create")

    def close(self):
        if self._handle:
            self._lib.femm_warm_start_destroy(self._handle)
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

    def fit(self, features, states, train: WarmStartTrain = None):
        features = np.ascontiguousarray(features, np.float64).reshape(-1, self.inputs)
        states = np.ascontiguousarray(states, np.float64).reshape(features.shape[0], self.plan.N)
        train = train or train_settings()
        _check(self._lib.femm_warm_start_fit(self._handle, features.shape[0], _dp(features),
                                             _dp(states), ctypes.byref(train)),
               "femm_warm_start_fit")
        return self.report()

    def fit_device(self, samples, features_ptr, states_ptr, stream_ptr, train: WarmStartTrain = None):
        train = train or train_settings()
        _check(self._lib.femm_warm_start_fit_device(
            self._handle, int(samples), ctypes.c_void_p(features_ptr), ctypes.c_void_p(states_ptr),
            ctypes.byref(train), ctypes.c_void_p(stream_ptr)), "femm_warm_start_fit_device")
        return self.report()

    def report(self):
        rep = WarmStartReport()
        _check(self._lib.femm_warm_start_report(self._handle, ctypes.byref(rep)),
               "femm_warm_start_report")
        return rep.as_dict()

    def predict(self, features):
        features = np.ascontiguousarray(features, np.float64).reshape(-1, self.inputs)
        a0 = np.zeros((features.shape[0], self.plan.N))
        _check(self._lib.femm_warm_start_predict(self._handle, features.shape[0], _dp(features),
                                                 _dp(a0)), "femm_warm_start_predict")
        return a0

    def solve_batch(self, features, currents, options: CNewtonOptions = None,
                    config: SolveConfig = None, warm_iterations=12, variant_of_item=None):
        """Learned-start solve (femm_newton_batch_learned); `variant_of_item` selects each
        item's material variant (Plan.set_variants) or None for the plan's materials."""
        features = np.ascontiguousarray(features, np.float64).reshape(-1, self.inputs)
        B = features.shape[0]
        currents = np.ascontiguousarray(currents, np.float64).reshape(B, self.plan.C)
        options = options or accepted_options()
        config = config or mixed_config()
        index = None if variant_of_item is None else \
            np.ascontiguousarray(variant_of_item, np.int32).reshape(B)
        a, flux = np.zeros((B, self.plan.N)), np.zeros((B, self.plan.C))
        results = np.zeros(B, RESULT_DTYPE)
        _check(self._lib.femm_newton_batch_learned(
            self.plan.handle, self._handle, B,
            None if index is None else index.ctypes.data_as(P(c_i32)), _dp(features),
            _dp(currents), ctypes.byref(options), ctypes.byref(config), int(warm_iterations),
            _dp(a), _dp(flux), results.ctypes.data_as(ctypes.c_void_p)),
            "femm_newton_batch_learned")
        return dict(a=a, flux=flux, results=results)

    def solve_batch_device(self, items, features_ptr, currents_ptr, a_ptr, flux_ptr, results_ptr,
                           stream_ptr, options: CNewtonOptions = None, config: SolveConfig = None,
                           warm_iterations=12, variant_ptr=None):
        options = options or accepted_options()
        config = config or mixed_config()
        _check(self._lib.femm_newton_batch_learned_device(
            self.plan.handle, self._handle, int(items),
            None if variant_ptr is None else ctypes.c_void_p(variant_ptr),
            ctypes.c_void_p(features_ptr), ctypes.c_void_p(currents_ptr), ctypes.byref(options),
            ctypes.byref(config), int(warm_iterations), ctypes.c_void_p(a_ptr),
            ctypes.c_void_p(flux_ptr), ctypes.c_void_p(results_ptr), ctypes.c_void_p(stream_ptr)),
            "femm_newton_batch_learned_device")


__all__ = ["Mlp", "WarmStart", "WarmStartTrain", "train_settings", "FemmError"]
