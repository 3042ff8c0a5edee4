"""CPU restatement of the warm-start MLP (test infrastructure only).

The reference has no learned part (its warm-start hook is
gpu_solver/src/solvers/nonlinear_solver.cu:667-698), so nothing here follows a
reference file: this is the checker for the library's own BF16 MFMA forward and
backward kernels.  ``quantized=True`` rounds to BF16 at exactly the points the device
does (standardised inputs, weights, hidden activations, output and hidden gradients);
``quantized=False`` is the same network in plain FP64, whose gradient the tests pin
against central finite differences of the FP64 loss.

Network: z1 = x W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2, h2 = silu(z2), y = h2 W3 + b3,
loss = mean over rows x outputs of (y - t)^2.  Parameters use the flat layout of
femm_mlp_layout (padded to multiples of 16; padded entries are zero).
"""

from __future__ import annotations

import numpy as np


def bf16(a):
    """Round-to-nearest-even FP32 -> BF16, returned widened to float32."""
    f = np.ascontiguousarray(a, np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return bits.astype(np.uint32).view(np.float32)


def _sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


class Layout:
    def __init__(self, inputs, hidden, outputs):
        self.inputs, self.hidden, self.outputs = inputs, hidden, outputs
        self.in_p = (inputs + 15) // 16 * 16
        self.out_p = (outputs + 15) // 16 * 16
        h = hidden
        self.w1 = 0
        self.b1 = self.in_p * h
        self.w2 = self.b1 + h
        self.b2 = self.w2 + h * h
        self.w3 = self.b2 + h
        self.b3 = self.w3 + h * self.out_p
        self.count = self.b3 + self.out_p

    def split(self, flat):
        h = self.hidden
        return (flat[self.w1:self.b1].reshape(self.in_p, h), flat[self.b1:self.w2],
                flat[self.w2:self.b2].reshape(h, h), flat[self.b2:self.w3],
                flat[self.w3:self.b3].reshape(h, self.out_p), flat[self.b3:self.count])


def standardise(x, mean, std, layout, quantized):
    xn = (np.asarray(x, np.float64) - mean) * (1.0 / np.asarray(std, np.float64))
    pad = np.zeros((xn.shape[0], layout.in_p))
    pad[:, : layout.inputs] = xn
    return bf16(pad.astype(np.float32)).astype(np.float64) if quantized else pad


def loss_and_gradients(flat, x, target, mean, std, layout: Layout, quantized=True):
    """Returns (loss, flat gradient) with the device's rounding points when quantized."""
    q = (lambda a: bf16(a).astype(np.float64)) if quantized else (lambda a: np.asarray(a, np.float64))
    flat = np.asarray(flat, np.float64)
    W1, b1, W2, b2, W3, b3 = layout.split(flat)
    W1q, W2q, W3q = q(W1), q(W2), q(W3)
    X = standardise(x, mean, std, layout, quantized)
    rows = X.shape[0]
    z1 = X @ W1q + b1
    h1 = q(z1 * _sigmoid(z1))
    z2 = h1 @ W2q + b2
    h2 = q(z2 * _sigmoid(z2))
    y = (h2 @ W3q + b3)[:, : layout.outputs]
    d = y - np.asarray(target, np.float64)
    loss = float(np.mean(d * d))
    dy = np.zeros((rows, layout.out_p))
    dy[:, : layout.outputs] = 2.0 * d / (rows * layout.outputs)
    dyq = q(dy)
    gW3 = h2.T @ dyq
    gb3 = dy.sum(0)
    dh2 = dyq @ W3q.T
    s2 = _sigmoid(z2)
    dz2 = dh2 * (s2 * (1.0 + z2 * (1.0 - s2)))
    dz2q = q(dz2)
    gW2 = h1.T @ dz2q
    gb2 = dz2.sum(0)
    dh1 = dz2q @ W2q.T
    s1 = _sigmoid(z1)
    dz1 = dh1 * (s1 * (1.0 + z1 * (1.0 - s1)))
    dz1q = q(dz1)
    gW1 = X.T @ dz1q
    gb1 = dz1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    # padded rows / columns of the weights carry no gradient on the device
    mask = np.zeros(layout.count, bool)
    mW1, _, _, _, mW3, mb3 = layout.split(mask)
    mW1[: layout.inputs, :] = True
    mW3[:, : layout.outputs] = True
    mb3[: layout.outputs] = True
    mask[layout.b1:layout.w2] = True
    mask[layout.w2:layout.w3] = True
    return loss, np.where(mask, grad, 0.0)


def forward(flat, x, mean, std, layout: Layout, quantized=True):
    q = (lambda a: bf16(a).astype(np.float64)) if quantized else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = layout.split(np.asarray(flat, np.float64))
    X = standardise(x, mean, std, layout, quantized)
    z1 = X @ q(W1) + b1
    h1 = q(z1 * _sigmoid(z1))
    z2 = h1 @ q(W2) + b2
    h2 = q(z2 * _sigmoid(z2))
    return (h2 @ q(W3) + b3)[:, : layout.outputs]
