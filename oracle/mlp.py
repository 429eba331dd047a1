"""ORACLE — test infrastructure only: a numpy restatement of the learned warm
start's network (the GPU's kh_mlp_*), used to check its forward pass and
gradients. No part of the product imports this file.

The network is n_in -> H -> H -> n_out with ReLU; the GPU feeds every matrix
product bf16 operands (inputs, activations, weights and back-propagated
gradients rounded to nearest even) and accumulates in FP32. `forward` and
`gradients` reproduce those roundings and accumulate in float64, so the GPU
differs only by its FP32 accumulation order. With bf16=False the same code is
plain float64 back-propagation, which the tests pin against finite
differences.
"""
from __future__ import annotations

import numpy as np


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(a, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta: np.ndarray, n_in: int, h: int, n_out: int):
    """Parameter vector (kh_mlp_copy order) -> [(W1, b1), (W2, b2), (W3, b3)]."""
    shapes = [(n_in, h), (h, h), (h, n_out)]
    out, off = [], 0
    for k, n in shapes:
        W = theta[off:off + k * n].reshape(k, n).astype(np.float64)
        off += k * n
        b = theta[off:off + n].astype(np.float64)
        off += n
        out.append((W, b))
    return out


def forward(theta, x, n_in, h, n_out, use_bf16=True):
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(theta, n_in, h, n_out)
    a = r(x)
    zs, ins = [], []
    for l, (W, b) in enumerate(layers):
        ins.append(a)
        z = a @ r(W) + b
        zs.append(z)
        if l < 2:
            a = r(np.maximum(z, 0.0).astype(np.float32)) if use_bf16 else np.maximum(z, 0.0)
    return zs[-1], zs, ins


def gradients(theta, x, t, n_in, h, n_out, use_bf16=True):
    """Mean squared error over B x n_out and its gradient (kh_mlp_copy order)."""
    r = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(theta, n_in, h, n_out)
    y, zs, ins = forward(theta, x, n_in, h, n_out, use_bf16)
    B = x.shape[0]
    e = y - t
    loss = float(np.sum(e * e) / (B * n_out))
    dz = 2.0 * e / (B * n_out)
    grads = [None] * 3
    for l in (2, 1, 0):
        W, _ = layers[l]
        dzr = r(dz.astype(np.float32)) if use_bf16 else dz
        gW = ins[l].T @ dzr
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        da = dzr @ r(W).T
        dz = da * (zs[l - 1] > 0)
    flat = []
    for gW, gb in grads:
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)
