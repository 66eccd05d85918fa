"""CPU restatement of the detector-intensity surrogate network (include/jms_nn.h)
for checks only: the device network has no counterpart in the reference, so
its forward pass and gradients are checked against this float64 evaluation.

Test infrastructure: only tests/ may use it.  With quantize=True the values
are rounded to bf16 at the points where the device stores bf16 (input,
weights' compute copies, every ReLU activation); with quantize=False the
network is evaluated in plain float64, which finite differences pin.  With
round_dy=True the gradient at each convolution's output is also rounded to
bf16 before it meets the weights and the layer input, as the matrix-core
kernels take it (the bias gradient sums the unrounded values)."""
from __future__ import annotations

import numpy as np

C0, H0, CH, HID, NOUT = 8, 32, 32, 64, 3
STRIDES = (1, 2, 2, 2)
CIN = (C0, CH, CH, CH)
FLAT = CH * (H0 // 8) ** 2


def layout():
    """[(name, shape)] of the parameter vector, in order."""
    out = []
    for l in range(4):
        out += [(f"conv{l + 1}.w", (CH, CIN[l], 3, 3)), (f"conv{l + 1}.b", (CH,))]
    out += [("dense1.w", (HID, FLAT)), ("dense1.b", (HID,)), ("dense2.w", (NOUT, HID)), ("dense2.b", (NOUT,))]
    return out


def num_params():
    return sum(int(np.prod(s)) for _, s in layout())


def unpack(p):
    d, o = {}, 0
    for name, shape in layout():
        k = int(np.prod(shape))
        d[name] = np.asarray(p[o:o + k], np.float64).reshape(shape)
        o += k
    return d


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in layout()])


def bf16(a):
    """float32 -> bf16 (round to nearest even) -> float64"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def bf16_bits(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _conv(x, w, b, s):
    n, ci, h, _ = x.shape
    ho = h // s
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    y = np.zeros((n, w.shape[0], ho, ho)) + b[None, :, None, None]
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + s * ho:s, kx:kx + s * ho:s]
            y += np.einsum("nchw,oc->nohw", patch, w[:, :, ky, kx])
    return y


def _conv_back(x, w, dy, s):
    n, ci, h, _ = x.shape
    ho = dy.shape[2]
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + s * ho:s, kx:kx + s * ho:s]
            dw[:, :, ky, kx] = np.einsum("nohw,nchw->oc", dy, patch)
            dxp[:, :, ky:ky + s * ho:s, kx:kx + s * ho:s] += np.einsum("nohw,oc->nchw", dy, w[:, :, ky, kx])
    return dw, dy.sum(axis=(0, 2, 3)), dxp[:, :, 1:-1, 1:-1]


def forward(p, x, quantize=True):
    """outputs [n, 3] and the activations"""
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    acts = [q(x)]
    a = acts[0]
    for l in range(4):
        a = q(np.maximum(_conv(a, q(P[f"conv{l + 1}.w"]), P[f"conv{l + 1}.b"], STRIDES[l]), 0.0))
        acts.append(a)
    f = a.reshape(a.shape[0], -1)
    h = q(np.maximum(f @ q(P["dense1.w"]).T + P["dense1.b"], 0.0))
    acts.append(h)
    y = h @ q(P["dense2.w"]).T + P["dense2.b"]
    return y, acts


def loss_and_grads(p, x, t, scale, quantize=True, round_dy=False):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    y, acts = forward(p, x, quantize)
    n = x.shape[0]
    sc = np.asarray(scale, np.float64)
    r = (y - t) / sc
    loss = np.mean(r * r)
    dy = 2 * r / sc / (n * NOUT)
    G = {}
    h = acts[5]
    G["dense2.w"] = dy.T @ h
    G["dense2.b"] = dy.sum(0)
    dh = (dy @ q(P["dense2.w"])) * (h > 0)
    f = acts[4].reshape(n, -1)
    G["dense1.w"] = dh.T @ f
    G["dense1.b"] = dh.sum(0)
    da = ((dh @ q(P["dense1.w"])) * (f > 0)).reshape(acts[4].shape)
    for l in range(3, -1, -1):
        dw, _, dx = _conv_back(acts[l], q(P[f"conv{l + 1}.w"]), bf16(da) if round_dy else da, STRIDES[l])
        G[f"conv{l + 1}.w"], G[f"conv{l + 1}.b"] = dw, da.sum(axis=(0, 2, 3))
        if l > 0:
            da = dx * (acts[l] > 0)
    return loss, pack(G)
