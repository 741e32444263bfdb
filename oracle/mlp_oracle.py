"""CPU restatement of the learned three-centre surrogate -- TEST INFRASTRUCTURE.

Only tests/ use this module: it checks the GPU training kernels (forward,
backward, loss) and the BF16 MFMA inference of thunder-master_amd/csrc/fb_mlp.*.
It is never called by, linked into or used as a fallback of the library.

The network (no counterpart in the reference):
    u  = (2x/xmax - 1, 2y/ymax - 1, cos theta)
    a1 = act(W1 u + b1)           W1: 64 x 3
    a2 = act(W2 a1 + b2)          W2: 64 x 64
    y  = W3 a2 + b3               W3: 64 x 64, output slots
    loss = mean over samples and valid slots of (y - t)^2
with act(z) = z s(z), s the quadratic sigmoid 1/2 + z/4 - z|z|/32 on
[-4, 4] (0 below, 1 above).
Output slot o holds (theta, ME) = slot(o, nme); the targets t are the theta
cells of getMEs_Fdata_3c (e.FDATA/M_Fdata_3c.f90:391-489) divided by a
per-slot scale.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NO = 3, 64, 64
P_W1 = 0
P_B1 = P_W1 + NH * NIN
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NO * NH
NP = P_B3 + NO


def slot(o, nme):
    """(valid, theta, me) of output slot o: lane group q = (o >> 2) & 3 holds
    the matrix elements q, q + 4, q + 8, each with its five theta values."""
    mt, q, j = o >> 4, (o >> 2) & 3, o & 3
    r = mt * 4 + j
    ml = r // 5
    th, m = r - ml * 5, q + 4 * ml
    return (ml < 3 and m < nme), th, m


def valid_mask(nme):
    return np.array([slot(o, nme)[0] for o in range(NO)])


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[P_W1:P_B1].reshape(NH, NIN), p[P_B1:P_W2], p[P_W2:P_B2].reshape(NH, NH), p[P_B2:P_W3],
            p[P_W3:P_B3].reshape(NO, NH), p[P_B3:NP])


SCLAMP = 4.0


def sig(z):
    """Quadratic sigmoid: s(0) = 1/2, s'(0) = 1/4, s(+-4) = 1 or 0 with zero slope."""
    zc = np.clip(z, -SCLAMP, SCLAMP)
    return 0.5 + zc * (0.25 - np.abs(zc) / 32.0)


def act(z):
    return z * sig(z)


def dact(z):
    ds = np.where(np.abs(z) < SCLAMP, 0.25 - np.abs(z) / 16.0, 0.0)
    return sig(z) + z * ds


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1.T + b1
    A1 = act(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = act(Z2)
    Y = A2 @ W3.T + b3
    return Z1, A1, Z2, A2, Y


def loss_and_grads(p, X, T, nme):
    """Binary64 loss and gradient (flat, same layout as the parameters)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    mask = valid_mask(nme).astype(np.float64)
    n = X.shape[0]
    cnt = n * mask.sum()
    Z1, A1, Z2, A2, Y = forward(p, X)
    D = (Y - T) * mask
    loss = float(np.sum(D * D) / cnt)
    dY = 2.0 * D / cnt
    g = np.zeros(NP)
    g[P_W3:P_B3] = (dY.T @ A2).ravel()
    g[P_B3:NP] = dY.sum(0)
    dZ2 = (dY @ W3) * dact(Z2)
    g[P_W2:P_B2] = (dZ2.T @ A1).ravel()
    g[P_B2:P_W3] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * dact(Z1)
    g[P_W1:P_B1] = (dZ1.T @ X).ravel()
    g[P_B1:P_W2] = dZ1.sum(0)
    return loss, g


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def act32(z):
    """The activation as the kernels evaluate it, in float32."""
    z = np.asarray(z, np.float32)
    zc = np.clip(z, np.float32(-SCLAMP), np.float32(SCLAMP))
    t = (np.abs(zc) * np.float32(-0.03125) + np.float32(0.25)).astype(np.float32)
    return (z * (zc * t + np.float32(0.5)).astype(np.float32)).astype(np.float32)


def split_bf16(a):
    """High and low BF16 parts of float32 values (a ~ hi + lo)."""
    hi = bf16(a)
    lo = bf16(np.asarray(a, np.float32) - hi)
    return hi, lo


def infer_bf16(p, X, oscale=None):
    """The inference arithmetic of the MFMA path (csrc/fb_mlp.h mlp_eval16):
    layer 1 with inputs and weights split into BF16 high and low parts
    (uh.wh + ul.wh + uh.wl), layer 2 with BF16 weights and activations,
    layer 3 with the slot scale folded into weights and bias and both sides
    split (hh.wh + hl.wh + hh.wl); FP32 accumulation from the FP32 bias.
    Products of BF16 values are exact in binary64, so each layer's sum is
    formed in binary64 and rounded once to FP32; the kernel's FP32 MFMA sum
    can differ from it in the last bits."""
    p = np.asarray(p, np.float32)
    osc = np.ones(NO, np.float32) if oscale is None else np.asarray(oscale, np.float32)
    W1 = p[P_W1:P_B1].reshape(NH, NIN)
    b1, b2 = p[P_B1:P_W2], p[P_B2:P_W3]
    W2 = bf16(p[P_W2:P_B2].reshape(NH, NH)).astype(np.float64)
    W3 = (p[P_W3:P_B3].reshape(NO, NH) * osc[:, None]).astype(np.float32)
    b3 = (p[P_B3:NP] * osc).astype(np.float32)
    W3h, W3l = split_bf16(W3)
    X = np.asarray(X, np.float32)
    uh, ul = split_bf16(X)
    w1h, w1l = split_bf16(W1)
    uh, ul, w1h, w1l = (a.astype(np.float64) for a in (uh, ul, w1h, w1l))
    z1 = (uh @ w1h.T + ul @ w1h.T + uh @ w1l.T + b1).astype(np.float32)
    a1 = bf16(act32(z1)).astype(np.float64)
    z2 = (a1 @ W2.T + b2).astype(np.float32)
    hh, hl = split_bf16(act32(z2))
    hh, hl = hh.astype(np.float64), hl.astype(np.float64)
    W3h, W3l = W3h.astype(np.float64), W3l.astype(np.float64)
    return hh @ W3h.T + hl @ W3h.T + hh @ W3l.T + b3
