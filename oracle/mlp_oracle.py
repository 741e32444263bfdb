"""CPU restatement of the learned three-centre surrogate -- TEST INFRASTRUCTURE.

Only tests/ use this module: it checks the GPU training kernels (forward,
backward, loss) and the BF16 MFMA inference of thunder-master_amd/csrc/fb_mlp.*.
It is never called by, linked into or used as a fallback of the library.

The network (no counterpart in the reference):
    u  = (2x/xmax - 1, 2y/ymax - 1, cos theta)
    a1 = silu(W1 u + b1)          W1: 64 x 3
    a2 = silu(W2 a1 + b2)         W2: 64 x 64
    y  = W3 a2 + b3               W3: 64 x 64, output slots
    loss = mean over samples and valid slots of (y - t)^2
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


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1.T + b1
    A1 = silu(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = silu(Z2)
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
    dZ2 = (dY @ W3) * dsilu(Z2)
    g[P_W2:P_B2] = (dZ2.T @ A1).ravel()
    g[P_B2:P_W3] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * dsilu(Z1)
    g[P_W1:P_B1] = (dZ1.T @ X).ravel()
    g[P_B1:P_W2] = dZ1.sum(0)
    return loss, g


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def infer_bf16(p, X, oscale=None):
    """The inference arithmetic of the MFMA path: FP32 first layer, BF16
    weights and activations into FP32-accumulated products, FP32 biases."""
    p = np.asarray(p, np.float32)
    W1 = p[P_W1:P_B1].reshape(NH, NIN)
    b1, b2, b3 = p[P_B1:P_W2], p[P_B2:P_W3], p[P_B3:NP]
    W2 = bf16(p[P_W2:P_B2].reshape(NH, NH)).astype(np.float64)
    W3 = bf16(p[P_W3:P_B3].reshape(NO, NH)).astype(np.float64)
    X = np.asarray(X, np.float32)
    z1 = (X[:, 0:1] * W1[:, 0] + X[:, 1:2] * W1[:, 1] + X[:, 2:3] * W1[:, 2] + b1).astype(np.float32)
    a1 = bf16((z1 / (1.0 + np.exp(-z1.astype(np.float64)))).astype(np.float32)).astype(np.float64)
    z2 = (a1 @ W2.T + b2).astype(np.float32)
    a2 = bf16((z2 / (1.0 + np.exp(-z2.astype(np.float64)))).astype(np.float32)).astype(np.float64)
    y = a2 @ W3.T + b3
    if oscale is not None:
        y = y * np.asarray(oscale, np.float64)
    return y
