"""The CPU restatement of the learned 3c surrogate (oracle/mlp_oracle.py),
pinned on its own: the output slot map is a bijection onto (theta, ME), the
activation has the closed-form values and derivatives its definition fixes,
the analytic gradient agrees with central finite differences, the BF16
rounding follows round-to-nearest-even, and the split-precision inference
emulation agrees with the float64 forward pass to its stated precision."""
import numpy as np
import pytest

import mlp_oracle as mo


@pytest.mark.parametrize("nme", [1, 3, 10, 12])
def test_slot_map_is_a_bijection(nme):
    seen = {}
    for o in range(mo.NO):
        ok, th, m = mo.slot(o, nme)
        if ok:
            assert (th, m) not in seen
            seen[(th, m)] = o
            # all five theta values of one ME sit in one lane group
            assert ((o >> 2) & 3) == m % 4
    assert set(seen) == {(th, m) for th in range(5) for m in range(nme)}


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mo.NP)
    X = rng.uniform(-1, 1, (40, 3))
    T = rng.normal(0, 1, (40, 64))
    nme = 10
    loss, g = mo.loss_and_grads(p, X, T, nme)
    assert loss > 0
    idx = np.concatenate([rng.choice(np.arange(s, e), 12, replace=False) for s, e in
                          ((mo.P_W1, mo.P_B1), (mo.P_B1, mo.P_W2), (mo.P_W2, mo.P_B2), (mo.P_B2, mo.P_W3),
                           (mo.P_W3, mo.P_B3), (mo.P_B3, mo.NP))])
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.NP)
        e[i] = h
        fd = (mo.loss_and_grads(p + e, X, T, nme)[0] - mo.loss_and_grads(p - e, X, T, nme)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # padding slots take no gradient through the output layer
    mask = mo.valid_mask(nme)
    W3g = g[mo.P_W3:mo.P_B3].reshape(64, 64)
    assert np.all(W3g[~mask] == 0) and np.all(g[mo.P_B3:][~mask] == 0)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 1e-3], np.float32)
    r = mo.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even mantissa
    assert r[2] == np.float32(1.0 + 2.0 ** -7)
    assert r[3] == -2.5
    assert abs(r[4] - 1e-3) <= 1e-3 * 2.0 ** -8


def test_activation_closed_form():
    """s(z) = 1/2 + z/4 - z|z|/32 on [-4, 4]: s(0) = 1/2, s'(0) = 1/4 (the
    logistic function's), s(4) = 1, s(-4) = 0 with zero slope there, odd
    about (0, 1/2), constant outside; act(z) = z s(z) and dact its derivative."""
    assert mo.sig(0.0) == 0.5
    assert mo.sig(4.0) == 1.0 and mo.sig(-4.0) == 0.0
    assert mo.sig(7.5) == 1.0 and mo.sig(-7.5) == 0.0
    z = np.linspace(-3.9, 3.9, 79)
    assert np.allclose(mo.sig(z) + mo.sig(-z), 1.0, rtol=0, atol=1e-15)
    h = 1e-6
    for z0, slope in ((0.0, 0.25), (4.0 - 1e-4, 0.0), (-4.0 + 1e-4, 0.0)):
        fd = (mo.sig(z0 + h) - mo.sig(z0 - h)) / (2 * h)
        assert abs(fd - slope) < 2e-5, (z0, fd)
    z = np.linspace(-5, 5, 201)
    z = z[(np.abs(np.abs(z) - 4.0) > 1e-3) & (np.abs(z) > 1e-3)]
    fd = (mo.act(z + h) - mo.act(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.dact(z))) < 1e-6
    # float32 evaluation of the kernels agrees with the binary64 definition
    assert np.max(np.abs(mo.act32(z.astype(np.float32)) - mo.act(z))) < 1e-6 * np.max(np.abs(z))


def test_split_inference_emulation_close_to_float64_forward():
    """With inputs and the last layer split into BF16 high + low parts, the
    only coarse rounding left is the BF16 of layer 2 (weights and
    activations): the emulated outputs stay within a few 2^-9 relative of the
    exact forward pass."""
    rng = np.random.default_rng(8)
    p = rng.normal(0, 0.3, mo.NP)
    X = rng.uniform(-1, 1, (500, 3))
    y = mo.forward(p, X)[4]
    ye = mo.infer_bf16(p.astype(np.float32), X.astype(np.float32))
    scale = np.max(np.abs(y))
    assert np.max(np.abs(ye - y)) < 8 * 2.0 ** -9 * scale
    assert np.median(np.abs(ye - y)) < 2.0 ** -9 * scale
