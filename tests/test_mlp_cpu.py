"""The CPU restatement of the learned 3c surrogate (oracle/mlp_oracle.py),
pinned on its own: the output slot map is a bijection onto (theta, ME), the
analytic gradient agrees with central finite differences, and the BF16
rounding follows round-to-nearest-even."""
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
