"""CPU tests pinning the warm-start network oracle (oracle/mlp_oracle.py):
BF16 rounding against hand-computed ties, and the back-propagated gradient
of the exact FP64 model against central finite differences of its loss."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        one: one,
        one + 2.0 ** -8: one,                  # tie between 1 and 1+2^-7: even mantissa (1)
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,  # tie between 1+2^-7 and 1+2^-6: even is 1+2^-6
        one + 2.0 ** -8 + 2.0 ** -12: one + 2.0 ** -7,  # above the tie rounds up
        -3.0: -3.0,
        2.0 ** -130: 2.0 ** -130,              # subnormal, exactly representable
    }
    x = np.array(list(cases), dtype=np.float32)
    want = np.array(list(cases.values()), dtype=np.float32)
    assert np.array_equal(M.to_bf16(x), want)


@pytest.mark.parametrize("nin,nout", [(M.IN, M.OUT), (66, 24)])
def test_gradient_matches_finite_differences(nin, nout):
    rng = np.random.default_rng(5)
    w1, b1, w2, b2, w3, b3, end = M.offsets(nin, nout)
    p = rng.standard_normal(end) * 0.1
    X = rng.standard_normal((6, nin))
    T = rng.standard_normal((6, nout))
    loss, g = M.loss_and_grad(p, X, T, bf16=False)
    assert g.size == end
    idx = np.concatenate([rng.choice(np.arange(w1, b1), 6, replace=False),
                          rng.choice(np.arange(b1, w2), 3, replace=False),
                          rng.choice(np.arange(w2, b2), 6, replace=False),
                          rng.choice(np.arange(b2, w3), 3, replace=False),
                          rng.choice(np.arange(w3, b3), 6, replace=False),
                          np.arange(b3, end)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, X, T, bf16=False)[0] - M.loss_and_grad(pm, X, T, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_model_is_close_to_exact():
    rng = np.random.default_rng(6)
    p = rng.standard_normal(M.NPARAM) * 0.1
    X = rng.standard_normal((64, M.IN))
    T = rng.standard_normal((64, M.OUT))
    l64, g64 = M.loss_and_grad(p, X, T, bf16=False)
    l16, g16 = M.loss_and_grad(p, X, T, bf16=True)
    assert abs(l16 - l64) < 2e-2 * l64
    assert np.linalg.norm(g16 - g64) < 3e-2 * np.linalg.norm(g64)
