"""The FP64 network oracle, pinned without the device: bfloat16 rounding against known answers
and the analytic gradient against central finite differences (a dropped term, a wrong mask or a
transposed product fails the difference check)."""
import numpy as np
import pytest

import pymlp

DIMS = (7, 9, 6, 5)


def _params(rng, dims=DIMS):
    d_in, h1, h2, d_out = dims
    n = h1 * d_in + h1 + h2 * h1 + h2 + d_out * h2 + d_out
    return rng.normal(0, 0.7, n)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3, 65504.0,
                  1.0 + 2 ** -8 + 2 ** -20], np.float32)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6; just above a tie rounds up
    want = [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 197 * 2.0 ** -16, 65536.0, 1.0 + 2 ** -7]
    np.testing.assert_array_equal(pymlp.bf16(x), np.array(want))


def test_parameter_layout_round_trip():
    rng = np.random.default_rng(1)
    p = _params(rng)
    W1, b1, W2, b2, W3, b3 = pymlp.split(p, DIMS)
    assert W1.shape == (9, 7) and W2.shape == (6, 9) and W3.shape == (5, 6)
    assert np.array_equal(np.concatenate([a.ravel() for a in (W1, b1, W2, b2, W3, b3)]), p)


def test_forward_is_the_composition():
    rng = np.random.default_rng(2)
    p = _params(rng)
    u = rng.normal(size=(4, 7))
    W1, b1, W2, b2, W3, b3 = pymlp.split(p, DIMS)
    want = np.maximum(np.maximum(u @ W1.T + b1, 0) @ W2.T + b2, 0) @ W3.T + b3
    np.testing.assert_allclose(pymlp.forward(p, DIMS, u, bf=False)[2], want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = _params(rng)
    u = rng.normal(size=(6, 7))
    t = rng.normal(size=(6, 5))

    def loss(q):
        y = pymlp.forward(q, DIMS, u, bf=False)[2]
        return ((y - t) ** 2).sum() / t.size

    g, l0 = pymlp.gradients(p, DIMS, u, t, bf=False)
    assert abs(l0 - loss(p)) < 1e-14 * max(1, l0)
    h = 1e-6
    fd = np.array([(loss(p + h * e) - loss(p - h * e)) / (2 * h) for e in np.eye(p.size)])
    # ReLU kinks are measure-zero for these draws; central differences are O(h^2)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_bf16_gradient_stays_close_to_exact():
    rng = np.random.default_rng(6)
    p = _params(rng)
    u = rng.normal(size=(16, 7))
    t = rng.normal(size=(16, 5))
    g64, _ = pymlp.gradients(p, DIMS, u, t, bf=False)
    g16, _ = pymlp.gradients(p, DIMS, u, t, bf=True)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < rel < 3e-2      # bfloat16 operands: visible, but a few parts in a thousand
