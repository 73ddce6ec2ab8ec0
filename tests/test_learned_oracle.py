"""Pins of the model oracle (CPU only)."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounds_to_nearest_even():
    f = lambda bits: np.array([bits], np.uint32).view(np.float32)
    assert mo.bf16(1.0)[()] == 1.0
    # exactly halfway between 1.0 and 1.0078125 -> even (1.0)
    assert mo.bf16(f(0x3F808000))[0] == 1.0
    # halfway between 1.0078125 and 1.015625 -> even (1.015625)
    assert mo.bf16(f(0x3F818000))[0] == 1.015625
    # just above halfway rounds up
    assert mo.bf16(f(0x3F808001))[0] == 1.0078125
    b = mo.bf16_bits(np.array([1.0, -2.5, 3.0e-3], np.float32))
    assert np.allclose(mo.from_bits(b), mo.bf16(np.array([1.0, -2.5, 3e-3])))


def test_zero_model_loss_is_log2_of_the_weighted_slots():
    rng = np.random.default_rng(0)
    n = 10
    X = rng.normal(size=(n, mo.IN))
    Y = (rng.random((n, mo.OUT)) < 0.3).astype(float)
    M = (rng.random((n, mo.OUT)) < 0.9).astype(float)
    loss, _ = mo.loss_and_grad(np.zeros(mo.NPARAM), X, Y, M, 3.0, False)
    w = np.where(Y > 0.5, 3.0, 1.0)
    assert np.isclose(loss, np.log(2) * (M * w).sum() / (n * mo.OUT))


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    n = 6
    params = mo.xavier(2)
    X = rng.normal(size=(n, mo.IN))
    Y = (rng.random((n, mo.OUT)) < 0.3).astype(float)
    M = (rng.random((n, mo.OUT)) < 0.9).astype(float)
    _, g = mo.loss_and_grad(params, X, Y, M, 4.0, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False)
                          for a, b in ((0, 6144), (6144, 6272),
                                       (6272, 22656), (22656, 22784),
                                       (22784, 26880), (26880, 26912))])
    eps = 1e-6
    for i in idx:
        p1, p0 = params.copy(), params.copy()
        p1[i] += eps
        p0[i] -= eps
        fd = (mo.loss_only(p1, X, Y, M, 4.0) -
              mo.loss_only(p0, X, Y, M, 4.0)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, i


def test_forward_closed_form():
    # one path through the network: x0 = 2 -> h1[0] = relu(3*2) = 6 ->
    # h2[1] = relu(0.5*6 + 1) = 4 (W2[0][1]; W2 is [in][out]) -> z[5] =
    # -1*4 + 0.25; a negative pre-activation is cut by the ReLU (h1[2])
    d = {n: np.zeros(s) for n, s in mo.SHAPES}
    d["W1"][0, 0] = 3.0
    d["W1"][0, 2] = -1.0
    d["W2"][0, 1] = 0.5
    d["W2"][2, 1] = 100.0
    d["b2"][1] = 1.0
    d["W3"][1, 5] = -1.0
    d["b3"][5] = 0.25
    d["b3"][7] = -2.0
    X = np.zeros((2, mo.IN))
    X[0, 0] = 2.0
    z = mo.forward(mo.pack(d), X)
    want = np.zeros((2, mo.OUT))
    want[:, 7] = -2.0
    want[0, 5] = -4.0 + 0.25
    want[1, 5] = -1.0 + 0.25  # h2[1] = relu(0 + 1) = 1 for the zero row
    assert np.array_equal(z, want)
    # bfloat16 rounding of the activations: x0 = 2.015625 (a bf16 value)
    # gives 3 * x0 = 6.046875, which needs 8 fraction bits and rounds
    # (half to even) to 6.0625; unrounded the logit differs
    X[0, 0] = 2.015625
    z = mo.forward(mo.pack(d), X)
    assert z[0, 5] == -(0.5 * 6.0625 + 1.0) + 0.25
    z = mo.forward(mo.pack(d), X, rounding=False)
    assert z[0, 5] == -(0.5 * 6.046875 + 1.0) + 0.25
