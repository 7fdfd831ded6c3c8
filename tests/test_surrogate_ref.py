"""The CPU reference MLP's gradients against central finite differences, so
that the GPU gradient check compares against something that is itself right."""

import numpy as np

import mlp_ref


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [7, 9, 8, 5]
    W = [rng.normal(0, 0.5, (dims[i + 1], dims[i])) for i in range(3)]
    b = [rng.normal(0, 0.1, dims[i + 1]) for i in range(3)]
    x = rng.normal(size=(6, 7))
    y = rng.normal(size=(6, 5))
    n, nout = 5, 4
    _, gW, gb = mlp_ref.loss_and_grads(W, b, x, y, n, nout)
    h = 1e-6
    for l in range(3):
        for (i, j) in [(0, 0), (1, 2), (dims[l + 1] - 1, dims[l] - 1)]:
            W[l][i, j] += h
            lp = mlp_ref.loss_and_grads(W, b, x, y, n, nout)[0]
            W[l][i, j] -= 2 * h
            lm = mlp_ref.loss_and_grads(W, b, x, y, n, nout)[0]
            W[l][i, j] += h
            assert abs((lp - lm) / (2 * h) - gW[l][i, j]) <= 1e-7 * (1 + abs(gW[l][i, j]))
        b[l][0] += h
        lp = mlp_ref.loss_and_grads(W, b, x, y, n, nout)[0]
        b[l][0] -= 2 * h
        lm = mlp_ref.loss_and_grads(W, b, x, y, n, nout)[0]
        b[l][0] += h
        assert abs((lp - lm) / (2 * h) - gb[l][0]) <= 1e-7 * (1 + abs(gb[l][0]))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -3.14159], dtype=np.float64)
    r = mlp_ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2**-7
    assert abs(r[3] + 3.14159) < 2**-7 * 4
