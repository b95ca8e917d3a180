"""CPU: the MLP oracle's backward is the derivative of its forward (central
finite differences of the unrounded float64 network), and its bfloat16
rounding matches torch's."""
import numpy as np

from oracle import mlp_oracle as mo


def _net(seed=0, sizes=(8, 16, 16, 5)):
    rng = np.random.default_rng(seed)
    Ws = [rng.standard_normal((a, b)) * (1.0 / np.sqrt(a)) for a, b in zip(sizes[:-1], sizes[1:])]
    bs = [rng.standard_normal(b) * 0.1 for b in sizes[1:]]
    return Ws, bs


def test_backward_matches_finite_differences():
    Ws, bs = _net()
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (7, 8))
    G = rng.standard_normal((7, 5))
    Y, cache = mo.forward(Ws, bs, X, rounding=False)
    gW, gb = mo.backward(Ws, cache, G, rounding=False)
    eps = 1e-6
    for l in range(len(Ws)):
        for (i, j) in [(0, 0), (3, 2), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (np.sum(mo.forward(Wp, bs, X, False)[0] * G) -
                  np.sum(mo.forward(Wm, bs, X, False)[0] * G)) / (2 * eps)
            assert abs(fd - gW[l][i, j]) < 1e-6 * max(1.0, abs(fd))
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][1] += eps
        bm[l][1] -= eps
        fd = (np.sum(mo.forward(Ws, bp, X, False)[0] * G) -
              np.sum(mo.forward(Ws, bm, X, False)[0] * G)) / (2 * eps)
        assert abs(fd - gb[l][1]) < 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(2).standard_normal(10000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)


def test_gelu_grad_is_derivative():
    z = np.linspace(-5, 5, 101)
    fd = (mo.gelu(z + 1e-6) - mo.gelu(z - 1e-6)) / 2e-6
    np.testing.assert_allclose(mo.gelu_grad(z), fd, atol=1e-8)
