"""GPU: the BF16 MFMA GEMM against a float64 product of the same bfloat16
operands, and the MLP's forward/backward against the CPU oracle."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('ta,tb', [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize('M,N,K', [(64, 64, 16), (100, 70, 37), (2048, 512, 8), (48, 256, 512),
                                   # large-tile kernel: aligned, and ragged edges in all dims
                                   (2048, 2048, 96), (2050, 2049, 37)])
def test_gemm_bf16_matches_float64(ta, tb, M, N, K):
    import torch
    from snoopy_amd import _native as nat
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    Ab, Bb = mo.bf16(A), mo.bf16(B)
    ref = Ab @ Bb + bias
    dA = torch.from_numpy(A.T.copy() if ta else A).cuda().to(torch.bfloat16)
    dB = torch.from_numpy(B.T.copy() if tb else B).cuda().to(torch.bfloat16)
    C = torch.empty(M, N, dtype=torch.float32, device='cuda')
    Z = torch.empty_like(C)
    lda = M if ta else K
    ldb = K if tb else N
    dbias = torch.from_numpy(bias).cuda()
    for epi in (0, 1):
        nat.check(nat.load().sn_gemm_bf16(ta, tb, epi, M, N, K, nat.ptr(dA), lda, nat.ptr(dB), ldb,
                                          nat.ptr(dbias), None,
                                          nat.ptr(Z), nat.ptr(C), None, N, nat.stream_handle()))
        got = C.cpu().numpy()
        want = ref if epi == 0 else mo.gelu(ref)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-5, (epi, err)
        if epi == 1:
            assert np.abs(Z.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-5
    # GELU-derivative epilogue
    Zin = torch.from_numpy(rng.standard_normal((M, N)).astype(np.float32)).cuda()
    nat.check(nat.load().sn_gemm_bf16(ta, tb, 2, M, N, K, nat.ptr(dA), lda, nat.ptr(dB), ldb,
                                      None, nat.ptr(Zin), None, nat.ptr(C), None, N,
                                      nat.stream_handle()))
    want = (Ab @ Bb) * mo.gelu_grad(Zin.cpu().numpy().astype(np.float64))
    assert np.abs(C.cpu().numpy() - want).max() / np.abs(want).max() < 1e-5


@pytest.mark.parametrize('ta,tb', [(1, 0), (0, 0), (1, 1)])
@pytest.mark.parametrize('M,N,K', [(512, 512, 2048), (512, 256, 2048), (200, 130, 1037)])
def test_gemm_bf16_split_k_matches_float64(ta, tb, M, N, K):
    """Plain products with few output tiles and a long K (the network's
    weight gradients) take the split-K path: layers of K summed in order."""
    import torch
    from snoopy_amd import _native as nat
    rng = np.random.default_rng(M * N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    ref = mo.bf16(A) @ mo.bf16(B)
    dA = torch.from_numpy(A.T.copy() if ta else A).cuda().to(torch.bfloat16)
    dB = torch.from_numpy(B.T.copy() if tb else B).cuda().to(torch.bfloat16)
    C = torch.full((M, N), float('nan'), dtype=torch.float32, device='cuda')
    nat.check(nat.load().sn_gemm_bf16(ta, tb, 0, M, N, K, nat.ptr(dA), M if ta else K,
                                      nat.ptr(dB), K if tb else N, None, None, None, nat.ptr(C),
                                      None, N, nat.stream_handle()))
    got = C.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5


def test_mlp_forward_backward_match_oracle():
    import torch
    from snoopy_amd.mlp import MLP
    net = MLP((8, 512, 512, 512, 512, 256), seed=3)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (256, 8)).astype(np.float32)
    T = rng.standard_normal((256, 256)).astype(np.float32)
    Y = net.forward(torch.from_numpy(X).cuda())
    dYf, dYb, loss = net.mse_grad(Y, torch.from_numpy(T).cuda())
    gW, gb = net.backward(dYf, dYb)
    Ws = [w.cpu().numpy() for w in net.W]
    bs = [b.cpu().numpy() for b in net.b]
    Yo, cache = mo.forward(Ws, bs, X)
    lo, dYo = mo.mse(Yo, T)
    assert np.abs(Y.cpu().numpy() - Yo).max() / np.abs(Yo).max() < 2e-3
    assert abs(loss - lo) < 1e-3 * lo
    gWo, gbo = mo.backward(Ws, cache, dYo)
    for l in range(net.L):
        for g, go in ((gW[l].cpu().numpy(), gWo[l]), (gb[l].cpu().numpy(), gbo[l])):
            err = np.linalg.norm(g - go) / np.linalg.norm(go)
            assert err < 2e-2, (l, err)


def test_mlp_training_reduces_loss():
    import torch
    from snoopy_amd.mlp import MLP
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (512, 8)).astype(np.float32)
    Wt = rng.standard_normal((8, 32))
    T = np.tanh(X @ Wt).astype(np.float32)
    net = MLP((8, 128, 128, 32), seed=1)
    hist = net.train(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), epochs=300, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0], (hist[0], hist[-1])


@pytest.mark.parametrize('M,N', [(2048, 512), (2048, 256), (333, 70), (5, 3)])
def test_colsum_matches_float64(M, N):
    """The bias gradient db = column sums of dZ (vector path for N % 4 == 0,
    element path otherwise; ragged row and column counts)."""
    import torch
    from snoopy_amd import _native as nat
    rng = np.random.default_rng(M + 7 * N)
    Z = rng.standard_normal((M, N)).astype(np.float32)
    dZ = torch.from_numpy(Z).cuda()
    db = torch.full((N,), float('nan'), dtype=torch.float32, device='cuda')
    nat.check(nat.load().sn_colsum_f32(M, N, nat.ptr(dZ), nat.ptr(db), nat.stream_handle()))
    want = Z.astype(np.float64).sum(axis=0)
    got = db.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(Z).sum(axis=0).max()
