"""Multilayer perceptron on the BF16 matrix cores (training and inference).

Layers: sizes[0] -> sizes[1] -> ... -> sizes[-1], GELU after every hidden
layer, linear output.  Parameters are kept in FP32 (master copy, Adam state)
and in BF16 (the GEMM operands); every product runs through ``sn_gemm_bf16``
(v_mfma_f32_32x32x16_bf16, FP32 accumulation) with the bias, GELU and
GELU-derivative applied in the GEMM epilogue.  Activations are stored in
BF16, pre-activations in FP32.

Shapes: W_l [K_in, N_out] row-major, b_l [N_out]; a batch is [B, features].
"""
import ctypes

import numpy as np

from . import _native as nat
from .device import device


class MLP:

    def __init__(self, sizes=(8, 512, 512, 512, 512, 256), seed=0):
        import torch
        self.sizes = tuple(int(s) for s in sizes)
        self.L = len(self.sizes) - 1
        g = torch.Generator().manual_seed(seed)
        dev = device()
        self.W, self.b = [], []
        for l in range(self.L):
            k, n = self.sizes[l], self.sizes[l + 1]
            std = (2.0 / k) ** 0.5 if l < self.L - 1 else (1.0 / k) ** 0.5
            self.W.append((torch.randn(k, n, generator=g) * std).to(dev))
            self.b.append(torch.zeros(n).to(dev))
        self.Wb = [w.to(torch.bfloat16) for w in self.W]
        self._adam = None
        self._act = {}
        self.step_count = 0

    # ------------------------------------------------------------ helpers
    def _buffers(self, B):
        import torch
        if B in self._act:
            return self._act[B]
        dev = device()
        bufs = {'A': [], 'Z': [], 'dZf': [], 'dZb': []}
        for l in range(self.L):
            n = self.sizes[l + 1]
            bufs['A'].append(torch.empty(B, n, dtype=torch.bfloat16, device=dev))
            bufs['Z'].append(torch.empty(B, n, dtype=torch.float32, device=dev))
            bufs['dZf'].append(torch.empty(B, n, dtype=torch.float32, device=dev))
            bufs['dZb'].append(torch.empty(B, n, dtype=torch.bfloat16, device=dev))
        bufs['Y'] = torch.empty(B, self.sizes[-1], dtype=torch.float32, device=dev)
        self._act[B] = bufs
        return bufs

    @staticmethod
    def _gemm(ta, tb, epi, M, N, K, A, lda, Bm, ldb, bias=None, Zin=None, Zout=None, Cf=None,
              Cb=None, ldc=None):
        nat.check(nat.load().sn_gemm_bf16(ta, tb, epi, M, N, K, nat.ptr(A), lda, nat.ptr(Bm), ldb,
                                          nat.ptr(bias), nat.ptr(Zin), nat.ptr(Zout), nat.ptr(Cf),
                                          nat.ptr(Cb), ldc if ldc is not None else N,
                                          nat.stream_handle()))

    # ------------------------------------------------------------ passes
    def forward(self, X):
        """X: [B, sizes[0]] (any float dtype, device) -> Y [B, sizes[-1]] float32."""
        import torch
        X = X.to(device=device(), dtype=torch.bfloat16).contiguous()
        B = X.shape[0]
        bufs = self._buffers(B)
        self._X = X
        a = X
        for l in range(self.L):
            k, n = self.sizes[l], self.sizes[l + 1]
            if l < self.L - 1:
                self._gemm(0, 0, 1, B, n, k, a, k, self.Wb[l], n, bias=self.b[l],
                           Zout=bufs['Z'][l], Cb=bufs['A'][l])
                a = bufs['A'][l]
            else:
                self._gemm(0, 0, 0, B, n, k, a, k, self.Wb[l], n, bias=self.b[l], Cf=bufs['Y'])
        return bufs['Y']

    def backward(self, dYf, dYb):
        """Gradients of the last forward pass for dL/dY (float32 and bfloat16)."""
        import torch
        B = dYf.shape[0]
        bufs = self._buffers(B)
        self.gW = [torch.empty_like(w) for w in self.W] if not hasattr(self, 'gW') else self.gW
        self.gb = [torch.empty_like(b) for b in self.b] if not hasattr(self, 'gb') else self.gb
        dzf, dzb = dYf, dYb
        for l in reversed(range(self.L)):
            k, n = self.sizes[l], self.sizes[l + 1]
            a_prev = self._X if l == 0 else bufs['A'][l - 1]
            # dW_l = A_{l-1}^T dZ_l, db_l = column sums of dZ_l
            self._gemm(1, 0, 0, k, n, B, a_prev, k, dzb, n, Cf=self.gW[l])
            nat.check(nat.load().sn_colsum_f32(B, n, nat.ptr(dzf), nat.ptr(self.gb[l]),
                                               nat.stream_handle()))
            if l > 0:
                # dZ_{l-1} = (dZ_l W_l^T) * GELU'(Z_{l-1})
                self._gemm(0, 1, 2, B, k, n, dzb, n, self.Wb[l], n, Zin=bufs['Z'][l - 1],
                           Cf=bufs['dZf'][l - 1], Cb=bufs['dZb'][l - 1])
                dzf, dzb = bufs['dZf'][l - 1], bufs['dZb'][l - 1]
        return self.gW, self.gb

    def mse_grad(self, Y, T):
        """dL/dY of L = mean((Y - T)^2) and L (host float)."""
        import torch
        dYf = torch.empty_like(Y)
        dYb = torch.empty(Y.shape, dtype=torch.bfloat16, device=Y.device)
        loss = ctypes.c_double()
        nat.check(nat.load().sn_mse_grad(int(Y.numel()), nat.ptr(Y), nat.ptr(T),
                                         float(1.0 / Y.numel()), nat.ptr(dYf), nat.ptr(dYb),
                                         ctypes.byref(loss), nat.stream_handle()))
        return dYf, dYb, loss.value

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        import torch
        if self._adam is None:
            self._adam = [(torch.zeros_like(p), torch.zeros_like(p))
                          for p in self.W + self.b]
        self.step_count += 1
        lib = nat.load()
        params = self.W + self.b
        grads = self.gW + self.gb
        copies = self.Wb + [None] * len(self.b)
        for p, g, (m1, m2), pb in zip(params, grads, self._adam, copies):
            nat.check(lib.sn_adam_f32(int(p.numel()), nat.ptr(p), nat.ptr(g), nat.ptr(m1),
                                      nat.ptr(m2), nat.ptr(pb), lr, b1, b2, eps, self.step_count,
                                      wd, nat.stream_handle()))

    def train(self, X, T, epochs=2000, lr=2e-3, lr_min=1e-5, batch=None, seed=0, log=None):
        """Adam on the MSE, cosine-decayed learning rate; returns the loss history."""
        import torch
        n = X.shape[0]
        batch = batch or n
        gen = torch.Generator().manual_seed(seed)
        hist = []
        for ep in range(epochs):
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * ep / max(epochs - 1, 1)))
            perm = torch.randperm(n, generator=gen).to(X.device) if batch < n else None
            for s in range(0, n, batch):
                idx = perm[s:s + batch] if perm is not None else None
                xb = X[idx] if idx is not None else X
                tb = T[idx] if idx is not None else T
                Y = self.forward(xb)
                dYf, dYb, loss = self.mse_grad(Y, tb)
                self.backward(dYf, dYb)
                self.adam_step(lr=float(lr_t))
            hist.append(loss)
            if log is not None and (ep % max(epochs // 10, 1) == 0 or ep == epochs - 1):
                log('  mlp epoch %d loss %.4e lr %.2e' % (ep, loss, lr_t))
        return hist

    def state(self):
        return {'sizes': np.array(self.sizes),
                **{'W%d' % l: w.cpu().numpy() for l, w in enumerate(self.W)},
                **{'b%d' % l: b.cpu().numpy() for l, b in enumerate(self.b)}}

    @classmethod
    def from_state(cls, st):
        import torch
        m = cls(tuple(int(s) for s in st['sizes']))
        for l in range(m.L):
            m.W[l] = torch.from_numpy(np.asarray(st['W%d' % l], dtype=np.float32)).to(device())
            m.b[l] = torch.from_numpy(np.asarray(st['b%d' % l], dtype=np.float32)).to(device())
        m.Wb = [w.to(torch.bfloat16) for w in m.W]
        return m
