"""Learned warm start for the Jacobi pressure solve: a 769 -> 512 -> 512 -> 256
GELU MLP trained and run on the GPU with the library's own BF16-MFMA GEMM and
elementwise kernels (FP32 master weights, FP32 accumulation).  torch tensors
only hold device memory.  No reference counterpart.

Inputs of one predictor-pass pressure solve (built on the GPU by the solver
kernel, see warmstart.py): the 256-section displacement d and velocity v of the
Euler predictor, each divided by its RMS over the sections; the residual
r_lin = Y - A p_lin of the extrapolated pressure p_lin = p + (dt / dt_prev)(p - p_prev),
divided by its RMS; and the input sample s divided by RMS(p)  (256+256+256+1).
The last pressure enters through p_lin and r_lin.  The output y is a
correction: the guess is p_lin + RMS(r_lin) c_out S(y), S = [1,2,1]/4, and the
training target is (A^-1 Y - p_lin) / (RMS(r_lin) c_out).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import check, lib, ptr

SIZES = (769, 512, 512, 256)
N_IN, N_OUT = SIZES[0], SIZES[-1]


class WarmStartMLP:
    def __init__(self, seed: int = 0, device: str = "cuda", params=None):
        import torch
        self.torch = torch
        self.dev = device
        self.L = lib()
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for i, (kin, kout) in enumerate(zip(SIZES[:-1], SIZES[1:])):
            if params is not None:
                w, b = params[2 * i], params[2 * i + 1]
            else:
                w = (rng.standard_normal((kin, kout)) * math.sqrt(2.0 / kin)).astype(np.float32)
                if i == len(SIZES) - 2:
                    w *= 0.1
                b = np.zeros(kout, np.float32)
            self.W.append(torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(device))
            self.b.append(torch.from_numpy(np.ascontiguousarray(b, np.float32)).to(device))
        self.Wbf = [torch.empty(w.shape, dtype=torch.bfloat16, device=device) for w in self.W]
        self._refresh_bf16()
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.t = 0
        self._cap = 0

    # ---------------------------------------------------------------- utils
    def _stream(self):
        lib().cm_set_stream(C.c_void_p(self.torch.cuda.current_stream().cuda_stream))

    def _refresh_bf16(self):
        for w, wb in zip(self.W, self.Wbf):
            check(self.L.cm_to_bf16(ptr(w), ptr(wb), w.numel()), "cm_to_bf16")

    def _ensure(self, B):
        if B <= self._cap:
            return
        t, d = self.torch, self.dev
        self._cap = B
        self.Xbf = t.empty((B, SIZES[0]), dtype=t.bfloat16, device=d)
        self.Z = [t.empty((B, n), dtype=t.float32, device=d) for n in SIZES[1:]]
        self.Gbf = [t.empty((B, n), dtype=t.bfloat16, device=d) for n in SIZES[1:-1]]
        self.dY = t.empty((B, SIZES[-1]), dtype=t.float32, device=d)
        self.dYbf = t.empty((B, SIZES[-1]), dtype=t.bfloat16, device=d)
        self.dG = [t.empty((B, n), dtype=t.float32, device=d) for n in SIZES[1:-1]]
        self.dZ = [t.empty((B, n), dtype=t.float32, device=d) for n in SIZES[1:-1]]
        self.dZbf = [t.empty((B, n), dtype=t.bfloat16, device=d) for n in SIZES[1:-1]]
        self.rowloss = t.empty(B, dtype=t.float32, device=d)
        self.gW = [t.empty_like(w) for w in self.W]
        self.gb = [t.empty_like(b) for b in self.b]

    def params_numpy(self):
        out = []
        for w, b in zip(self.W, self.b):
            out += [w.cpu().numpy(), b.cpu().numpy()]
        return out

    # --------------------------------------------------------- fwd / bwd
    def forward(self, X):
        """X: float32 (B, 769) device tensor -> (B, 256) float32."""
        B = X.shape[0]
        self._ensure(B)
        self._stream()
        L = self.L
        check(L.cm_to_bf16(ptr(X), ptr(self.Xbf), X.numel()), "to_bf16")
        inp = self.Xbf
        for i in range(3):
            kin, kout = SIZES[i], SIZES[i + 1]
            check(L.cm_gemm_bf16(0, 0, B, kout, kin, ptr(inp), kin, ptr(self.Wbf[i]), kout, ptr(self.Z[i]), kout,
                                 ptr(self.b[i]), 0.0), "gemm")
            if i < 2:
                check(L.cm_gelu_fwd(ptr(self.Z[i]), ptr(self.Gbf[i]), None, B * kout), "gelu")
                inp = self.Gbf[i]
        return self.Z[2][:B]

    def backward(self, T, B):
        """MSE against T (B, 256); fills gradients, returns the loss."""
        L = self.L
        check(L.cm_mse_grad(ptr(self.Z[2]), ptr(T), ptr(self.dY), ptr(self.dYbf), ptr(self.rowloss), B, N_OUT), "mse")
        dcur_bf, dcur_f = self.dYbf, self.dY
        for i in (2, 1, 0):
            kin, kout = SIZES[i], SIZES[i + 1]
            a_in = self.Xbf if i == 0 else self.Gbf[i - 1]
            # dW = a_in^T dcur, db = colsum(dcur)
            check(L.cm_gemm_bf16(1, 0, kin, kout, B, ptr(a_in), kin, ptr(dcur_bf), kout, ptr(self.gW[i]), kout, None,
                                 0.0), "gemm dW")
            check(L.cm_colsum(ptr(dcur_f), ptr(self.gb[i]), B, kout), "colsum")
            if i > 0:
                # dG = dcur W^T ; dZ = dG * gelu'(Z)
                check(L.cm_gemm_bf16(0, 1, B, kin, kout, ptr(dcur_bf), kout, ptr(self.Wbf[i]), kout,
                                     ptr(self.dG[i - 1]), kin, None, 0.0), "gemm dG")
                check(L.cm_gelu_bwd(ptr(self.dG[i - 1]), ptr(self.Z[i - 1]), ptr(self.dZbf[i - 1]),
                                    ptr(self.dZ[i - 1]), B * kin), "gelu bwd")
                dcur_bf, dcur_f = self.dZbf[i - 1], self.dZ[i - 1]
        return float(self.rowloss[:B].sum().item()) / (B * N_OUT)

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.t += 1
        L = self.L
        for i in range(3):
            check(L.cm_adam(ptr(self.W[i]), ptr(self.Wbf[i]), ptr(self.gW[i]), ptr(self.mW[i]), ptr(self.vW[i]),
                            self.W[i].numel(), lr, b1, b2, eps, self.t, wd), "adam")
            check(L.cm_adam(ptr(self.b[i]), None, ptr(self.gb[i]), ptr(self.mb[i]), ptr(self.vb[i]),
                            self.b[i].numel(), lr, b1, b2, eps, self.t, 0.0), "adam")

    def train(self, X, T, epochs=3, batch=1024, lr=1e-3, seed=0, log=None, decay=0.5):
        """Mini-batch Adam over device tensors X (N, 769), T (N, 256)."""
        torch = self.torch
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot, cnt = 0.0, 0
            lr_ep = lr * (decay ** ep)
            for s0 in range(0, n - batch + 1, batch):
                idx = perm[s0:s0 + batch]
                xb = X.index_select(0, idx).contiguous()
                tb = T.index_select(0, idx).contiguous()
                self.forward(xb)
                tot += self.backward(tb, batch)
                cnt += 1
                self.adam_step(lr_ep)
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: mse {hist[-1]:.4e}")
        return hist
