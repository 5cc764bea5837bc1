"""Spectral surrogate: an MLP trained and evaluated by the project's own
BF16 MFMA kernels (include/pmmlp.h), with FP32 master weights.

Layers: ``n_in -> hidden x depth (GELU) -> n_out``.  Forward: one GEMM per
layer with the bias and GELU fused into its epilogue (the pre-activations
are kept for the backward pass).  Backward: per layer one GEMM for the
weight gradient (dY^T A over the batch), one column sum for the bias
gradient and one GEMM for the input gradient with the GELU derivative fused
into its epilogue.  Loss: weighted MSE.  Optimiser: Adam (decoupled weight
decay) on the FP32 masters; the GEMMs round operands to BF16 as they stage
them.  ``SurrogateEnsemble`` trains several members from different seeds;
their spread is the predictive uncertainty.
"""

from __future__ import annotations

import numpy as np

from .. import _native as N


def _torch():
    import torch

    return torch


class SpectralMLP:
    def __init__(self, n_in: int = 4, n_out: int = 256, hidden: int = 256, depth: int = 4,
                 seed: int = 0, device: int = 0):
        torch = _torch()
        self.lib = N.load()
        self.device = torch.device("cuda", device)
        self.sizes = [n_in] + [hidden] * depth + [n_out]
        rng = np.random.default_rng(seed)
        # every weight and bias is a view of one flat buffer (all weights,
        # then all biases), the gradients of another of the same layout, so
        # one Adam launch updates everything (two with weight decay)
        self.shapes = [(self.sizes[k + 1], self.sizes[k]) for k in range(len(self.sizes) - 1)]
        self._nW = sum(fo * fi for fo, fi in self.shapes)
        total = self._nW + sum(fo for fo, _ in self.shapes)
        self._carve(torch.zeros(total, dtype=torch.float32, device=self.device),
                    torch.zeros(total, dtype=torch.float32, device=self.device))
        for k, (fo, fi) in enumerate(self.shapes):
            std = np.sqrt(2.0 / fi) if k < len(self.shapes) - 1 else np.sqrt(1.0 / fi)
            self.W[k].copy_(torch.from_numpy(rng.normal(0, std, (fo, fi)).astype(np.float32)))
        self._adam = None
        self.t = 0
        self._fixed_stream = None  # set during fit: the stream looked up once

    def _carve(self, flat, gflat) -> None:
        """Weights and biases as views of ``flat``, their gradients of
        ``gflat`` (both 1-D, all weights then all biases)."""
        self._flat, self._gflat = flat, gflat
        self.W, self.b, self._gW, self._gb = [], [], [], []
        ow, ob = 0, self._nW
        for fo, fi in self.shapes:
            self.W.append(flat[ow:ow + fo * fi].view(fo, fi))
            self.b.append(flat[ob:ob + fo])
            self._gW.append(gflat[ow:ow + fo * fi].view(fo, fi))
            self._gb.append(gflat[ob:ob + fo])
            ow += fo * fi
            ob += fo

    def woff(self, k) -> int:
        """Offset of layer k's weights in the flat buffer."""
        return sum(fo * fi for fo, fi in self.shapes[:k])

    def boff(self, k) -> int:
        """Offset of layer k's biases in the flat buffer."""
        return self._nW + sum(fo for fo, _ in self.shapes[:k])

    # ---------------------------------------------------------------- helpers
    def _stream(self):
        if self._fixed_stream is not None:
            return self._fixed_stream
        return _torch().cuda.current_stream(self.device).cuda_stream

    def _check(self, rc, what):
        if rc != 0:
            raise N.NativeError(f"{what} failed ({rc})")

    def _gemm(self, mode, M, Nn, K, A, lam, lak, B, lbk, lbn, C, ldc, bias=None, Z=None):
        rc = self.lib.pm_mlp_gemm(self._stream(), mode, M, Nn, K, A.data_ptr(), lam, lak,
                                  B.data_ptr(), lbk, lbn, C.data_ptr(), ldc,
                                  bias.data_ptr() if bias is not None else None,
                                  Z.data_ptr() if Z is not None else None)
        self._check(rc, "pm_mlp_gemm")

    def params(self):
        return [(w.cpu().numpy(), b.cpu().numpy()) for w, b in zip(self.W, self.b)]

    # ---------------------------------------------------------------- passes
    def forward(self, X, keep: bool = False):
        """X: (M, n_in) float32 CUDA tensor.  Returns Y, and with ``keep``
        the cache (inputs of every layer, hidden pre-activations)."""
        torch = _torch()
        X = X.contiguous()
        M = X.shape[0]
        acts, zs = [X], []
        a = X
        L = len(self.W)
        for k in range(L):
            fo, fi = self.W[k].shape
            out = torch.empty((M, fo), dtype=torch.float32, device=self.device)
            if k < L - 1:
                z = torch.empty_like(out)
                # C = A W^T: B[k][n] = W[n][k]
                self._gemm(N.MLP_GELU, M, fo, fi, a, fi, 1, self.W[k], 1, fi, out, fo,
                           self.b[k], z)
                zs.append(z)
            else:
                self._gemm(N.MLP_LINEAR, M, fo, fi, a, fi, 1, self.W[k], 1, fi, out, fo,
                           self.b[k])
            acts.append(out)
            a = out
        return (a, (acts, zs)) if keep else a

    def backward(self, cache, dY):
        """Gradients [(dW, db)] of sum(dY * Y), as views of the network's
        gradient buffer (overwritten by the next call)."""
        torch = _torch()
        acts, zs = cache
        g = dY.contiguous()
        M = g.shape[0]
        grads = [None] * len(self.W)
        for k in range(len(self.W) - 1, -1, -1):
            fo, fi = self.W[k].shape
            dW = self._gW[k]
            # dW[o][i] = sum_m g[m][o] a[m][i]: A[m'=o][k'=m] = g[m][o]
            self._gemm(N.MLP_PLAIN, fo, fi, M, g, 1, fo, acts[k], fi, 1, dW, fi)
            db = self._gb[k]
            self._check(self.lib.pm_mlp_colsum(self._stream(), M, fo, g.data_ptr(), fo,
                                               db.data_ptr()), "pm_mlp_colsum")
            grads[k] = (dW, db)
            if k > 0:
                gp = torch.empty((M, fi), dtype=torch.float32, device=self.device)
                # g_prev[m][i] = (sum_o g[m][o] W[o][i]) * gelu'(z_prev[m][i])
                self._gemm(N.MLP_DGELU, M, fi, fo, g, fo, 1, self.W[k], fi, 1, gp, fi,
                           None, zs[k - 1])
                g = gp
        return grads

    def mse(self, Y, T, w=None):
        torch = _torch()
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        M, Nn = Y.shape
        self._check(self.lib.pm_mlp_mse(self._stream(), M, Nn, Y.data_ptr(), T.data_ptr(),
                                        w.data_ptr() if w is not None else None,
                                        dY.data_ptr(), loss.data_ptr()), "pm_mlp_mse")
        return loss, dY

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        torch = _torch()
        if self._adam is None:
            self._adam = (torch.zeros_like(self._flat), torch.zeros_like(self._flat))
        mf, vf = self._adam
        self.t += 1
        flat = all(gW.data_ptr() == a.data_ptr() and gb.data_ptr() == c.data_ptr()
                   for (gW, gb), a, c in zip(grads, self._gW, self._gb))
        if flat:
            # the whole buffer at once (weights then biases; decay on weights)
            spans = [(0, self._flat.numel(), wd)] if wd == 0.0 else \
                [(0, self._nW, wd), (self._nW, self._flat.numel(), 0.0)]
            for a, e, decay in spans:
                p, g, m, v = (t[a:e] for t in (self._flat, self._gflat, mf, vf))
                self._check(self.lib.pm_mlp_adam(self._stream(), e - a, p.data_ptr(),
                                                 g.data_ptr(), m.data_ptr(), v.data_ptr(),
                                                 lr, b1, b2, eps, decay, self.t), "pm_mlp_adam")
            return
        # gradients given elsewhere: one launch per tensor, moments from the
        # same flat buffers
        ow, ob = 0, self._nW
        for (W, b), (gW, gb) in zip(zip(self.W, self.b), grads):
            for p, g, o, decay in ((W, gW, ow, wd), (b, gb, ob, 0.0)):
                n = p.numel()
                self._check(self.lib.pm_mlp_adam(self._stream(), n, p.data_ptr(),
                                                 g.contiguous().data_ptr(),
                                                 mf[o:o + n].data_ptr(), vf[o:o + n].data_ptr(),
                                                 lr, b1, b2, eps, decay, self.t), "pm_mlp_adam")
            ow += W.numel()
            ob += b.numel()

    def fit(self, X, T, epochs: int = 200, batch: int = 256, lr: float = 2e-3, w=None,
            seed: int = 0, lr_final: float = 1e-4, wd: float = 0.0):
        """Minibatch Adam with a cosine learning-rate decay; returns the
        per-epoch mean loss (host floats)."""
        rng = np.random.default_rng(seed)
        hist = []
        self._fixed_stream = self._stream()
        try:
            self._fit_epochs(X, T, epochs, batch, lr, w, rng, lr_final, wd, hist)
        finally:
            self._fixed_stream = None
        return hist

    def _fit_epochs(self, X, T, epochs, batch, lr, w, rng, lr_final, wd, hist):
        torch = _torch()
        M = X.shape[0]
        steps = epochs * ((M + batch - 1) // batch)
        s = 0
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(M)).to(self.device)
            tot = torch.zeros(1, dtype=torch.float32, device=self.device)
            nb = 0
            for i0 in range(0, M, batch):
                idx = perm[i0:i0 + batch]
                xb, tb = X[idx], T[idx]
                Y, cache = self.forward(xb, keep=True)
                loss, dY = self.mse(Y, tb, w)
                grads = self.backward(cache, dY)
                lr_s = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / max(1, steps)))
                self.adam_step(grads, lr=float(lr_s), wd=wd)
                tot += loss
                nb += 1
                s += 1
            hist.append(float(tot.item()) / nb)


class SurrogateEnsemble:
    """Members trained from different seeds on the same data; prediction =
    member mean, uncertainty = member standard deviation.

    The members' parameters live in one (members, P) buffer (each member's
    ``SpectralMLP`` views its row), and ``fit`` trains them in lockstep:
    every GEMM, column sum and loss of a step is one batched launch over the
    members and Adam one launch over the buffer.  Each member sees what its
    own ``fit(seed=k)`` would (same initial weights, same minibatch order):
    the parameters come out the same."""

    def __init__(self, n_members: int = 4, **kw):
        torch = _torch()
        self.members = [SpectralMLP(seed=1000 + k, **kw) for k in range(n_members)]
        m0 = self.members[0]
        self.P = torch.stack([m._flat for m in self.members]).contiguous()
        self.G = torch.zeros_like(self.P)
        for k, m in enumerate(self.members):
            m._carve(self.P[k], self.G[k])
        self._adam = None
        self.t = 0
        self.lib, self.device = m0.lib, m0.device

    def _check(self, rc, what):
        if rc != 0:
            raise N.NativeError(f"{what} failed ({rc})")

    def fit(self, X, T, epochs: int = 200, batch: int = 256, lr: float = 2e-3, w=None,
            lr_final: float = 1e-4, wd: float = 0.0):
        """Minibatch Adam with a cosine learning-rate decay, member k on the
        minibatch order of seed k; returns each member's per-epoch mean
        loss (host floats)."""
        torch = _torch()
        L, nm = self.members[0], len(self.members)
        lib, dev = self.lib, self.device
        st = torch.cuda.current_stream(dev).cuda_stream
        Pn = self.P.shape[1]
        M = X.shape[0]
        rngs = [np.random.default_rng(k) for k in range(nm)]
        hist = [[] for _ in range(nm)]
        steps = epochs * ((M + batch - 1) // batch)
        if self._adam is None:
            self._adam = (torch.zeros_like(self.P), torch.zeros_like(self.P))
        mA, vA = self._adam
        P0, G0 = self.P.data_ptr(), self.G.data_ptr()
        f4 = 4  # bytes per element
        s = 0
        for ep in range(epochs):
            perms = torch.stack([torch.from_numpy(r.permutation(M)) for r in rngs]).to(dev)
            tot = torch.zeros(nm, dtype=torch.float32, device=dev)
            nb = 0
            for i0 in range(0, M, batch):
                idx = perms[:, i0:i0 + batch]
                B = idx.shape[1]
                a = X[idx].contiguous()  # (nm, B, n_in)
                tb = T[idx].contiguous()
                acts, zs = [a], []
                for k, (fo, fi) in enumerate(L.shapes):
                    out = torch.empty((nm, B, fo), dtype=torch.float32, device=dev)
                    last = k == len(L.shapes) - 1
                    z = None if last else torch.empty_like(out)
                    self._check(lib.pm_mlp_gemm_batched(
                        st, N.MLP_LINEAR if last else N.MLP_GELU, nm, B, fo, fi,
                        a.data_ptr(), fi, 1, B * fi, P0 + f4 * L.woff(k), 1, fi, Pn,
                        out.data_ptr(), fo, B * fo, P0 + f4 * L.boff(k), Pn,
                        z.data_ptr() if z is not None else None, B * fo), "gemm")
                    if z is not None:
                        zs.append(z)
                    acts.append(out)
                    a = out
                Y = a
                nout = Y.shape[2]
                dY = torch.empty_like(Y)
                loss = torch.zeros(nm, dtype=torch.float32, device=dev)
                self._check(lib.pm_mlp_mse_batched(
                    st, nm, B, nout, Y.data_ptr(), tb.data_ptr(), B * nout,
                    w.data_ptr() if w is not None else None, dY.data_ptr(), loss.data_ptr()),
                    "mse")
                g = dY
                for k in range(len(L.shapes) - 1, -1, -1):
                    fo, fi = L.shapes[k]
                    # dW[o][i] = sum_m g[m][o] a[m][i]
                    self._check(lib.pm_mlp_gemm_batched(
                        st, N.MLP_PLAIN, nm, fo, fi, B, g.data_ptr(), 1, fo, B * fo,
                        acts[k].data_ptr(), fi, 1, B * fi, G0 + f4 * L.woff(k), fi, Pn,
                        None, 0, None, 0), "gemm")
                    self._check(lib.pm_mlp_colsum_batched(
                        st, nm, B, fo, g.data_ptr(), fo, B * fo, G0 + f4 * L.boff(k), Pn),
                        "colsum")
                    if k > 0:
                        gp = torch.empty((nm, B, fi), dtype=torch.float32, device=dev)
                        self._check(lib.pm_mlp_gemm_batched(
                            st, N.MLP_DGELU, nm, B, fi, fo, g.data_ptr(), fo, 1, B * fo,
                            P0 + f4 * L.woff(k), fi, 1, Pn, gp.data_ptr(), fi, B * fi,
                            None, 0, zs[k - 1].data_ptr(), B * fi), "gemm")
                        g = gp
                lr_s = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / max(1, steps)))
                self.t += 1
                if wd == 0.0:
                    spans = [(0, self.P.numel(), 0.0)]
                else:
                    spans = []
                    for q in range(nm):
                        spans += [(q * Pn, q * Pn + L._nW, wd), (q * Pn + L._nW, (q + 1) * Pn, 0.0)]
                for a0, e0, decay in spans:
                    self._check(lib.pm_mlp_adam(
                        st, e0 - a0, P0 + f4 * a0, G0 + f4 * a0, mA.data_ptr() + f4 * a0,
                        vA.data_ptr() + f4 * a0, float(lr_s), 0.9, 0.999, 1e-8, decay, self.t),
                        "adam")
                tot += loss
                nb += 1
                s += 1
            h = tot.cpu().numpy()
            for q in range(nm):
                hist[q].append(float(h[q]) / nb)
        for m in self.members:
            m.t = self.t
        return hist

    def predict(self, X, chunk: int = 16384):
        torch = _torch()
        outs = []
        for m in self.members:
            parts = [m.forward(X[i:i + chunk]) for i in range(0, X.shape[0], chunk)]
            outs.append(torch.cat(parts))
        S = torch.stack(outs)
        return S.mean(0), S.std(0)
