"""GPU: the surrogate's BF16 MFMA kernels (include/pmmlp.h) against the O2
oracle (oracle/mlp.py), which rounds the same operands to BF16 and sums in
FP64.

Layer by layer the oracle is fed the device's own inputs, so the comparison
isolates the kernel arithmetic (FP32 sums of BF16 products): bar 1e-5
relative (max-norm).  End to end (rounding of independently computed
activations can flip single BF16 roundings) the bar is 1e-2.  Training with
the device kernels reduces the loss on a smooth target by 20x.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("M,Nn,K", [(70, 45, 37), (128, 256, 256), (5, 300, 4), (257, 64, 96),
                                    (33, 70, 300)])
def test_gemm_modes_match_oracle(mi355x, M, Nn, K):
    import torch
    from oracle import mlp as O2
    from prismo_amd import _native as N

    lib = N.load()
    rng = np.random.default_rng(M * 7 + Nn)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, Nn)).astype(np.float32)
    bias = rng.normal(size=Nn).astype(np.float32)
    Zin = rng.normal(size=(M, Nn)).astype(np.float32)
    ref = O2.bf16(A) @ O2.bf16(B)
    s = torch.cuda.current_stream().cuda_stream
    # operand layouts: row-major A, B given as B^T row-major (k fastest)
    layouts = {
        "A_rows_Bt": (_t(A), K, 1, _t(B.T), 1, K),
        "A_cols_B": (_t(A.T), 1, M, _t(B), Nn, 1),
    }
    for name, (a, lam, lak, b, lbk, lbn) in layouts.items():
        for mode in (N.MLP_LINEAR, N.MLP_GELU, N.MLP_DGELU, N.MLP_PLAIN):
            C = torch.zeros((M, Nn), dtype=torch.float32, device="cuda")
            Z = _t(Zin) if mode == N.MLP_DGELU else torch.zeros_like(C)
            bb = _t(bias) if mode in (N.MLP_LINEAR, N.MLP_GELU) else None
            rc = lib.pm_mlp_gemm(s, mode, M, Nn, K, a.data_ptr(), lam, lak, b.data_ptr(), lbk,
                                 lbn, C.data_ptr(), Nn, bb.data_ptr() if bb is not None else None,
                                 Z.data_ptr())
            assert rc == 0
            got = C.cpu().numpy()
            if mode == N.MLP_LINEAR:
                want = ref + bias
            elif mode == N.MLP_GELU:
                want = O2.gelu(ref + bias)
                assert _rel(Z.cpu().numpy(), ref + bias) <= 1e-5, (name, "Z")
            elif mode == N.MLP_DGELU:
                want = ref * O2.dgelu(Zin.astype(np.float64))
            else:
                want = ref
            assert _rel(got, want) <= 1e-5, (name, mode)


def test_forward_backward_layerwise(mi355x):
    import torch
    from oracle import mlp as O2
    from prismo_amd.surrogate.mlp import SpectralMLP

    net = SpectralMLP(n_in=4, n_out=20, hidden=64, depth=3, seed=5)
    rng = np.random.default_rng(9)
    X = _t(rng.uniform(-1, 1, (37, 4)))
    T = _t(rng.normal(size=(37, 20)))
    Y, cache = net.forward(X, keep=True)
    loss, dY = net.mse(Y, T)
    grads = net.backward(cache, dY)
    acts = [a.cpu().numpy() for a in cache[0]]
    zs = [z.cpu().numpy() for z in cache[1]]
    P = net.params()
    L = len(P)
    for k, (W, b) in enumerate(P):
        z = O2.bf16(acts[k]) @ O2.bf16(W).T + b
        if k < L - 1:
            assert _rel(zs[k], z) <= 1e-5, k
            assert _rel(acts[k + 1], O2.gelu(zs[k].astype(np.float64))) <= 1e-5, k
        else:
            assert _rel(acts[k + 1], z) <= 1e-5
    l_ref, d_ref = O2.mse(acts[-1], T.cpu().numpy())
    assert abs(loss.item() - l_ref) <= 1e-5 * l_ref
    assert _rel(dY.cpu().numpy(), d_ref) <= 1e-5
    # backward, layer by layer from the device's own upstream gradient
    g = dY.cpu().numpy().astype(np.float64)
    for k in range(L - 1, -1, -1):
        W, _ = P[k]
        dW, db = (t.cpu().numpy() for t in grads[k])
        assert _rel(dW, O2.bf16(g).T @ O2.bf16(acts[k])) <= 1e-5, k
        assert _rel(db, g.sum(0)) <= 1e-5, k
        if k > 0:
            g = (O2.bf16(g) @ O2.bf16(W)) * O2.dgelu(zs[k - 1].astype(np.float64))
            # continue from the oracle's value rounded as the device stores it
            g = g.astype(np.float32).astype(np.float64)
    # end to end against the oracle on its own activations
    Yo, co = O2.forward(P, X.cpu().numpy())
    _, do = O2.mse(Yo, T.cpu().numpy())
    go = O2.backward(P, co, do)
    assert _rel(Y.cpu().numpy(), Yo) <= 1e-2
    for k in range(L):
        assert _rel(grads[k][0].cpu().numpy(), go[k][0]) <= 1e-2, k


def test_adam_matches_oracle(mi355x):
    import torch
    from oracle import mlp as O2
    from prismo_amd import _native as N

    lib = N.load()
    rng = np.random.default_rng(4)
    n = 1000
    p, g, m, v = (rng.normal(size=n).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    P, G, Mm, V = _t(p), _t(g), _t(m), _t(v)
    s = torch.cuda.current_stream().cuda_stream
    assert lib.pm_mlp_adam(s, n, P.data_ptr(), G.data_ptr(), Mm.data_ptr(), V.data_ptr(),
                           1e-3, 0.9, 0.999, 1e-8, 0.01, 3) == 0
    pr, mr, vr = O2.adam(p.astype(np.float64), g, m, v, 1e-3, 0.9, 0.999, 1e-8, 0.01, 3)
    assert _rel(P.cpu().numpy(), pr) <= 1e-6
    assert _rel(Mm.cpu().numpy(), mr) <= 1e-6
    assert _rel(V.cpu().numpy(), vr) <= 1e-6


def test_training_reduces_loss(mi355x):
    from prismo_amd.surrogate.mlp import SpectralMLP

    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (512, 4))
    A = rng.normal(size=(4, 8))
    T = np.sin(1.5 * X @ A)
    net = SpectralMLP(n_in=4, n_out=8, hidden=64, depth=3, seed=2)
    hist = net.fit(_t(X), _t(T), epochs=120, batch=64, lr=3e-3)
    assert hist[-1] < hist[0] / 20, (hist[0], hist[-1])


def test_ensemble_lockstep_equals_members_one_by_one(mi355x):
    """Batched lockstep training of the ensemble (one launch per GEMM,
    column sum and loss for all members) gives every member the parameters
    its own fit(seed=k) gives; the loss histories agree to rounding (the loss
    sum uses atomics)."""
    from prismo_amd.surrogate.mlp import SpectralMLP, SurrogateEnsemble

    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (200, 4))
    T = np.tanh(X @ rng.normal(size=(4, 12)))
    kw = dict(n_in=4, n_out=12, hidden=48, depth=3)
    ens = SurrogateEnsemble(n_members=3, **kw)
    hist = ens.fit(_t(X), _t(T), epochs=4, batch=64, lr=3e-3)
    for k, m in enumerate(ens.members):
        ref = SpectralMLP(seed=1000 + k, **kw)
        h = ref.fit(_t(X), _t(T), epochs=4, batch=64, lr=3e-3, seed=k)
        for (w1, b1), (w2, b2) in zip(m.params(), ref.params()):
            assert np.array_equal(w1, w2) and np.array_equal(b1, b2), k
        assert np.allclose(hist[k], h, rtol=1e-5), (k, hist[k], h)
