"""Pins of the learned selector's CPU oracle (oracle/mlp.py), CPU only.

bf16 rounding is checked against torch's float32 -> bfloat16 conversion; the
gradient against central finite differences and against torch autograd (an
independent backpropagation); the Adam step against torch.optim.Adam; top-k
against a brute-force ordering. A dropped term, wrong transpose or wrong mask
in the backward fails the finite-difference and autograd checks.
"""
import numpy as np
import pytest
import torch

from oracle import mlp as M


def rand_params(rng, nfeat=20):
    p = np.zeros(M.NPARAM)
    W1 = np.zeros((M.FPAD, M.HID))
    W1[:nfeat] = rng.normal(0, np.sqrt(2.0 / nfeat), (nfeat, M.HID))
    p[M.OW1:M.OB1] = W1.ravel()
    p[M.OB1:M.OW2] = rng.normal(0, 0.1, M.HID)
    p[M.OW2:M.OB2] = rng.normal(0, np.sqrt(2.0 / M.HID), M.HID * M.HID)
    p[M.OB2:M.OW3] = rng.normal(0, 0.1, M.HID)
    p[M.OW3:M.OB3] = rng.normal(0, 1.0 / 16, M.HID)
    p[M.OB3] = -3.0
    return p


def rand_batch(rng, n, nfeat=20, nocc=6):
    X = np.zeros((n, M.FPAD))
    for i in range(n):
        X[i, rng.choice(nfeat, nocc, replace=False)] = 1.0
    return X, rng.normal(-4.0, 1.0, n)


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(0)
    x = rng.normal(0, 10, 4000).astype(np.float32)
    # exact ties (low half 0x8000) with even and odd kept parts, plus near-ties
    u = rng.integers(0, 2**31 - 1, 1000, dtype=np.uint64).astype(np.uint32)
    ties = ((u & np.uint32(0x7F7F0000)) | np.uint32(0x8000)).view(np.float32)
    near = ((u & np.uint32(0x7F7F0000)) | np.uint32(0x7FFF)).view(np.float32)
    allx = np.concatenate([x, ties, near, -ties, np.array([0.0, -0.0, 1.0, 65504.0], np.float32)])
    ref = torch.from_numpy(allx).to(torch.bfloat16).to(torch.float32).numpy()
    got = M.bf16(allx)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_features_bit_layout():
    d = np.array([0b1011, 0b110000], dtype=np.uint64)
    X = M.features("genci", 8, d)
    assert X.shape == (2, 64)
    assert list(np.nonzero(X[0])[0]) == [0, 1, 3] and list(np.nonzero(X[1])[0]) == [4, 5]
    X = M.features("fullci", 6, np.array([[0b11, 0b101]], dtype=np.uint64))
    assert list(np.nonzero(X[0])[0]) == [0, 1, 6, 8]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rand_params(rng)
    X, t = rand_batch(rng, 40)
    _, g = M.loss_grad(p, X, t, rounding=False)
    blocks = [(M.OW1, 20 * M.HID), (M.OB1, M.HID), (M.OW2, M.HID * M.HID), (M.OB2, M.HID), (M.OW3, M.HID),
              (M.OB3, 1)]
    h = 1e-6
    for off, size in blocks:
        for i in rng.choice(size, min(size, 12), replace=False):
            q = p.copy()
            q[off + i] += h
            lp, _ = M.loss_grad(q, X, t, rounding=False)
            q[off + i] -= 2 * h
            lm, _ = M.loss_grad(q, X, t, rounding=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[off + i]) <= 1e-6 * max(1.0, abs(fd)), (off, i, fd, g[off + i])


def test_gradient_matches_torch_autograd():
    rng = np.random.default_rng(2)
    p = rand_params(rng)
    X, t = rand_batch(rng, 200)
    loss, g = M.loss_grad(p, X, t, rounding=False)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = P[M.OW1:M.OB1].view(M.FPAD, M.HID)
    W2 = P[M.OW2:M.OB2].view(M.HID, M.HID)
    h1 = torch.relu(torch.tensor(X) @ W1 + P[M.OB1:M.OW2])
    h2 = torch.relu(h1 @ W2 + P[M.OB2:M.OW3])
    y = h2 @ P[M.OW3:M.OB3] + P[M.OB3]
    L = ((y - torch.tensor(t)) ** 2).mean()
    L.backward()
    assert abs(float(L) - loss) < 1e-12 * max(1.0, loss)
    np.testing.assert_allclose(g, P.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_rounded_gradient_close_to_exact():
    # the bf16 rounding points perturb the gradient only at bf16 precision
    rng = np.random.default_rng(3)
    p = rand_params(rng)
    X, t = rand_batch(rng, 300)
    _, g0 = M.loss_grad(p, X, t, rounding=False)
    _, g1 = M.loss_grad(p, X, t, rounding=True)
    assert np.linalg.norm(g1 - g0) <= 2e-2 * np.linalg.norm(g0)


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=50)
    P = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([P], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 6):
        g = rng.normal(size=50)
        P.grad = torch.tensor(g)
        opt.step()
        p, m, v = M.adam_step(p, g, m, v, t, 1e-2)
    np.testing.assert_allclose(p, P.detach().numpy(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("k", [0, 1, 7, 50, 200])
def test_topk_select_brute_force(k):
    rng = np.random.default_rng(5)
    s = np.round(rng.normal(size=120), 1).astype(np.float32)  # many ties
    got = M.topk_select(s, k)
    ranked = sorted(range(s.size), key=lambda i: (-float(s[i]), i))[: min(k, s.size)]
    assert list(got) == sorted(ranked)
