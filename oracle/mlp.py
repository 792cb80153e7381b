"""CPU oracle of the learned selector's network (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this
module, as the checker; the product (pyci_amd) never imports it.

The network has no counterpart in the reference. Its definition is fixed by
include/pyci_engine.h (PYCI_MLP_* block) and restated here in numpy:

  x   occupation bits, 64 features (bit i of the determinant; for FullCI alpha
      bits 0..n-1, then beta bits n..2n-1; the rest zero)
  h1  = bf16(relu(x @ bf16(W1) + b1))             W1: 64 x 256
  h2  = bf16(relu(h1 @ bf16(W2) + b2))            W2: 256 x 256
  y   = h2 @ w3 + b3                              w3: 256 (FP32)
  L   = mean((y - t)^2)
Backward with the same rounding points:
  dy  = 2 (y - t) / n,  db3 = sum dy,  dw3 = h2^T dy
  dz2 = bf16(dy w3^T [h2 > 0]),  db2 = sum dz2,  dW2 = h1^T dz2
  dz1 = bf16((dz2 @ bf16(W2)^T) [h1 > 0]),  db1 = sum dz1,  dW1 = x^T dz1
Products accumulate in FP32 on the GPU (float64 here, so the oracle is the
exact value of the same rounded operands).

Pins (tests/test_mlp_oracle.py): bf16() against torch's float32 -> bfloat16
conversion (round to nearest even); loss_grad(rounding=False) against central
finite differences and against torch autograd; adam_step against
torch.optim.Adam; topk_select against a brute-force sort.
"""
import numpy as np

HID = 256
FPAD = 64
OW1 = 0
OB1 = OW1 + FPAD * HID
OW2 = OB1 + HID
OB2 = OW2 + HID * HID
OW3 = OB2 + HID
OB3 = OW3 + HID
NPARAM = OB3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    if p.dtype != np.float64:
        p = p.astype(np.float32)
    assert p.size == NPARAM
    return (p[OW1:OB1].reshape(FPAD, HID), p[OB1:OW2], p[OW2:OB2].reshape(HID, HID), p[OB2:OW3], p[OW3:OB3],
            p[OB3])


def features(kind, nbasis, dets):
    """Occupation-bit features [n, 64] (float32 0/1) of determinants."""
    d = np.asarray(dets, dtype=np.uint64)
    if kind == "fullci":
        d = d.reshape(-1, 2)
        n = d.shape[0]
        X = np.zeros((n, FPAD), np.float32)
        sh = np.arange(nbasis, dtype=np.uint64)
        X[:, :nbasis] = (d[:, :1] >> sh) & np.uint64(1)
        X[:, nbasis:2 * nbasis] = (d[:, 1:2] >> sh) & np.uint64(1)
        return X
    d = d.reshape(-1)
    X = np.zeros((d.size, FPAD), np.float32)
    X[:, :nbasis] = (d[:, None] >> np.arange(nbasis, dtype=np.uint64)) & np.uint64(1)
    return X


def _r(rounding):
    return bf16 if rounding else (lambda a: np.asarray(a, np.float64))


def forward(p, X, rounding=True):
    r = _r(rounding)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    f = np.float64
    z1 = r(X).astype(f) @ r(W1).astype(f) + b1
    h1 = r(np.maximum(z1, 0.0)).astype(f)
    z2 = h1 @ r(W2).astype(f) + b2
    h2 = r(np.maximum(z2, 0.0)).astype(f)
    y = h2 @ w3.astype(f) + f(b3)
    return y, (h1, h2)


def loss_grad(p, X, t, rounding=True):
    """Loss and flat gradient (float64) of the mean-squared error."""
    r = _r(rounding)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    f = np.float64
    y, (h1, h2) = forward(p, X, rounding)
    t = np.asarray(t, f)
    n = t.size
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = np.zeros(NPARAM)
    g[OB3] = dy.sum()
    g[OW3:OB3] = h2.T @ dy
    dz2 = r(dy[:, None] * w3.astype(f)[None, :] * (h2 > 0)).astype(f)
    g[OB2:OW3] = dz2.sum(0)
    g[OW2:OB2] = (h1.T @ dz2).ravel()
    dz1 = r((dz2 @ r(W2).astype(f).T) * (h1 > 0)).astype(f)
    g[OB1:OW2] = dz1.sum(0)
    g[OW1:OB1] = (r(X).astype(f).T @ dz1).ravel()
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t = 1-based step); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def topk_select(scores, k):
    """Indices of the k largest scores, ties broken by position; returned in
    ascending position (the order the selection appends them)."""
    s = np.asarray(scores, np.float32)
    k = min(int(k), s.size)
    if k <= 0:
        return np.zeros(0, np.int64)
    order = np.lexsort((np.arange(s.size), -s.astype(np.float64)))
    return np.sort(order[:k])
