"""CPU restatement of the Krylov-step controller's network and of the
Krylov residual estimate (test infrastructure only: only tests/ and
bench.py's checks import it; the product never does).

The network (8 -> 64 -> 64 -> 2, tanh, tanh, sigmoid) is written out with
numpy in two modes:

* ``bf16=True`` rounds exactly where the device kernels round (inputs,
  weights, hidden activations and back-propagated dz to bfloat16, round to
  nearest even), summing in float32, so that device results agree up to
  the order of float32 sums;
* ``bf16=False`` computes in float64 without rounding: the mathematics the
  kernels implement, which the tests pin with finite differences.

No reference counterpart: the controller is an addition to the short-
iterative Lanczos propagator (sil.f / plnz.f), whose error estimates are
restated here with scipy's matrix exponential: the residual estimate
err(m, tau) = beta0 beta_m |[exp(-i tau T_m)]_{m,1}| and the difference
beta0 |y_m - y_M| of the order-m and order-M approximations
y_j = exp(-i tau T_j) e1.
"""
import numpy as np

NIN, NHID, NOUT = 8, 64, 2
W1 = 0
B1 = W1 + NHID * NIN
W2 = B1 + NHID
B2 = W2 + NHID * NHID
W3 = B2 + NHID
B3 = W3 + NOUT * NHID
NPAR = B3 + NOUT  # 4866
BATCH = 64


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, np.uint64(0x7FC00000), r).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p)
    return (p[W1:B1].reshape(NHID, NIN), p[B1:W2], p[W2:B2].reshape(NHID, NHID), p[B2:W3],
            p[W3:B3].reshape(NOUT, NHID), p[B3:NPAR])


def _mode(bf):
    if bf:
        return (lambda v: bf16(np.asarray(v, dtype=np.float32))), np.float32
    return (lambda v: np.asarray(v, dtype=np.float64)), np.float64


def forward(p, x, bf16_mode=True):
    """Outputs sigmoid(y) [n][2] and the cached activations."""
    r, dt = _mode(bf16_mode)
    w1, b1, w2, b2, w3, b3 = (r(a) for a in unpack(p))
    x = r(np.asarray(x).reshape(-1, NIN))
    h1 = r(np.tanh((x @ w1.T).astype(dt) + b1))
    h2 = r(np.tanh((h1 @ w2.T).astype(dt) + b2))
    y = (h2 @ w3.T).astype(dt) + b3
    o = (1.0 / (1.0 + np.exp(-y))).astype(dt)
    return o, (x, h1, h2, w1, w2, w3)


def grad(p, x, t, bf16_mode=True):
    """Gradient of mean((sigmoid(y) - t)^2) over all outputs, and the loss."""
    r, dt = _mode(bf16_mode)
    o, (x, h1, h2, w1, w2, w3) = forward(p, x, bf16_mode)
    t = np.asarray(t, dtype=dt).reshape(o.shape)
    n = o.size
    dy = r(2.0 * (o - t) / n * o * (1.0 - o))
    g = np.zeros(NPAR, dtype=dt)
    g[W3:B3] = (dy.T @ h2).ravel()
    g[B3:NPAR] = dy.sum(axis=0)
    d2 = r((dy @ w3) * (1.0 - h2 * h2))
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(axis=0)
    d1 = r((d2 @ w2) * (1.0 - h1 * h1))
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(axis=0)
    loss = float(np.mean((o - t) ** 2))
    return g, loss


def loss(p, x, t):
    """float64 loss without rounding (for finite differences)."""
    o, _ = forward(p, x, bf16_mode=False)
    return float(np.mean((o - np.asarray(t, dtype=np.float64).reshape(o.shape)) ** 2))


def krylov_est(alpha, beta, beta0, m, tau):
    """beta0 beta_m |[exp(-i tau T_m)]_{m,1}| for the leading m x m Lanczos
    tridiagonal (alpha[0..m), beta[0..m-1)) with beta_m = beta[m-1]."""
    from scipy.linalg import expm
    T = np.diag(alpha[:m]) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
    E = expm(-1j * tau * T)
    return beta0 * beta[m - 1] * abs(E[m - 1, 0])


def krylov_diff(alpha, beta, beta0, m, M, tau):
    """beta0 |y_m - y_M| with y_j = exp(-i tau T_j) e1 (zero padded): the
    error of the order-m Krylov approximation against the order-M one."""
    from scipy.linalg import expm

    def y(j):
        T = np.diag(alpha[:j]) + np.diag(beta[:j - 1], 1) + np.diag(beta[:j - 1], -1)
        return expm(-1j * tau * T)[:, 0]
    d = y(M)
    d[:m] -= y(m)
    return beta0 * np.linalg.norm(d)
