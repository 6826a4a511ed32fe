"""The controller network's CPU restatement and the label rule, pinned
without a GPU: bfloat16 rounding against torch's conversion, the
back-propagation against central finite differences in float64, and the
cheapest-(m, s) label rule against a brute-force loop."""
import numpy as np
import pytest
import torch

import mlp_oracle as M
from espec_amd import krylov


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0, 1e-40,
                         np.inf, -np.inf]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = M.bf16(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(M.bf16(np.array([np.nan], dtype=np.float32)))[0]


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.5, 0.5, M.NPAR)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.standard_normal((M.BATCH, 8))
    t = rng.uniform(0, 1, (M.BATCH, 2))
    g, _ = M.grad(p, x, t, bf16_mode=False)
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.B1), 6), rng.choice(np.arange(M.B1, M.W2), 3),
                          rng.choice(np.arange(M.W2, M.B2), 8), rng.choice(np.arange(M.B2, M.W3), 3),
                          rng.choice(np.arange(M.W3, M.B3), 4), [M.B3, M.B3 + 1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, x, t) - M.loss(pm, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.standard_normal((M.BATCH, 8))
    t = rng.uniform(0, 1, (M.BATCH, 2))
    g64, l64 = M.grad(p, x, t, bf16_mode=False)
    g16, l16 = M.grad(p, x, t, bf16_mode=True)
    assert abs(l16 - l64) < 2e-2 * l64
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)


def test_forward_outputs_in_unit_interval():
    o, _ = M.forward(_params(5) * 20, np.random.default_rng(5).standard_normal((100, 8)))
    assert np.all((o >= 0) & (o <= 1))


def _brute_targets(lab, tol, overhead):
    ns, nv = lab.shape[:2]
    out = np.zeros((ns, nv // 2, 2))
    for k in range(ns):
        for pk in range(nv // 2):
            best = (np.inf, 20, 0.5)
            for mi, m in enumerate(range(4, 21)):
                for si, s in enumerate((0.5, 1.0, 2.0)):
                    e = max(lab[k, 2 * pk, mi, si], lab[k, 2 * pk + 1, mi, si])
                    lim = {0.5: tol / 2, 1.0: tol, 2.0: 2 * tol}[s]
                    cost = {0.5: 2 * (m + 1 + overhead), 1.0: m + 1 + overhead,
                            2.0: (m + 1 + overhead + 1) / 2}[s]
                    if e <= lim and cost < best[0]:
                        best = (cost, m, s)
            out[k, pk] = ((best[1] - 4) / 16, (np.log2(best[2]) + 1) / 2)
    return out


def test_label_rule_matches_brute_force():
    rng = np.random.default_rng(6)
    ms = np.arange(4, 21)
    lab = np.zeros((40, 6, 17, 3))
    for k in range(40):
        for v in range(6):
            rate = rng.uniform(0.5, 3.0)
            for si, t in enumerate((0.5, 1.0, 2.0)):
                lab[k, v, :, si] = 10.0 ** (-(ms - 4) * rate) * t ** rng.uniform(2, 8)
    lab[3, 2] = 1.0  # nothing feasible for packet 1 of step 3
    for tol in (1e-10, 1e-6):
        got = krylov.targets(lab, tol, 4.0)
        ref = _brute_targets(lab, tol, 4.0)
        assert np.allclose(got, ref)
    assert np.allclose(krylov.targets(lab, 1e-10, 4.0)[3, 1], (1.0, 0.0))


def test_krylov_estimate_oracle_two_level_closed_form():
    """m = 2: exp(-i t T) of a 2x2 symmetric matrix in closed form."""
    a, b, c, beta0, t = 0.3, -0.2, 0.15, 1.7, 2.5
    est = M.krylov_est(np.array([a, b]), np.array([c, 0.4]), beta0, 2, t)
    mean, half = (a + b) / 2, np.hypot((a - b) / 2, c)
    off = -1j * np.exp(-1j * mean * t) * np.sin(half * t) * c / half
    assert est == pytest.approx(beta0 * 0.4 * abs(off), rel=1e-12)


def test_krylov_difference_is_the_true_error_at_full_order():
    """With M = n (the Krylov space is the whole space) beta0 |y_m - y_M| is
    the exact error of the order-m Lanczos approximation of exp(-i t H) u."""
    from scipy.linalg import expm
    rng = np.random.default_rng(7)
    n = 12
    A = rng.standard_normal((n, n))
    H = 0.3 * (A + A.T)
    u = rng.standard_normal(n)
    beta0 = np.linalg.norm(u)
    V = np.zeros((n, n))
    al, bt = np.zeros(n), np.zeros(n)
    V[:, 0] = u / beta0
    for j in range(n):  # Lanczos with full reorthogonalisation
        w = H @ V[:, j]
        al[j] = V[:, j] @ w
        w -= V[:, :j + 1] @ (V[:, :j + 1].T @ w)
        w -= V[:, :j + 1] @ (V[:, :j + 1].T @ w)
        bt[j] = np.linalg.norm(w)
        if j + 1 < n:
            V[:, j + 1] = w / bt[j]
    t = 1.3
    exact = expm(-1j * t * H) @ u
    for m in (3, 5, 8):
        T = np.diag(al[:m]) + np.diag(bt[:m - 1], 1) + np.diag(bt[:m - 1], -1)
        approx = beta0 * V[:, :m] @ expm(-1j * t * T)[:, 0]
        err = np.linalg.norm(exact - approx)
        assert M.krylov_diff(al, bt, beta0, m, n, t) == pytest.approx(err, rel=1e-8)
