"""Pins of the learned-preconditioner oracle (oracle/pc_oracle.py): the
convolution convention by brute force, adjoints by inner-product identities,
the backward pass by finite differences of the float64 graph, BF16 rounding
and Adam against torch's CPU implementations, flexible CG against its closed
forms (exact-inverse and scalar preconditioners), and the gain of a trained
net over plain CG."""
import numpy as np
import pytest
import torch

import pc_oracle as P

F5 = (1 / (0.5 * 0.501)) ** 2
F4 = 4 + F5


def test_conv_matches_brute_force_loop():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((9, 4, 3))
    out = P.conv(x, w)
    ref = np.zeros((2, 5, 6, 4))
    for b in range(2):
        for j in range(5):
            for i in range(6):
                for t in range(9):
                    dy, dx = t // 3 - 1, t % 3 - 1
                    ref[b, j, i] += w[t] @ x[b, (j + dy) % 5, (i + dx) % 6]
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_adjoint_identities():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 8, 6, 3))
    w = rng.standard_normal((9, 5, 3))
    g = rng.standard_normal((2, 8, 6, 5))
    lhs = (P.conv(x, w) * g).sum()
    assert np.isclose(lhs, (x * P.conv_t(g, w)).sum(), rtol=1e-12)
    assert np.isclose(lhs, (w * P.conv_dw(x, g)).sum(), rtol=1e-12)
    c = rng.standard_normal((2, 4, 3, 3))
    assert np.isclose((P.up(c) * x).sum(), (c * P.up_t(x)).sum(), rtol=1e-12)
    assert np.isclose((P.pool(x) * c).sum(), (x * 0.25 * P.up(c)).sum(), rtol=1e-12)


def test_identity_net_is_scaled_identity():
    """lift/proj centre taps (1, 1/f4) and every other weight zero: z = r/f4."""
    th = np.zeros(P.NPARAM, np.float32)
    o = P.OFFSETS
    th[o["lift"][0] + 4 * P.C0 + 0] = 1.0
    th[o["proj"][0] + 4 * P.C0 + 0] = np.float32(1.0 / F4)
    r = np.random.default_rng(3).standard_normal((3, 16, 24))
    z, _ = P.forward(th, r)
    assert np.allclose(z, r * np.float32(1.0 / F4), rtol=1e-6, atol=0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    th = P.init_params(F4, 5).astype(np.float64) + 0.05 * rng.standard_normal(P.NPARAM)
    r = rng.standard_normal((3, 16, 16))
    e = rng.standard_normal((3, 16, 16)) * 0.05
    L, g = P.loss_grad(th, r, e, bf16_path=False)
    idx = []
    for name, (o, ci, co) in P.OFFSETS.items():
        idx += list(o + rng.choice(9 * ci * co, 4, replace=False))
    h = 1e-6
    for k in idx:
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (P.loss_grad(tp, r, e, False)[0] - P.loss_grad(tm, r, e, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g).max()), (k, fd, g[k])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 10000),
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(P.bf16(x), ref)


def test_adam_matches_torch():
    rng = np.random.default_rng(7)
    th = rng.standard_normal(50).astype(np.float32)
    t = torch.nn.Parameter(torch.from_numpy(th.copy()))
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    a = P.Adam(50, lr=1e-2)
    for _ in range(5):
        g = rng.standard_normal(50).astype(np.float32)
        t.grad = torch.from_numpy(g.copy())
        opt.step()
        th = a.step(th, g)
    assert np.allclose(th, t.detach().numpy(), rtol=1e-5, atol=1e-6)


def _exact_inverse(r):
    ky = 2 * np.pi * np.fft.fftfreq(r.shape[1])[:, None]
    kx = 2 * np.pi * np.fft.fftfreq(r.shape[2])[None, :]
    lam = F4 - 2 * np.cos(kx) - 2 * np.cos(ky)
    return np.real(np.fft.ifft2(np.fft.fft2(r) / lam))


def test_fcg_closed_forms():
    rng = np.random.default_rng(8)
    b = rng.standard_normal((2, 16, 16))
    x0 = np.zeros_like(b)
    # exact inverse: one iteration, exact solution
    x, its = P.fcg(b, x0.copy(), F4, 1e-10, _exact_inverse)
    assert list(its) == [1, 1]
    assert np.allclose(P.apply_a(x, F4), b, atol=1e-10 * np.abs(b).max())
    # scalar preconditioner: same iterations as plain CG
    _, its_cg, _ = P.cg(b, x0.copy(), F4, 1e-10)
    _, its_s = P.fcg(b, x0.copy(), F4, 1e-10, lambda r: r / F4)
    assert list(its_s) == list(its_cg)


def test_trained_net_halves_iterations():
    rng = np.random.default_rng(9)
    H = W = 32

    def rhs():
        return rng.standard_normal((3, H, W))

    pairs = []
    for _ in range(4):
        xf, its, snaps = P.cg(rhs(), np.zeros((3, H, W)), F4, 1e-10)
        pairs += [(r, xf[k] - xk) for k, r, xk in snaps]
    th = P.init_params(F4, 1)
    opt = P.Adam(P.NPARAM, lr=3e-3)
    n = 200
    for s in range(n):
        idx = rng.choice(len(pairs), 3, replace=False)
        opt.lr = 3e-3 * 0.5 * (1 + np.cos(np.pi * s / n))
        L, g = P.loss_grad(th, np.stack([pairs[i][0] for i in idx]),
                           np.stack([pairs[i][1] for i in idx]))
        th = opt.step(th, g)
    b = rhs()
    _, its_cg, _ = P.cg(b, np.zeros_like(b), F4, 1e-10)
    x, its = P.fcg(b, np.zeros_like(b), F4, 1e-10, lambda r: P.forward(th, r)[0])
    assert np.all(2 * its <= its_cg + 1), (its, its_cg)
    res = np.linalg.norm((b - P.apply_a(x, F4)).reshape(3, -1), axis=1)
    assert np.all(res <= 1e-10 * np.linalg.norm(b.reshape(3, -1), axis=1))
