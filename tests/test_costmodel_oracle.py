"""Pins the cost-model oracle (oracle/costmodel_ref.py) against facts it does not restate:
a direct-definition DFT, brute-force counting loops, a pure tone at a known bin, torch's
float64 Linear/GELU/autograd, central finite differences and torch.optim.Adam."""
import numpy as np
import pytest
import torch

from oracle import costmodel_ref as C


def test_band_power_matches_direct_dft():
    rng = np.random.default_rng(11)
    for n in (50, 256, 700, 1025):
        x = rng.standard_normal(n)
        k = max(1, n // 256)
        segs = []
        for s in range(k):
            seg = np.zeros(256)
            part = x[s * 256:(s + 1) * 256]
            seg[:len(part)] = part
            segs.append(seg)
        j = np.arange(256)
        ref = np.zeros(129)
        for seg in segs:
            for b in range(129):
                X = np.sum(seg * np.exp(-2j * np.pi * b * j / 256))
                ref[b] += abs(X) ** 2
        ref /= len(segs)
        assert np.allclose(C.band_power(x), ref, rtol=1e-10, atol=1e-9)


def test_tone_lands_in_its_band():
    n = np.arange(2048)
    for k0, band in ((3, 1), (20, 4), (40, 5), (100, 7)):
        f = C.features(np.cos(2 * np.pi * k0 * n / 256))
        assert np.argmax(f[4:]) == band
        assert f[4 + band] > -1e-6           # fraction ~ 1 -> log ~ 0
        others = np.delete(f[4:], band)
        assert np.all(others < -20)          # fraction ~ 0 -> log ~ log(1e-10)


def test_counts_and_rms_brute_force():
    rng = np.random.default_rng(12)
    for _ in range(50):
        n = int(rng.integers(3, 400))
        x = np.round(rng.standard_normal(n) * 3)  # ties and zeros
        ext = sum(1 for i in range(1, n - 1) if (x[i] - x[i - 1]) * (x[i + 1] - x[i]) < 0)
        zc = sum(1 for i in range(n - 1) if x[i] * x[i + 1] < 0)
        f = C.features(x)
        assert f[0] == pytest.approx(np.log(n))
        assert f[1] == pytest.approx(np.log1p(ext)) and f[2] == pytest.approx(np.log1p(zc))
        assert f[3] == pytest.approx(np.log(np.sqrt(np.sum(x ** 2) / n) + 1e-12))


def _torch_model(params):
    P = C.unpack(params)
    l1, l2, l3 = (torch.nn.Linear(12, 64).double(), torch.nn.Linear(64, 64).double(),
                  torch.nn.Linear(64, 3).double())
    with torch.no_grad():
        for lin, w, b in ((l1, "W1", "b1"), (l2, "W2", "b2"), (l3, "W3", "b3")):
            lin.weight.copy_(torch.from_numpy(P[w].copy()))
            lin.bias.copy_(torch.from_numpy(P[b].copy()))
    gelu = torch.nn.GELU(approximate="none")
    return torch.nn.Sequential(l1, gelu, l2, gelu, l3)


def _case(seed, M=37):
    rng = np.random.default_rng(seed)
    params = rng.standard_normal(C.N_PARAMS) * 0.3
    norm = np.concatenate([rng.standard_normal(12), rng.uniform(0.5, 2.0, 12)])
    F = rng.standard_normal((M, 12)) * 2
    T = rng.standard_normal((M, 3))
    tw = np.array([1.0, 0.5, 2.0])
    return params, norm, F, T, tw


def test_forward_and_backward_match_torch_autograd():
    params, norm, F, T, tw = _case(13)
    model = _torch_model(params)
    X = torch.from_numpy((F - norm[:12]) / norm[12:])
    y = model(X)
    assert np.allclose(C.forward(params, norm, F)[0], y.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss = (torch.from_numpy(tw) * (y - torch.from_numpy(T)) ** 2).sum() / len(F)
    loss.backward()
    L, g = C.loss_and_grad(params, norm, F, T, tw)
    assert L == pytest.approx(loss.item(), rel=1e-12)
    ref = C.pack({"W1": model[0].weight.grad.numpy(), "b1": model[0].bias.grad.numpy(),
                  "W2": model[2].weight.grad.numpy(), "b2": model[2].bias.grad.numpy(),
                  "W3": model[4].weight.grad.numpy(), "b3": model[4].bias.grad.numpy()})
    assert np.allclose(g, ref, rtol=1e-10, atol=1e-12)


def test_gradient_finite_differences():
    params, norm, F, T, tw = _case(14, M=9)
    _, g = C.loss_and_grad(params, norm, F, T, tw)
    rng = np.random.default_rng(15)
    idx = rng.choice(C.N_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (C.loss_and_grad(p1, norm, F, T, tw)[0] - C.loss_and_grad(p2, norm, F, T, tw)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-8)


def test_adam_matches_torch():
    rng = np.random.default_rng(16)
    p0 = rng.standard_normal(C.N_PARAMS)
    grads = [rng.standard_normal(C.N_PARAMS) for _ in range(3)]
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step, g in enumerate(grads, 1):
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = C.adam(p, m, v, g, step, 3e-3)
    assert np.allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)
