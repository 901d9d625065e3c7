"""Cost-model kernels (features, MLP forward / fused backward / Adam) against the fp64
oracle, training on the method's own counts, and the ablation: the learned plan changes
queue order and LDS tile only, never the decomposition."""
import numpy as np
import pytest
import torch

from cuda_emd import EMDKernelConfig, costmodel, synth
from oracle import costmodel_ref as C

pytestmark = pytest.mark.gpu


def _csr(traces):
    off = np.zeros(len(traces) + 1, np.int64)
    off[1:] = np.cumsum([len(t) for t in traces])
    vals = np.concatenate(traces) if traces else np.zeros(0)
    return np.ascontiguousarray(vals), off


def test_features_match_oracle():
    rng = np.random.default_rng(21)
    traces = [np.zeros(0), np.ones(1), np.array([1.0, -1.0]), rng.standard_normal(5),
              np.zeros(300), rng.standard_normal(255), rng.standard_normal(256),
              np.round(rng.standard_normal(700) * 2)]
    traces += [synth.chirp_mix(1, int(n), seed=int(n))[0] for n in (100, 1024, 2001, 4097)]
    v, off = _csr(traces)
    got = costmodel.features(torch.from_numpy(v), torch.from_numpy(off)).cpu().numpy()
    ref = C.features_batch(v, off)
    assert np.all(np.abs(got - ref) <= 1e-5 * np.maximum(1.0, np.abs(ref))), \
        np.max(np.abs(got - ref))


def _mlp_case(seed, M):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(C.N_PARAMS) * 0.2).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(12), rng.uniform(0.5, 2.0, 12)]).astype(np.float32)
    F = (rng.standard_normal((M, 12)) * 2).astype(np.float32)
    T = rng.standard_normal((M, 3)).astype(np.float32)
    tw = np.array([1.0, 0.5, 2.0], np.float32)
    return params, norm, F, T, tw


def _model(params, norm):
    cm = costmodel.CostModel()
    cm.params = torch.from_numpy(params).cuda()
    cm.norm = torch.from_numpy(norm).cuda()
    return cm


@pytest.mark.parametrize("M", [1, 63, 64, 200])
def test_forward_matches_oracle(M):
    params, norm, F, T, tw = _mlp_case(22 + M, M)
    got = _model(params, norm).forward_std(torch.from_numpy(F)).cpu().numpy()
    ref = C.forward(params.astype(np.float64), norm.astype(np.float64), F.astype(np.float64))[0]
    assert np.max(np.abs(got - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("M", [1, 100, 300])
def test_backward_matches_oracle(M):
    params, norm, F, T, tw = _mlp_case(30 + M, M)
    cm = _model(params, norm)
    g = cm.loss_grad(torch.from_numpy(F).cuda(), torch.from_numpy(T).cuda(),
                     torch.from_numpy(tw).cuda()).cpu().numpy().astype(np.float64)
    L, gref = C.loss_and_grad(params.astype(np.float64), norm.astype(np.float64),
                              F.astype(np.float64), T.astype(np.float64), tw.astype(np.float64))
    assert g[-1] == pytest.approx(L, rel=1e-5)
    assert np.max(np.abs(g[:-1] - gref)) <= 1e-5 * np.max(np.abs(gref))


def test_adam_matches_oracle():
    rng = np.random.default_rng(40)
    p0 = rng.standard_normal(C.N_PARAMS).astype(np.float32)
    cm = _model(p0.copy(), np.concatenate([np.zeros(12), np.ones(12)]).astype(np.float32))
    p, m, v = p0.astype(np.float64), np.zeros(C.N_PARAMS), np.zeros(C.N_PARAMS)
    for step in range(1, 4):
        g = rng.standard_normal(C.N_PARAMS + 1).astype(np.float32)
        cm.adam(torch.from_numpy(g).cuda(), 3e-3)
        p, m, v = C.adam(p, m, v, g[:-1].astype(np.float64), step, 3e-3)
    assert np.max(np.abs(cm.params.cpu().numpy() - p)) <= 1e-6


def _decompose(launcher, values, offsets):  # FIFO persistent launch
    out = launcher.run_ragged(torch.from_numpy(values).cuda(), torch.from_numpy(offsets),
                              threads=256, persistent=True)
    return out


def test_training_learns_the_method_counts():
    """Train on ragged mixed traces (seed 3), evaluate on held-out ones (seed 4): the loss
    falls and the predictions beat the constant (training-mean) predictor on every target."""
    launcher = EMDKernelConfig().build()
    res = {}
    for name, seed, B in (("train", 3, 3072), ("test", 4, 1024)):
        v, off, _ = synth.ragged_mix(B, seed=seed, n_lo=256, n_hi=4096)
        out = launcher.run_ragged(torch.from_numpy(v).cuda(), torch.from_numpy(off),
                                  threads=256, persistent=True)
        f = costmodel.features(torch.from_numpy(v).cuda(), torch.from_numpy(off).cuda())
        res[name] = (f, costmodel.targets_from(out[7], out[2]))
    cm = costmodel.CostModel(seed=0)
    hist = cm.fit(res["train"][0], res["train"][1], epochs=40, batch=256, lr=3e-3)
    pred = cm.predict(res["test"][0])
    true = res["test"][1]
    m_model = costmodel.mape(pred, true)
    const = np.broadcast_to(res["train"][1].mean(0), true.shape)
    m_const = costmodel.mape(const, true)
    print("loss", hist[0], "->", hist[-1], "MAPE model", m_model, "constant", m_const)
    assert hist[-1] < 0.5 * hist[0]
    for k in costmodel.TARGETS:
        assert m_model[k] < m_const[k], (k, m_model, m_const)


def test_learned_plan_leaves_outputs_identical():
    values, offsets, _ = synth.ragged_mix(512, seed=7, n_lo=64, n_hi=3000)
    launcher = EMDKernelConfig().build()
    base = _decompose(launcher, values, offsets)
    cm = costmodel.CostModel(seed=1)
    f = costmodel.features(torch.from_numpy(values).cuda(), torch.from_numpy(offsets).cuda())
    cm.fit(f, costmodel.targets_from(base[7], base[2]), epochs=5, batch=128)
    plan = cm.plan(f, np.diff(offsets))
    got = launcher.run_ragged(torch.from_numpy(values).cuda(), torch.from_numpy(offsets),
                              order=plan.order, threads=256, tile_slots=plan.tile_slots,
                              persistent=True)
    for k in range(2, 8):
        a, b = base[k].cpu().numpy(), got[k].cpu().numpy()
        if k == 7:
            a, b = a[:, :3], b[:, :3]  # flags column records which tile a sift used
        assert np.array_equal(a, b), k
    n_imfs = base[2].cpu().numpy()
    lengths = np.diff(offsets)
    ia, ib = base[0].cpu().numpy(), got[0].cpu().numpy()
    oo = base[1].cpu().numpy()
    for s in range(len(lengths)):
        seg = slice(int(oo[s]), int(oo[s] + n_imfs[s] * lengths[s]))
        assert np.array_equal(ia[seg], ib[seg])
