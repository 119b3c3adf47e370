"""Learned cost model: the GPU's FP32 gradient against an FP64 numpy
reference (analytic and finite differences), the BF16 MFMA inference against
a numpy forward pass, training progress, the model file, and the learned
schedule's invariance and fallback."""
import numpy as np
import pytest


def np_forward(P, x, dtype=np.float64, bf16=False):
    def rb(a):  # round to bfloat16 (nearest even), as the inference kernel does
        a = np.asarray(a, np.float32)
        u = a.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
        return (u.astype(np.uint32) << 16).view(np.float32).astype(dtype)

    xn = (x.astype(dtype) - P["mu"]) / P["sigma"]
    W1, W2 = P["W1"].astype(dtype), P["W2"].astype(dtype)
    if bf16:
        xn, W1, W2 = rb(xn), rb(W1), rb(W2)
    h1 = np.maximum(xn @ W1 + P["b1"], 0)
    if bf16:
        h1 = rb(h1)
    h2 = np.maximum(h1 @ W2 + P["b2"], 0)
    return xn, h1, h2, h2 @ P["w3"].astype(dtype) + P["b3"]


def np_grad(P, x, y):
    xn, h1, h2, yhat = np_forward(P, x)
    n = x.shape[0]
    dy = 2.0 * (yhat - y) / n
    dh2 = np.outer(dy, P["w3"]) * (h2 > 0)
    dh1 = (dh2 @ P["W2"].T) * (h1 > 0)
    g = np.concatenate([(xn.T @ dh1).ravel(), dh1.sum(0), (h1.T @ dh2).ravel(), dh2.sum(0), h2.T @ dy, [dy.sum()]])
    return g, float(np.mean((yhat - y) ** 2))


def data(rows, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(rows, 16)).astype(np.float32) * rng.uniform(0.5, 3, 16).astype(np.float32) + 2
    x[:, 15] = 1.0
    y = (np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2] - 0.1 * x[:, 3] ** 2).astype(np.float32)
    return x, y


def test_init_save_load(mce, tmp_path):
    x, _ = data(500, 1)
    m = mce.mlp_init(x, seed=3)
    P = mce.mlp_params(m)
    np.testing.assert_allclose(P["mu"], x.mean(0), rtol=1e-4, atol=1e-4)
    assert np.all(P["sigma"] > 0)
    p = str(tmp_path / "m.bin")
    mce.mlp_save(p, m)
    m2 = mce.mlp_load(p)
    assert bytes(m2) == bytes(m)
    raw = bytearray(open(p, "rb").read())
    raw[20] ^= 0xFF
    open(p, "wb").write(bytes(raw[:-7]))
    with pytest.raises(mce.MCEError) as e:
        mce.mlp_load(p)
    assert e.value.code == -7


@pytest.mark.gpu
def test_gradient_matches_fp64(mce):
    x, y = data(3000, 2)
    m = mce.mlp_init(x, seed=5)
    P = {k: (v.astype(np.float64) if isinstance(v, np.ndarray) else v) for k, v in mce.mlp_params(m).items()}
    g, loss = mce.mlp_grad(m, x, y)
    ref, ref_loss = np_grad(P, x.astype(np.float64), y.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-4 * max(1.0, ref_loss)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=2e-3, atol=2e-4 * scale)
    # central finite differences on a few parameters (float64 forward)
    flat = np.ctypeslib.as_array(m.params).astype(np.float64)
    rng = np.random.default_rng(0)
    for i in rng.choice(flat.size, 12, replace=False):
        def loss_at(val):
            q = flat.copy()
            q[i] = val
            Q = dict(P)
            o = [0, 1024, 1088, 5184, 5248, 5312]
            Q.update(W1=q[o[0]:o[1]].reshape(16, 64), b1=q[o[1]:o[2]], W2=q[o[2]:o[3]].reshape(64, 64),
                     b2=q[o[3]:o[4]], w3=q[o[4]:o[5]], b3=q[o[5]])
            return float(np.mean((np_forward(Q, x.astype(np.float64))[3] - y) ** 2))
        h = 1e-4 * max(1.0, abs(flat[i]))
        fd = (loss_at(flat[i] + h) - loss_at(flat[i] - h)) / (2 * h)
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(fd)) + 2e-4 * scale, (i, fd, g[i])


@pytest.mark.gpu
def test_bf16_mfma_inference(mce):
    x, y = data(10007, 3)  # not a multiple of the 16-row tile
    m = mce.mlp_init(x, seed=7)
    mce.mlp_train(m, x, y, epochs=3, batch=2048, lr=3e-3)
    P = mce.mlp_params(m)
    out = mce.mlp_predict(m, x)
    ref_bf = np_forward(P, x, dtype=np.float32, bf16=True)[3]
    ref_fp = np_forward(P, x, dtype=np.float64)[3]
    # same bf16 rounding of inputs, weights and the hidden layer: only the
    # FP32 accumulation order differs
    np.testing.assert_allclose(out, ref_bf, rtol=2e-3, atol=2e-3 * np.abs(ref_bf).max())
    # against exact FP64, within bf16 precision (8 mantissa bits)
    assert np.abs(out - ref_fp).max() <= 0.05 * np.abs(ref_fp).max() + 0.05


@pytest.mark.gpu
def test_training_reduces_loss(mce):
    x, y = data(20000, 4)
    m = mce.mlp_init(x, seed=9)
    _, l0 = mce.mlp_grad(m, x, y)
    l1 = mce.mlp_train(m, x, y, epochs=15, batch=4096, lr=3e-3, seed=2)
    _, l2 = mce.mlp_grad(m, x, y)
    assert l2 < 0.2 * l0 and np.isfinite(l1)
    # deterministic for a given seed
    m2 = mce.mlp_init(x, seed=9)
    mce.mlp_train(m2, x, y, epochs=15, batch=4096, lr=3e-3, seed=2)
    assert bytes(m2) == bytes(m)


def np_adam(p, grads, lr, wd, bias_mask, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma and Ba, 2015) with bias correction and L2 weight decay
    added to the gradient of the weights (not the biases), in FP64."""
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t, g in enumerate(grads, 1):
        g = g + np.where(bias_mask, 0.0, wd * p)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p


def flat(P):
    return np.concatenate([P["W1"].ravel(), P["b1"], P["W2"].ravel(), P["b2"], P["w3"], [P["b3"]]]).astype(np.float64)


@pytest.mark.gpu
def test_adam_steps_match_numpy(mce):
    """One and two full-batch optimiser steps of the GPU trainer against an
    FP64 numpy Adam on the FP64 gradient: after one step every weight moves by
    lr * g / (|g| + eps), so the directions must agree wherever the gradient
    is not vanishingly small; after two steps the bias-corrected moments set
    the size of the second move."""
    x, y = data(4096, 11)
    lr, wd = 1e-3, 1e-2
    m0 = mce.mlp_init(x, seed=5)
    P0 = mce.mlp_params(m0)
    p0 = flat(P0)
    sizes = [16 * 64, 64, 64 * 64, 64, 64, 1]
    bias = np.concatenate([np.full(k, b) for k, b in zip(sizes, (False, True, False, True, False, True))])
    g1, _ = np_grad(P0, x, y.astype(np.float64))
    m1 = mce.mlp_init(x, seed=5)
    mce.mlp_train(m1, x, y, epochs=1, batch=x.shape[0], lr=lr, weight_decay=wd, seed=1)
    p1 = flat(mce.mlp_params(m1))
    want1 = np_adam(p0, [g1], lr, wd, bias)
    gw = np.abs(g1 + np.where(bias, 0.0, wd * p0))
    ok = gw > 1e-4
    assert ok.mean() > 0.9
    np.testing.assert_allclose(p1[ok], want1[ok], rtol=0, atol=1e-6)
    assert np.all(np.abs(p1 - p0) <= lr * (1 + 1e-5))
    # second step from the numpy-updated weights
    P1 = dict(P0)
    o = np.cumsum([0] + sizes)
    w = want1
    P1.update(W1=w[o[0]:o[1]].reshape(16, 64), b1=w[o[1]:o[2]], W2=w[o[2]:o[3]].reshape(64, 64), b2=w[o[3]:o[4]],
              w3=w[o[4]:o[5]], b3=float(w[o[5]]))
    g2, _ = np_grad(P1, x, y.astype(np.float64))
    m2 = mce.mlp_init(x, seed=5)
    mce.mlp_train(m2, x, y, epochs=2, batch=x.shape[0], lr=lr, weight_decay=wd, seed=1)
    p2 = flat(mce.mlp_params(m2))
    want2 = np_adam(p0, [g1, g2], lr, wd, bias)
    g2w = np.abs(g2 + np.where(bias, 0.0, wd * want1))
    ok2 = ok & (g2w > 1e-4)
    np.testing.assert_allclose(p2[ok2], want2[ok2], rtol=0, atol=1e-5)


@pytest.mark.gpu
def test_learned_schedule_is_result_invariant(mce, oracle, tmp_path):
    g = mce.gen_rmat(12, 16, seed=13)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        F = dg.features()
        assert F.shape == (g.n, 16) and np.all(np.isfinite(F)) and np.all(F[:, 15] == 1)
        lab = dg.count(level=1, induced=2, sub_nodes=True)
        y = np.log1p(lab["sub_nodes"].astype(np.float64)).astype(np.float32)
        m = mce.mlp_init(F, seed=1)
        mce.mlp_train(m, F, y, epochs=10, batch=1024, lr=3e-3)
        for wl in (1, 2):
            # whole vertex subproblems only: the same tree as the oracle's
            r = dg.count(level=1, induced=2, worklist=wl, model=m)
            assert r["fallback_events"] == 0 and r["split_vertices"] == 0
            assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
                ref["count"], ref["hash"], ref["hist"], ref["tree_nodes"])
            # predicted-heavy vertices as level-2 arc subproblems (the cut
            # picked automatically): other trees, the same cliques
            r = dg.count(level=1, induced=2, worklist=wl, model=m, split_heavy=-1)
            assert r["fallback_events"] == 0
            assert (r["count"], r["hash"], r["hist"]) == (ref["count"], ref["hash"], ref["hist"])
        # a corrupt or missing model falls back to the plain schedule, logged
        bad = tmp_path / "bad.bin"
        bad.write_bytes(b"garbage")
        r = dg.count(level=1, induced=2, model_path=str(bad))
        assert r["fallback_events"] == 1 and r["count"] == ref["count"]
        r = dg.count(level=1, induced=2, model_path=str(tmp_path / "none.bin"))
        assert r["fallback_events"] == 1 and r["count"] == ref["count"]


@pytest.mark.gpu
def test_features_match_numpy(mce, oracle):
    g = mce.gen_rmat(10, 8, seed=3)
    with mce.DeviceGraph(g) as dg:
        info = dg.order()
        F = dg.features()
        prio, split, oc = dg.download()
    rp = g.row_ptr.astype(np.int64)
    dplus = rp[1:] - split.astype(np.int64)
    for v in range(0, g.n, 41):
        s, e = int(split[v]), int(rp[v + 1])
        P = oc[s:e]
        Pset = set(P.tolist())
        eP = sum(len(Pset.intersection(oc[int(split[u]):int(rp[u + 1])].tolist())) for u in P)
        p = e - s
        assert abs(F[v, 1] - np.log1p(p)) < 1e-5
        assert abs(F[v, 5] - np.log1p(dplus[P].sum())) < 1e-4
        assert abs(F[v, 7] - np.log1p(eP)) < 1e-4
        if p > 1:
            assert abs(F[v, 8] - 2 * eP / (p * (p - 1))) < 1e-5


@pytest.mark.gpu
def test_auto_variant_is_result_invariant(mce, oracle):
    """-w auto: the committed SBM model predicts no heavy tail on an SBM batch
    (no worker list, so no donations); any pick leaves the results alone."""
    import os
    model = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mce-gpu_amd",
                         "cost_model_sbm.mlp")
    off, g = mce.gen_sbm_batch(48, n_min=64, n_max=700, blocks=4, p_in=0.3, p_out=0.01, seed=21)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        r = dg.count(level=1, induced=2, worklist=mce.WL_AUTO, model_path=model)
        assert r["fallback_events"] == 0 and r["donations"] == 0
        assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
            ref["count"], ref["hash"], ref["hist"], ref["tree_nodes"])
        # the plain schedule under auto takes the worker list
        r = dg.count(level=1, induced=2, worklist=mce.WL_AUTO)
        assert (r["count"], r["hash"]) == (ref["count"], ref["hash"])
    g = mce.gen_rmat(12, 16, seed=13)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        r = dg.count(level=1, induced=2, worklist=mce.WL_AUTO, model_path=model)
        assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
            ref["count"], ref["hash"], ref["hist"], ref["tree_nodes"])


@pytest.mark.gpu
def test_split_heavy_vertices(mce, oracle):
    """A model that predicts one vertex far heavier than the rest splits it
    into its arc subproblems; the per-root clique counts still match."""
    g = mce.gen_rmat(11, 16, seed=4)
    ref = oracle.mce(g, level=1, ipx=True, sub_cliques=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        F = dg.features()
        lab = dg.count(level=1, induced=2, sub_nodes=True)
        y = np.log1p(lab["sub_nodes"].astype(np.float64)).astype(np.float32)
        m = mce.mlp_init(F, seed=2)
        mce.mlp_train(m, F, y, epochs=30, batch=512, lr=3e-3)
        for wl in (1, 2):
            r = dg.count(level=1, induced=2, worklist=wl, model=m, sub_cliques=True, split_heavy=64)
            assert r["split_vertices"] > 0
            assert (r["count"], r["hash"], r["hist"]) == (ref["count"], ref["hash"], ref["hist"])
            assert np.array_equal(r["sub_cliques"], ref["sub_cliques"])
