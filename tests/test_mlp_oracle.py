"""The FP64 network oracle against finite differences and a closed form (CPU)."""
import numpy as np

from oracle import mlp_ref


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = [5, 7, 6, 1]
    count = sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))
    p = rng.standard_normal(count) * 0.5
    X = rng.standard_normal((11, 5))
    t = rng.standard_normal(11)
    loss, g = mlp_ref.loss_and_grad(p, widths, X, t)
    h = 1e-6
    for k in rng.choice(count, 25, replace=False):
        e = np.zeros(count)
        e[k] = h
        num = (mlp_ref.loss_and_grad(p + e, widths, X, t)[0] - mlp_ref.loss_and_grad(p - e, widths, X, t)[0]) / (2 * h)
        assert abs(num - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, num, g[k])


def test_linear_model_closed_form():
    """One linear layer: gradient of mean((Xw + b - t)^2) is 2/n X^T r, 2/n sum r."""
    rng = np.random.default_rng(1)
    X = rng.standard_normal((9, 3))
    w = rng.standard_normal(3)
    b = 0.3
    t = rng.standard_normal(9)
    p = np.concatenate([w, [b]])
    loss, g = mlp_ref.loss_and_grad(p, [3, 1], X, t)
    r = X @ w + b - t
    assert np.isclose(loss, np.mean(r * r))
    assert np.allclose(g[:3], 2.0 / 9 * X.T @ r)
    assert np.isclose(g[3], 2.0 / 9 * r.sum())


def test_shipped_cost_model_on_held_out_expressions():
    """The shipped cost model against the exhaustive autotuner's own timings (the saved sweep
    of the 512 C5 expressions, seed 5001; the model saw the first 384): on the 128 held-out
    expressions its picks cost at most 10% more than the best configuration of each, among
    all 45 configurations and among the 9 with 256 threads that the per-expression launches
    use, and among the latter no more than the untuned default (V=4, U=1). The network is
    evaluated by the FP64 oracle's forward pass on the saved parameters."""
    import gzip
    import json
    import os

    from compute import Tensors, costmodel, workloads
    data = os.path.join(os.path.dirname(costmodel.__file__), "data")
    sweep = json.load(gzip.open(os.path.join(data, "c5_timings.json.gz")))
    model = np.load(os.path.join(data, "costmodel.npz"))
    widths = [int(w) for w in model["widths"]]
    cfgs = sweep["configs"]
    times = {(i, c): us for i, c, us in sweep["times"]}
    exprs = workloads.uneven_dags(Tensors(offline=True), 512, 5001)
    default = cfgs.index({"vector_width": 4, "items_per_thread": 1, "block_size": 256})
    b256 = [c for c, cfg in enumerate(cfgs) if cfg["block_size"] == 256]
    for cand in (list(range(len(cfgs))), b256):
        best = chosen = untuned = 0.0
        for i in range(384, 512):
            X = np.stack([costmodel.features(exprs[i], cfgs[c]) for c in cand])
            pred = mlp_ref.forward(model["params"].astype(np.float64), widths, (X - model["mean"]) / model["std"])[-1]
            best += min(times[(i, c)] for c in cand)
            chosen += times[(i, cand[int(np.argmin(pred[:, 0]))])]
            untuned += times[(i, default)]
        assert chosen <= 1.10 * best, (len(cand), chosen / best)
        if cand is b256:
            assert chosen <= untuned, (chosen, untuned)
