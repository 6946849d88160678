"""The cost model's GPU MLP (BF16 MFMA forward/backward, FP32 accumulation,
Adam) against the CPU restatement in oracle/mlp_ref.py."""

import os
import sys

import numpy as np
import pytest

from bitstreamzoo_amd.costmodel import CostModel, GpuMlp
from hip_models import ctx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu

DIMS = [(28, 64, 64), (5, 20, 9), (33, 16, 48)]


def _net(dims, seed=3):
    n_in, h1, h2 = dims
    net = GpuMlp(ctx(), n_in, h1, h2, seed)
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=net.n_params) * 0.3).astype(np.float32)
    net.set_params(p)
    return net, p, rng


def test_parameter_roundtrip():
    net, p, _ = _net(DIMS[0])
    assert net.n_params == mlp_ref.n_params(*DIMS[0])
    assert np.array_equal(net.get_params(), p)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("n", [1, 37, 300])
def test_forward_matches_cpu(dims, n):
    net, p, rng = _net(dims)
    X = rng.normal(size=(n, dims[0])).astype(np.float32)
    got = net.predict(X)
    _, _, _, ref = mlp_ref.forward(p, X, *dims, bf16=True)
    # bf16 operands, fp32 accumulation in a different order: a flipped bf16
    # rounding of an activation moves the output by ~2^-8 relative
    assert np.allclose(got, ref, rtol=2e-2, atol=2e-2 * (1 + np.abs(ref).mean())), np.abs(got - ref).max()


@pytest.mark.parametrize("dims", DIMS)
def test_gradients_match_cpu(dims):
    net, p, rng = _net(dims)
    X = rng.normal(size=(200, dims[0])).astype(np.float32)
    y = rng.normal(size=200).astype(np.float32)
    loss, g = net.gradients(X, y)
    ref_loss, ref_g = mlp_ref.loss_and_grad(p, X, y, *dims, bf16=True)
    assert abs(loss - ref_loss) <= 2e-2 * (1 + ref_loss)
    cos = float(np.dot(g, ref_g) / (np.linalg.norm(g) * np.linalg.norm(ref_g)))
    assert cos > 0.999, cos
    # per parameter block
    n_in, h1, h2 = dims
    sizes = [n_in * h1, h1, h1 * h2, h2, h2, 1]
    o = 0
    for s in sizes:
        a, b = g[o:o + s], ref_g[o:o + s]
        assert np.linalg.norm(a - b) <= 3e-2 * (np.linalg.norm(b) + 1e-3), (o, s)
        o += s


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(2000, 8)).astype(np.float32)
    y = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * np.abs(X[:, 3])).astype(np.float32)
    net = GpuMlp(ctx(), 8, 64, 64, seed=1)
    before = float(np.mean((net.predict(X) - y) ** 2))
    net.train(X, y, epochs=60, batch=128, lr=3e-3, seed=2)
    after = float(np.mean((net.predict(X) - y) ** 2))
    assert after < before / 10 and after < 0.02, (before, after)


def test_cost_model_ranks_by_a_known_law():
    # runtime law over 6 features; the model must recover the ranking
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 4, size=(1500, 6))
    ms = np.exp(0.5 * X[:, 0] - 0.3 * X[:, 1] + 0.2 * X[:, 2] * X[:, 3] + 0.05 * rng.normal(size=1500))
    m = CostModel(ctx(), hidden=32, members=3, seed=0)
    m.fit(X[:1200], ms[:1200], epochs=150, batch=64, lr=3e-3)
    pred = m.predict_ms(X[1200:])
    err = np.abs(np.log(pred) - np.log(ms[1200:]))
    assert np.median(err) < 0.1, np.median(err)
    r = np.corrcoef(np.argsort(np.argsort(pred)), np.argsort(np.argsort(ms[1200:])))[0, 1]
    assert r > 0.95, r


def test_trained_model_beats_the_default_schedule_on_held_out_programs():
    """The gain over the method without the model: the cost model, trained on
    the committed autotuner measurements, ranks 32 sampled schedules of each
    held-out generated program from plan features alone, and its pick (the
    fastest of the three best-predicted schedules that compile within their
    register file) runs faster than the untuned default schedule on every
    program and in total; every pick gives the default schedule's outputs."""
    from bitstreamzoo_amd import api, costmodel, schedsearch, tuning
    X, ms = schedsearch.load_training_data()
    model = costmodel.CostModel(ctx(), seed=0)
    model.fit(X, ms)
    cands = schedsearch.sample_space(32, 4321)
    specs = schedsearch.random_specs(3, 99, nodes=(800, 1600), log2_bytes=(20, 21))
    built = [sp.build() for sp in specs]
    c = ctx()
    picks = []
    for p, case in built:
        F = costmodel.features_batch(api.Program(p.source), tuning.shape_case([case]), cands,
                                     (case.bitlength + 63) // 64, 1)
        picks.append(schedsearch.model_pick(c, p.source, case, cands, model.predict_ms(F)).options[0])
    tuning.precompile_many([(p.source, [case], {}) for p, case in built], processes=16)
    t_pick = t_default = 0.0
    slower = []
    for (p, case), o in zip(built, picks):
        res = tuning.Resident(c, [case])
        prog, shape = api.Program(p.source), tuning.shape_case([case])
        pl = api.Plan(c, prog, shape, o)
        d = api.Plan(c, prog, shape, {})
        res.bind(pl)
        res.bind(d)
        nw = (case.bitlength + 63) // 64
        pl.run()
        d.run()
        c.sync()
        for out in p.outputs:
            assert np.array_equal(pl.download_output(0, out, nw), d.download_output(0, out, nw)), out
        tp = tuning.time_runs(c, [pl], 10, 2)
        td = tuning.time_runs(c, [d], 10, 2)
        if tp >= td:
            slower.append((o, tp, td))
        t_pick += tp
        t_default += td
        pl.close()
        d.close()
        res.free()
    assert not slower, slower
    assert t_pick < t_default, (t_pick, t_default)
