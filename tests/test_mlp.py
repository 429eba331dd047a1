"""GPU checks of the learned warm start (kh_mlp_* and kronspice_amd/warmstart.py):
the network's bf16 matrix-core forward pass and gradients against the numpy
oracle (oracle/mlp.py, which rounds the same operands to bf16), training, and
what the trained network gains over the cold start."""
import json

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _net(n_in, n_out, hidden, seed=7):
    from kronspice_amd import api
    rng = np.random.default_rng(seed)
    net = api.MLP(n_in, n_out, hidden=hidden, seed=seed)
    theta = rng.normal(0, 0.3, net.count).astype(np.float32)
    net.set_parameters(theta)
    return net, theta, rng


@pytest.mark.parametrize("n_in,n_out,hidden,B", [(60, 200, 256, 300), (5, 3, 64, 1), (70, 130, 128, 129)])
def test_forward_matches_oracle(n_in, n_out, hidden, B):
    """Shapes that exercise every padding (inputs, outputs, batch): the MFMA
    operand and result layouts and the strided tile loads are right when the
    product matches the oracle to FP32-accumulation round-off."""
    import torch
    net, theta, rng = _net(n_in, n_out, hidden)
    assert np.array_equal(net.parameters(), theta)
    x = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    yo, _, _ = omlp.forward(theta, x, n_in, hidden, n_out)
    assert np.max(np.abs(y - yo)) <= 1e-3 * np.max(np.abs(yo))


@pytest.mark.parametrize("n_in,n_out,hidden,B", [(60, 200, 256, 512), (7, 9, 64, 70)])
def test_gradients_match_oracle(n_in, n_out, hidden, B):
    import torch
    net, theta, rng = _net(n_in, n_out, hidden, seed=11)
    x = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    t = rng.normal(0, 1, (B, n_out)).astype(np.float32)
    loss = net.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = net.parameters(grads=True)
    lo, go = omlp.gradients(theta, x, t, n_in, hidden, n_out)
    assert abs(loss - lo) <= 1e-4 * lo
    off = 0
    for k, n in [(n_in, hidden), (hidden, hidden), (hidden, n_out)]:
        for size in (k * n, n):  # weights, then biases of each layer
            a, b = g[off:off + size], go[off:off + size]
            assert np.max(np.abs(a - b)) <= 2e-3 * np.max(np.abs(b)) + 1e-9, (off, size)
            off += size


def test_training_reduces_loss():
    import torch
    from kronspice_amd import api
    rng = np.random.default_rng(5)
    n_in, n_out, B = 8, 4, 2048
    A = rng.normal(0, 1, (n_in, n_out))
    x = rng.uniform(-1, 1, (B, n_in)).astype(np.float32)
    t = np.tanh(x @ A).astype(np.float32)
    net = api.MLP(n_in, n_out, hidden=64, seed=3)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    first = net.gradients(xd, td)
    for _ in range(300):
        net.gradients(xd, td)
        net.adam_step(3e-3)
    last = net.gradients(xd, td)
    assert last < 0.05 * first, (first, last)


def test_warm_start_saves_newton_iterations():
    """Trained on the amplifier's own operating points, the network's guesses
    (Options.NodeSet) take fewer Newton iterations than the cold start on
    fresh circuits, never fail where the cold start succeeds, and the polished
    answers agree to 1e-9."""
    from kronspice_amd import warmstart as WS
    data = WS.make_data(32768)
    assert np.mean(data.status == 0) > 0.99
    ws = WS.WarmStart(data.genes.shape[1], data.volts.shape[1])
    hist = ws.fit(data, steps=3000)
    assert np.mean(hist[-50:]) < 0.1 * hist[0]
    rep = WS.evaluate(ws, 4096)
    msg = json.dumps(rep)
    assert rep["warm_status_ok"] >= rep["cold_status_ok"], msg
    assert rep["iters_saved_fraction"] > 0.15, msg
    assert rep["max_rel_diff_polished"] <= 1e-9, msg
