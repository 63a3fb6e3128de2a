"""The guess network's gfx950 kernels against its CPU oracle (oracle/unet_oracle.py).

Bar: with the same bf16 weights, every stored bf16 tensor equals the oracle's rounding of the
same layer applied to the device's own input tensor, up to one bf16 unit in the last place at a
small fraction of entries (fp32 accumulation on the device, float64 in the oracle, rounded once
to bf16 by both); the fp32 output within 1e-5 relative; every weight and bias gradient within
2e-2 relative (Frobenius) of the oracle's backward through the device's forward tensors, with
cosine similarity above 0.9999.  Adam steps on the device lower the loss of a fixed pair."""
import numpy as np
import pytest

from oracle import unet_oracle as un

pytestmark = pytest.mark.gpu

TSPEC = [(0, 8), (0, 16), (0, 16), (1, 16), (1, 32), (1, 32), (2, 32), (2, 64), (2, 64), (3, 64), (3, 64),
         (3, 64), (2, 64), (2, 32), (1, 32), (1, 16), (0, 16), (0, 16)]
TNAMES = ["x0", "a0", "s0", "x1", "a1", "s1", "x2", "a2", "s2", "x3", "a3", "bb", "u2", "g2", "u1", "g1", "u0",
          "g0"]
LNAMES = [name for name, _, _ in un.LAYERS]


def levels(n):
    lv = [tuple(n)]
    for _ in range(3):
        lv.append(tuple(un.coarse_n(k) for k in lv[-1]))
    return lv


def net_with_params(n, seed=1):
    from saras_amd import guess
    rng = np.random.default_rng(seed)
    params = un.init_params(seed=seed)
    params = {k: (W.astype(np.float32).astype(np.float64), (rng.standard_normal(W.shape[1]) * 0.05).astype(
        np.float32).astype(np.float64)) for k, (W, b) in params.items()}
    net = guess.GuessNet(n)
    for l, name in enumerate(LNAMES):
        W, b = params[name]
        net.set_layer(l, W, b)
    return net, params


def device_tensors(net, n):
    lv = levels(n)
    return {name: net.tensor(i).astype(np.float64).reshape(lv[TSPEC[i][0]] + (TSPEC[i][1],))
            for i, name in enumerate(TNAMES)}


def assert_bf16_close(got, want, what):
    tol = 2.0 ** -7 * np.maximum(np.abs(got), np.abs(want)) + 1e-30
    bad = np.abs(got - want) > tol
    assert not bad.any(), (what, float(np.max(np.abs(got - want))))
    frac = float(np.mean(got != want))
    assert frac < 0.01, (what, frac)


def random_input(n, seed=3):
    rng = np.random.default_rng(seed)
    x0 = np.zeros(tuple(n) + (un.C_IN,), np.float32)
    x0[..., :3] = rng.standard_normal(tuple(n) + (3,))
    return x0


@pytest.mark.parametrize("n", [(17, 17, 17), (33, 17, 9)])
def test_forward_layers_match_oracle(n):
    net, params = net_with_params(n)
    x0 = random_input(n)
    y = net.forward(x0)
    T = device_tensors(net, n)
    Wq = {k: un.bf16(W).astype(np.float64) for k, (W, b) in params.items()}
    B = {k: b for k, (W, b) in params.items()}
    assert np.array_equal(T["x0"], un.bf16(x0).astype(np.float64))
    conv_io = [("e0a", "x0", "a0"), ("e0b", "a0", "s0"), ("e1a", "x1", "a1"), ("e1b", "a1", "s1"),
               ("e2a", "x2", "a2"), ("e2b", "a2", "s2"), ("b_a", "x3", "a3"), ("b_b", "a3", "bb"),
               ("d2", "u2", "g2"), ("d1", "u1", "g1"), ("d0a", "u0", "g0")]
    for lname, tin, tout in conv_io:
        want = un.bf16(un.relu(un.conv_fwd(T[tin], Wq[lname], B[lname]))).astype(np.float64)
        assert_bf16_close(T[tout], want, lname)
    for src, dst in (("s0", "x1"), ("s1", "x2"), ("s2", "x3")):
        assert_bf16_close(T[dst], un.bf16(un.restrict(T[src])).astype(np.float64), dst)
    for c, skip, dst in (("bb", "s2", "u2"), ("g2", "s1", "u1"), ("g1", "s0", "u0")):
        want = un.bf16(un.prolong(T[c], T[skip].shape) + T[skip]).astype(np.float64)
        assert_bf16_close(T[dst], want, dst)
    yo = un.conv_fwd(T["g0"], Wq["out"], B["out"])[..., 0]
    assert np.max(np.abs(y - yo)) <= 1e-5 * max(np.max(np.abs(yo)), 1e-30)
    # end to end against the fully rounded oracle: the rare one-unit flips stay small
    ye, _ = un.forward(params, x0.astype(np.float64), quantize=True)
    assert np.linalg.norm(y - ye) <= 2e-2 * np.linalg.norm(ye)
    net.close()


@pytest.mark.parametrize("n", [(17, 17, 17), (17, 33, 17)])
def test_backward_matches_oracle(n):
    net, params = net_with_params(n, seed=2)
    x0 = random_input(n, seed=4)
    net.forward(x0)
    T = device_tensors(net, n)
    dy = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    net.backward(dy)
    grads = un.backward(params, T, dy.astype(np.float64), quantize=True)
    for l, name in enumerate(LNAMES):
        dW, db = net.get_grad(l)
        eW, eb = grads[name]
        if name == "e0a":
            dW, eW = dW[:, :, :3], eW[:, :, :3]
        if name == "out":
            dW, eW, db, eb = dW[:, :1], eW[:, :1], db[:1], eb[:1]
        for got, want, what in ((dW, eW, "W"), (db, eb, "b")):
            nw = np.linalg.norm(want)
            assert nw > 0, (name, what)
            err = np.linalg.norm(got - want) / nw
            cos = float(np.sum(got * want) / (np.linalg.norm(got) * nw))
            assert err < 2e-2 and cos > 0.9999, (name, what, err, cos)
    net.close()


def test_adam_lowers_loss_on_a_fixed_pair():
    from saras_amd import guess
    n = (17, 17, 17)
    net = guess.GuessNet(n, seed=7, lr=2e-3)
    x0 = random_input(n, seed=8)
    ax = np.linspace(0, np.pi, 17)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    target = (np.sin(X) * np.sin(Y) * np.sin(Z)).astype(np.float32)
    losses = []
    for _ in range(40):
        y = net.forward(x0)
        r = y - target
        losses.append(float(np.mean(r * r)))
        net.backward((2.0 * r / r.size).astype(np.float32))
        net.adam()
    assert losses[-1] < 0.3 * losses[0], losses[::8]
    net.close()
