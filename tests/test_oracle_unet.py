"""Pins of the guess-network oracle (oracle/unet_oracle.py).  CPU only.

The network has no counterpart in the reference, so its oracle is pinned by brute force, by the
adjoint identities its backward pass must satisfy, by central finite differences of the whole
network in float64, and by the multigrid oracle's restriction and prolongation (which the
reference's testProlong / testTransfer KATs already pin)."""
import numpy as np
import pytest

from oracle import unet_oracle as un
from helpers import Problem
from oracle import pyoracle as po


def test_bf16_rounding_ties_to_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.float32([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0, 65504.0])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(un.bf16(x), ref)
    assert un.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0            # tie -> even mantissa


def test_conv_forward_brute_force():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, 4, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    out = un.conv_fwd(X, W, b)
    ref = np.zeros((3, 4, 5, 3))
    for i in range(3):
        for j in range(4):
            for k in range(5):
                for t, (dx, dy, dz) in enumerate(un.TAPS):
                    a, c, e = i + dx, j + dy, k + dz
                    if 0 <= a < 3 and 0 <= c < 4 and 0 <= e < 5:
                        ref[i, j, k] += W[t] @ X[a, c, e]
                ref[i, j, k] += b
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_conv_backward_adjoint_identities():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((5, 4, 6, 3))
    W = rng.standard_normal((27, 4, 3))
    Y = rng.standard_normal((5, 4, 6, 4))
    lhs = np.sum(un.conv_fwd(X, W) * Y)
    assert np.isclose(lhs, np.sum(X * un.conv_bwd_data(Y, W)), rtol=1e-12)
    dW, db = un.conv_bwd_weight(Y, X)
    assert np.isclose(lhs, np.sum(W * dW), rtol=1e-12)
    assert np.allclose(db, Y.reshape(-1, 4).sum(0))


def test_transfer_operators_match_multigrid_oracle():
    """Restriction = the reference coarsen with zero pads; prolongation = the reference prolong."""
    pb = Problem(ind=(4, 4, 4), domain="NNN", mesh="UUU", depth=2)
    o = pb.oracle()
    rng = np.random.default_rng(3)
    X = rng.standard_normal((17, 17, 17))
    T = o.array(po.T, 0)
    T[...] = 0.0
    T[1:-1, 1:-1, 1:-1] = X
    o.level = 0
    o.coarsen()
    got = o.array(po.R, 1)[1:-1, 1:-1, 1:-1]
    assert np.allclose(un.restrict(X[..., None])[..., 0], got, rtol=1e-14, atol=1e-14)
    Xc = rng.standard_normal((9, 9, 9))
    o.array(po.P, 1)[1:-1, 1:-1, 1:-1] = Xc
    o.level = 1
    o.prolong()
    got = o.array(po.P, 0)[1:-1, 1:-1, 1:-1]
    assert np.allclose(un.prolong(Xc[..., None], (17, 17, 17))[..., 0], got, rtol=1e-14, atol=1e-14)
    o.close()


def test_transfer_transposes():
    rng = np.random.default_rng(4)
    F = rng.standard_normal((9, 5, 17, 2))
    C = rng.standard_normal((5, 3, 9, 2))
    assert np.isclose(np.sum(un.restrict(F) * C), np.sum(F * un.restrict_T(C, F.shape)), rtol=1e-12)
    assert np.isclose(np.sum(un.prolong(C, F.shape) * F), np.sum(C * un.prolong_T(F)), rtol=1e-12)


def test_network_gradient_finite_differences():
    """Whole-network backward against central differences in float64 (no rounding)."""
    rng = np.random.default_rng(5)
    params = un.init_params(seed=6)
    # non-zero biases so that ReLU kinks are away from the sample
    params = {k: (W, rng.standard_normal(W.shape[1]) * 0.05) for k, (W, b) in params.items()}
    x0 = np.zeros((9, 9, 9, un.C_IN))
    x0[..., :3] = rng.standard_normal((9, 9, 9, 3))
    dy = rng.standard_normal((9, 9, 9))
    y, cache = un.forward(params, x0, quantize=False)
    grads = un.backward(params, cache, dy, quantize=False)

    def f(p):
        return float(np.sum(un.forward(p, x0, quantize=False)[0] * dy))

    eps = 1e-6
    checked = 0
    for name, ci, co in un.LAYERS:
        W, b = params[name]
        for _ in range(3):
            t, o_, i_ = rng.integers(27), rng.integers(co if name != "out" else 1), rng.integers(
                min(ci, 3) if name == "e0a" else ci)
            Wp, Wm = W.copy(), W.copy()
            Wp[t, o_, i_] += eps
            Wm[t, o_, i_] -= eps
            pp, pm = dict(params), dict(params)
            pp[name], pm[name] = (Wp, b), (Wm, b)
            fd = (f(pp) - f(pm)) / (2 * eps)
            an = grads[name][0][t, o_, i_]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (name, t, o_, i_, fd, an)
            checked += 1
        o_ = rng.integers(co if name != "out" else 1)
        bp, bm = b.copy(), b.copy()
        bp[o_] += eps
        bm[o_] -= eps
        pp, pm = dict(params), dict(params)
        pp[name], pm[name] = (W, bp), (W, bm)
        fd = (f(pp) - f(pm)) / (2 * eps)
        assert abs(fd - grads[name][1][o_]) <= 1e-6 * max(1.0, abs(fd)), (name, "bias", fd)
    assert checked == 3 * len(un.LAYERS)
