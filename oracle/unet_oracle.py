"""CPU ORACLE -- test infrastructure only.  Nothing in the product path imports this module; tests
and the benchmark's checks use it to verify the device kernels of the learned initial guess.

A numpy restatement of the guess network: a 4-level 3D U-Net on node-centred grids
(n, (n-1)/2+1, ...), channels-last tensors of shape (nx, ny, nz, C), 3x3x3 convolutions with zero
padding, ReLU, fixed full-weighting restriction and trilinear prolongation between levels, skip
connections by addition, a linear single-channel output.  The reference solver has no such
network (it starts every pressure solve from the previous step's correction), so nothing in
roshansamuel/saras pins it; the functions here are pinned instead by brute force on tiny grids,
by adjoint identities, by central finite differences of the whole network in float64, and (for
the transfer operators) by the multigrid oracle's own restriction and prolongation.

With quantize=True every tensor the device stores in bfloat16 (weights, activations, masked
gradients) is rounded to bfloat16 at the same point, so that device results can be compared
layer by layer; accumulation is float64 here and float32 on the device.
"""
import numpy as np

# tap t = (dx+1)*9 + (dy+1)*3 + (dz+1): input voxel v + (dx, dy, dz) feeds output voxel v
TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]

C_IN = 8          # input channels as stored (3 used: previous two pressures, divergence)
C_OUT = 16        # output channels of the last convolution (1 used)
LAYERS = [("e0a", C_IN, 16), ("e0b", 16, 16),
          ("e1a", 16, 32), ("e1b", 32, 32),
          ("e2a", 32, 64), ("e2b", 64, 64),
          ("b_a", 64, 64), ("b_b", 64, 64),
          ("d2", 64, 32), ("d1", 32, 16), ("d0a", 16, 16), ("out", 16, C_OUT)]


def bf16(x):
    """Round to the nearest bfloat16 (ties to even); returned as float32 values."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _q(x, quantize):
    return bf16(x).astype(np.float64) if quantize else x


def coarse_n(n):
    return (n - 1) // 2 + 1


def shift(X, d):
    """Y[v] = X[v + d] inside the grid, 0 where v + d falls outside."""
    Y = np.zeros_like(X)
    src, dst = [], []
    for ax, s in enumerate(d):
        n = X.shape[ax]
        if s >= 0:
            src.append(slice(s, n))
            dst.append(slice(0, n - s))
        else:
            src.append(slice(0, n + s))
            dst.append(slice(-s, n))
    Y[tuple(dst)] = X[tuple(src)]
    return Y


def conv_fwd(X, W, b=None):
    """out[v, co] = b[co] + sum_t sum_ci W[t, co, ci] X[v + tap_t, ci]; W has shape (27, co, ci)."""
    out = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, d in enumerate(TAPS):
        out += shift(X, d) @ W[t].T
    if b is not None:
        out += b
    return out


def conv_bwd_data(dY, W):
    """The adjoint of conv_fwd in X: dX[v, ci] = sum_t sum_co W[t, co, ci] dY[v - tap_t, co]."""
    out = np.zeros(dY.shape[:3] + (W.shape[2],))
    for t, d in enumerate(TAPS):
        out += shift(dY, tuple(-s for s in d)) @ W[t]
    return out


def conv_bwd_weight(dY, X):
    """dW[t, co, ci] = sum_v dY[v, co] X[v + tap_t, ci];  db[co] = sum_v dY[v, co]."""
    co, ci = dY.shape[3], X.shape[3]
    dYf = dY.reshape(-1, co)
    dW = np.zeros((27, co, ci))
    for t, d in enumerate(TAPS):
        dW[t] = dYf.T @ shift(X, d).reshape(-1, ci)
    return dW, dYf.sum(axis=0)


def _restrict_1d(nf):
    nc = coarse_n(nf)
    R = np.zeros((nc, nf))
    for i in range(nc):
        R[i, 2 * i] = 0.5
        if 2 * i - 1 >= 0:
            R[i, 2 * i - 1] = 0.25
        if 2 * i + 1 < nf:
            R[i, 2 * i + 1] = 0.25
    return R


def _prolong_1d(nf):
    nc = coarse_n(nf)
    P = np.zeros((nf, nc))
    for i in range(nf):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, i // 2] = 0.5
            P[i, i // 2 + 1] = 0.5
    return P


def _apply3(X, mats):
    """Apply a 1D linear map along each of the three grid axes."""
    Y = np.einsum("ai,ijkh->ajkh", mats[0], X)
    Y = np.einsum("bj,ajkh->abkh", mats[1], Y)
    return np.einsum("ck,abkh->abch", mats[2], Y)


def restrict(X):
    """27-point full weighting (1/8, 1/16, 1/32, 1/64) onto every second node, zero outside."""
    return _apply3(X, [_restrict_1d(n) for n in X.shape[:3]])


def restrict_T(dC, fine_shape):
    return _apply3(dC, [_restrict_1d(n).T for n in fine_shape[:3]])


def prolong(Xc, fine_shape):
    """Trilinear interpolation from every second node: injection at even indices, means of 2, 4
    or 8 coarse neighbours elsewhere."""
    return _apply3(Xc, [_prolong_1d(n) for n in fine_shape[:3]])


def prolong_T(dF):
    return _apply3(dF, [_prolong_1d(n).T for n in dF.shape[:3]])


def relu(x):
    return np.maximum(x, 0.0)


def init_params(seed=0, scale=1.0):
    """He-style initial weights, zero biases: {name: (W (27, co, ci), b (co,))}."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co in LAYERS:
        std = scale * np.sqrt(2.0 / (27 * ci))
        W = rng.standard_normal((27, co, ci)) * std
        if name == "out":
            W[:, 1:, :] = 0.0
            W *= 0.1
        if name == "e0a":
            W[:, :, 3:] = 0.0
        params[name] = (W, np.zeros(co))
    return params


def forward(params, x0, quantize=True):
    """x0: (n, n, n, C_IN) input; returns (y (n, n, n) float, cache of stored tensors)."""
    Q = lambda a: _q(a, quantize)
    Wq = {k: (Q(W), b) for k, (W, b) in params.items()}
    c = {"x0": Q(x0)}

    def layer(name, X, act=True):
        W, b = Wq[name]
        y = conv_fwd(X, W, b)
        return Q(relu(y)) if act else y

    c["a0"] = layer("e0a", c["x0"])
    c["s0"] = layer("e0b", c["a0"])
    c["x1"] = Q(restrict(c["s0"]))
    c["a1"] = layer("e1a", c["x1"])
    c["s1"] = layer("e1b", c["a1"])
    c["x2"] = Q(restrict(c["s1"]))
    c["a2"] = layer("e2a", c["x2"])
    c["s2"] = layer("e2b", c["a2"])
    c["x3"] = Q(restrict(c["s2"]))
    c["a3"] = layer("b_a", c["x3"])
    c["bb"] = layer("b_b", c["a3"])
    c["u2"] = Q(prolong(c["bb"], c["s2"].shape) + c["s2"])
    c["g2"] = layer("d2", c["u2"])
    c["u1"] = Q(prolong(c["g2"], c["s1"].shape) + c["s1"])
    c["g1"] = layer("d1", c["u1"])
    c["u0"] = Q(prolong(c["g1"], c["s0"].shape) + c["s0"])
    c["g0"] = layer("d0a", c["u0"])
    y = layer("out", c["g0"], act=False)[..., 0]
    return y, c


def backward(params, cache, dy, quantize=True):
    """Gradients of sum(dy * y) with respect to every weight and bias: {name: (dW, db)}."""
    Q = lambda a: _q(a, quantize)
    Wq = {k: Q(W) for k, (W, b) in params.items()}
    grads = {}
    c = cache

    def back(name, g, X, need_dx=True):
        dW, db = conv_bwd_weight(g, X)
        grads[name] = (dW, db)
        return conv_bwd_data(g, Wq[name]) if need_dx else None

    def mask(d, Y):
        return Q(d * (Y > 0))

    dy16 = np.zeros(dy.shape + (C_OUT,))
    dy16[..., 0] = dy
    dy16 = Q(dy16)
    d = back("out", dy16, c["g0"])
    d = back("d0a", mask(d, c["g0"]), c["u0"])           # d/du0
    ds0 = d.copy()
    d = back("d1", mask(prolong_T(d), c["g1"]), c["u1"])  # d/du1
    ds1 = d.copy()
    d = back("d2", mask(prolong_T(d), c["g2"]), c["u2"])  # d/du2
    ds2 = d.copy()
    d = back("b_b", mask(prolong_T(d), c["bb"]), c["a3"])
    d = back("b_a", mask(d, c["a3"]), c["x3"])
    ds2 = ds2 + restrict_T(d, c["s2"].shape)
    d = back("e2b", mask(ds2, c["s2"]), c["a2"])
    d = back("e2a", mask(d, c["a2"]), c["x2"])
    ds1 = ds1 + restrict_T(d, c["s1"].shape)
    d = back("e1b", mask(ds1, c["s1"]), c["a1"])
    d = back("e1a", mask(d, c["a1"]), c["x1"])
    ds0 = ds0 + restrict_T(d, c["s0"].shape)
    d = back("e0b", mask(ds0, c["s0"]), c["a0"])
    back("e0a", mask(d, c["a0"]), c["x0"], need_dx=False)
    return grads


def mse_loss(y, target):
    r = y - target
    return float(np.mean(r * r)), 2.0 * r / r.size
