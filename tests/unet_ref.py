"""FP64 numpy reference of the warm-start network's operations (test
infrastructure: the checker of the agdu_* kernels). Channels-last arrays
[n][d][h][w][c]; conv weights [27][co][ci], tap = (kd * 3 + kh) * 3 + kw."""
import numpy as np


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3(x, w, b=None, relu=False):
    n, D, H, W, ci = x.shape
    co = w.shape[1]
    xp = np.zeros((n, D + 2, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((n, D, H, W, co))
    for kd in range(3):
        for kh in range(3):
            for kw in range(3):
                tap = (kd * 3 + kh) * 3 + kw
                y += np.einsum("ndhwi,oi->ndhwo", xp[:, kd:kd + D, kh:kh + H, kw:kw + W],
                               w[tap])
    if b is not None:
        y += b
    if relu:
        y = np.maximum(y, 0.0)
    return y


# Combined taps over a nearest x2 upsampled input: along one axis, output
# parity p and pair index a read the fine row o - 1 + p + a, which carries
# the taps UP_SETS[p, a] of the 3-tap kernel.
UP_SETS = {(0, 0): (0,), (0, 1): (1, 2), (1, 0): (0, 1), (1, 1): (2,)}


def up_tap(pd, ph, a, b, kw):
    return ((pd * 2 + ph) * 4 + a * 2 + b) * 3 + kw


def up_weights(w, cu, rounded=True):
    """[48][co][cu] combined taps of the first cu input channels of w
    [27][co][ci]: per (pd, ph, a, b, kw) the sum over kd in UP_SETS[pd, a]
    and kh in UP_SETS[ph, b] of w[(kd * 3 + kh) * 3 + kw], in float32 in that
    order and rounded to bf16 (the kernel's arithmetic), or exact in FP64
    with rounded=False."""
    co = w.shape[1]
    out = np.zeros((48, co, cu))
    for pd in range(2):
        for ph in range(2):
            for a in range(2):
                for b in range(2):
                    for kw in range(3):
                        taps = [(kd * 3 + kh) * 3 + kw for kd in UP_SETS[pd, a]
                                for kh in UP_SETS[ph, b]]
                        if rounded:
                            acc = np.zeros((co, cu), np.float32)
                            for t in taps:
                                acc = (acc + np.asarray(w[t][:, :cu], np.float32)).astype(
                                    np.float32)
                            out[up_tap(pd, ph, a, b, kw)] = bf16_round(acc)
                        else:
                            out[up_tap(pd, ph, a, b, kw)] = sum(
                                np.asarray(w[t][:, :cu], np.float64) for t in taps)
    return out


def conv3_up(xu, xs, w, wcu, b=None, relu=False):
    """conv3 of [upsample(xu), xs] (D, H, W even) with xu's channels on the
    combined taps wcu (up_weights) and xs's on w's."""
    n, D, H, W, cs = xs.shape
    cu = xu.shape[-1]
    y = conv3(xs, np.asarray(w)[:, :, cu:])
    up = np.zeros((n, D + 2, H + 2, W + 2, cu))
    up[:, 1:-1, 1:-1, 1:-1] = upsample(xu)
    for pd in range(2):
        for ph in range(2):
            for ia in range(2):
                for ib in range(2):
                    for kw in range(3):
                        # outputs d = pd + 2 k read padded plane d + pd + ia
                        xs_ = up[:, 2 * pd + ia:2 * pd + ia + D:2,
                                 2 * ph + ib:2 * ph + ib + H:2, kw:kw + W]
                        y[:, pd::2, ph::2] += np.einsum("ndhwi,oi->ndhwo", xs_,
                                                        wcu[up_tap(pd, ph, ia, ib, kw)])
    if b is not None:
        y += b
    if relu:
        y = np.maximum(y, 0.0)
    return y


def conv3_input_grad(dy, w):
    """d(sum dy * conv3(x, w)) / dx"""
    n, D, H, W, co = dy.shape
    ci = w.shape[2]
    dyp = np.zeros((n, D + 2, H + 2, W + 2, co))
    dyp[:, 1:-1, 1:-1, 1:-1] = dy
    dx = np.zeros((n, D, H, W, ci))
    for kd in range(3):
        for kh in range(3):
            for kw in range(3):
                tap = (kd * 3 + kh) * 3 + kw
                # y[v] += x[v + t - 1] w[t]  =>  dx[u] += dy[u - t + 1] w[t]
                dx += np.einsum("ndhwo,oi->ndhwi",
                                dyp[:, 2 - kd:2 - kd + D, 2 - kh:2 - kh + H, 2 - kw:2 - kw + W],
                                w[tap])
    return dx


def conv3_weight_grad(x, dy):
    n, D, H, W, ci = x.shape
    co = dy.shape[-1]
    xp = np.zeros((n, D + 2, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dw = np.zeros((27, co, ci))
    for kd in range(3):
        for kh in range(3):
            for kw in range(3):
                tap = (kd * 3 + kh) * 3 + kw
                dw[tap] = np.einsum("ndhwo,ndhwi->oi", dy, xp[:, kd:kd + D, kh:kh + H, kw:kw + W])
    return dw


def maxpool(x):
    n, D, H, W, c = x.shape
    v = x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c)
    return v.max(axis=(2, 4, 6))


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


# The warm-start network: name, input channels, output channels. Encoder
# e1a e1b | pool | e2a e2b | pool | b1 b2 | up, concat e2b | d2a d2b | up,
# concat e1b | d1a d1b | head (linear). ReLU after every layer but the head.
LAYERS = (("e1a", 16, 16), ("e1b", 16, 16), ("e2a", 16, 32), ("e2b", 32, 32),
          ("b1", 32, 64), ("b2", 64, 64), ("d2a", 96, 32), ("d2b", 32, 32),
          ("d1a", 48, 16), ("d1b", 16, 16), ("head", 16, 16))


def unet_forward(x, P):
    """x [n][d][h][w][16]; P[name] = (w [27][co][ci], b [co]). Returns the
    output and the saved activations."""
    A = {"x": x}

    def cv(name, inp, relu=True):
        w, b = P[name]
        A[name] = conv3(inp, w, b, relu)
        return A[name]

    s1 = cv("e1b", cv("e1a", x))
    p1 = A["p1"] = maxpool(s1)
    s2 = cv("e2b", cv("e2a", p1))
    p2 = A["p2"] = maxpool(s2)
    bt = cv("b2", cv("b1", p2))
    c2 = A["c2"] = np.concatenate([upsample(bt), s2], axis=-1)
    u2 = cv("d2b", cv("d2a", c2))
    c1 = A["c1"] = np.concatenate([upsample(u2), s1], axis=-1)
    u1 = cv("d1b", cv("d1a", c1))
    return cv("head", u1, relu=False), A


def _pool_bwd(x, dy):
    """gradient of maxpool at x (first maximum in (d, h, w) order wins)"""
    n, D, H, W, c = x.shape
    v = x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).transpose(0, 1, 3, 5, 7, 2, 4, 6)
    v = v.reshape(n, D // 2, H // 2, W // 2, c, 8)
    k = np.argmax(v, axis=-1)
    g = np.zeros(v.shape)
    np.put_along_axis(g, k[..., None], dy[..., None], axis=-1)
    g = g.reshape(n, D // 2, H // 2, W // 2, c, 2, 2, 2).transpose(0, 1, 5, 2, 6, 3, 7, 4)
    return g.reshape(x.shape)


def _up_bwd(dy):
    n, D, H, W, c = dy.shape
    return dy.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, c).sum(axis=(2, 4, 6))


def unet_backward(dout, P, A):
    """Gradients {name: (dw, db)} of sum(dout * output)."""
    G = {}

    def layer(name, inp, dy, gate):
        """dy: gradient at the layer's pre-activation; returns the gradient
        at its input times (gate > 0) when gate is given"""
        w, _ = P[name]
        G[name] = (conv3_weight_grad(A[inp] if isinstance(inp, str) else inp, dy),
                   dy.sum(axis=(0, 1, 2, 3)))
        if gate is None:
            return None
        return conv3_input_grad(dy, w) * (gate > 0)

    du1 = layer("head", A["d1b"], dout, A["d1b"])
    dd1a = layer("d1b", A["d1a"], du1, A["d1a"])
    dc1 = layer("d1a", A["c1"], dd1a, A["c1"])
    c_u2 = A["d2b"].shape[-1]
    du2 = _up_bwd(dc1[..., :c_u2])
    dd2a = layer("d2b", A["d2a"], du2, A["d2a"])
    dc2 = layer("d2a", A["c2"], dd2a, A["c2"])
    c_b = A["b2"].shape[-1]
    db2 = _up_bwd(dc2[..., :c_b])
    db1 = layer("b2", A["b1"], db2, A["b1"])
    dp2 = layer("b1", A["p2"], db1, A["p2"])
    ds2 = _pool_bwd(A["e2b"], dp2) + dc2[..., c_b:]
    de2a = layer("e2b", A["e2a"], ds2, A["e2a"])
    dp1 = layer("e2a", A["p1"], de2a, A["p1"])
    ds1 = _pool_bwd(A["e1b"], dp1) + dc1[..., c_u2:]
    de1a = layer("e1b", A["e1a"], ds1, A["e1a"])
    layer("e1a", A["x"], de1a, None)
    return G


def unet_loss(out, t):
    """mean over voxels of (out[..., 0] - t)^2, and its gradient"""
    e = out[..., 0] - t
    nvox = e.size
    dout = np.zeros(out.shape)
    dout[..., 0] = 2.0 * e / nvox
    return float(np.sum(e * e)) / nvox, dout
