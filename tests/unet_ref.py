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
