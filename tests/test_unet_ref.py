"""Pins the FP64 numpy reference of the warm-start network (the checker of
the agdu_* kernels) against independent routines: scipy's N-d correlation
for the convolution, the adjoint identities <A x, y> = <x, A^T y> for the
input and weight gradients, pooling and upsampling, and finite differences
of the loss."""
import numpy as np
import pytest
from scipy import signal

from tests import unet_ref as R


def test_conv3_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 4, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    y = R.conv3(x, w, b)
    k = w.reshape(3, 3, 3, 3, 2)  # [kd][kh][kw][co][ci]
    for co in range(3):
        ref = b[co] + sum(signal.correlate(x[0, ..., ci], k[..., co, ci], mode="same",
                                           method="direct") for ci in range(2))
        assert np.allclose(y[0, ..., co], ref, rtol=1e-12, atol=1e-12)
    assert np.array_equal(R.conv3(x, w, b, relu=True), np.maximum(y, 0))


def test_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 4, 5, 3))
    w = rng.standard_normal((27, 4, 3))
    dy = rng.standard_normal((2, 3, 4, 5, 4))
    lhs = np.sum(R.conv3(x, w) * dy)
    assert np.isclose(lhs, np.sum(x * R.conv3_input_grad(dy, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * R.conv3_weight_grad(x, dy)), rtol=1e-12)


def test_pool_upsample_and_loss():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 2, 6, 2))
    p = R.maxpool(x)
    assert p.shape == (1, 2, 1, 3, 2)
    assert p[0, 1, 0, 2, 1] == x[0, 2:4, 0:2, 4:6, 1].max()
    u = rng.standard_normal((1, 8, 4, 12, 2))
    lhs = np.sum(R.upsample(x) * u)
    pooled_sum = u.reshape(1, 4, 2, 2, 2, 6, 2, 2).sum(axis=(2, 4, 6))
    assert np.isclose(lhs, np.sum(x * pooled_sum), rtol=1e-12)


@pytest.mark.parametrize("v", [1.0, -3.5, 0.1, 65504.0, 1.00390625, 1.01171875])
def test_bf16_round_to_nearest_even(v):
    got = R.bf16_round(np.array([v]))[0]
    # brute force: the nearest of all bf16 values with that exponent, ties even
    u = np.array([v], np.float32).view(np.uint32)[0]
    lo = np.array([u & 0xFFFF0000], np.uint32).view(np.float32)[0]
    hi = np.array([(u & 0xFFFF0000) + 0x10000], np.uint32).view(np.float32)[0]
    dl, dh = abs(float(v) - float(lo)), abs(float(hi) - float(v))
    want = lo if dl < dh else hi if dh < dl else (lo if ((u >> 16) & 1) == 0 else hi)
    assert got == float(want)


def _small_net(rng, scale=0.3):
    P = {}
    for name, ci, co in R.LAYERS:
        P[name] = (rng.standard_normal((27, co, ci)) * scale / np.sqrt(27 * ci),
                   rng.standard_normal(co) * 0.05)
    return P


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(4)
    P = _small_net(rng)
    x = np.zeros((1, 4, 4, 8, 16))
    x[..., 0] = rng.standard_normal((1, 4, 4, 8))
    t = rng.standard_normal((1, 4, 4, 8))
    out, A = R.unet_forward(x, P)
    loss, dout = R.unet_loss(out, t)
    G = R.unet_backward(dout, P, A)
    for name, _, _ in R.LAYERS:
        w, b = P[name]
        for which, arr in ((0, w), (1, b)):
            flat = arr.reshape(-1)
            for k in rng.choice(flat.size, 3, replace=False):
                old = flat[k]
                h = 1e-6
                flat[k] = old + h
                lp = R.unet_loss(R.unet_forward(x, P)[0], t)[0]
                flat[k] = old - h
                lm = R.unet_loss(R.unet_forward(x, P)[0], t)[0]
                flat[k] = old
                fd = (lp - lm) / (2 * h)
                an = G[name][which].reshape(-1)[k]
                assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3), (name, which, fd, an)


def test_combined_taps_equal_the_upsampled_convolution():
    """A 3x3x3 convolution of [upsample(xu), xs] equals the xs part plus the
    48 combined taps over xu (exact FP64 sums: an identity), for every
    output parity and at the borders; rounded combined taps are the float32
    sums of the bf16 weights, rounded once."""
    rng = np.random.default_rng(17)
    n, D, H, W, cu, cs, co = 1, 6, 4, 10, 3, 2, 4
    xu = rng.standard_normal((n, D // 2, H // 2, W // 2, cu))
    xs = rng.standard_normal((n, D, H, W, cs))
    w = rng.standard_normal((27, co, cu + cs))
    b = rng.standard_normal(co)
    cat = np.concatenate([R.upsample(xu), xs], axis=-1)
    ref = R.conv3(cat, w, b, relu=True)
    got = R.conv3_up(xu, xs, w, R.up_weights(w, cu, rounded=False), b, relu=True)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # a plausible slip (the parity sets swapped) breaks it
    wrong = R.up_weights(w[::-1].copy(), cu, rounded=False)
    assert np.abs(R.conv3_up(xu, xs, w, wrong, b) - R.conv3(cat, w, b)).max() > 1e-3
    wb = R.bf16_round(w)
    wr = R.up_weights(wb, cu)
    assert np.array_equal(wr, R.bf16_round(wr))
    assert np.abs(wr - R.up_weights(wb, cu, rounded=False)).max() <= 2 ** -8 * np.abs(wr).max()



def test_convolution_shape_limits_are_refused_on_the_host():
    """The kernels index a d plane in 32-bit unsigned arithmetic (H * W *
    channels < 2^32) and launch one grid row per (n, d block): larger shapes
    are refused before any device work (no GPU needed for the refusal)."""
    from agdeblend_amd import lib as L
    lib = L.load()
    # H * W * 16 = 2^36: over the plane limit
    assert lib.agdu_conv3(None, None, None, None, None, 1, 1, 65535, 65536, 16, 16, 3, None) == 1
    # H over 65535 grid rows
    assert lib.agdu_conv3(None, None, None, None, None, 1, 1, 65536, 64, 16, 16, 3, None) == 1
    # N * D over 65535
    assert lib.agdu_conv3(None, None, None, None, None, 2, 40000, 8, 64, 16, 16, 3, None) == 1
    assert lib.agdu_conv_in(None, None, 16, None, None, 1, 1, 65535, 65536, 1.0, None) == 1
    assert lib.agdu_conv_out(None, None, 16, None, None, None, 1, 1, 65535, 65536, 1.0,
                             None) == 1
