"""Every compile-time wavelet kernel against the FP64 oracle.

The wavelet convolution runs with a real FFT length Mr >= nconv + nfk - 1
(any such length gives the same linear convolution as the reference's
length-nconv circular one, wavelet.c:239-301). The host prefers a length
with a compile-time kernel (half lengths 512 ... 6144) when it is at most
25% longer than the smallest even 7-smooth length. Here a small vibroseis
geometry (2 x 16 sources, windows (2, 16, 64)) gets a wavelet length that
lands on each of those lengths in turn.

CPU: the host picks the intended length (and falls back to the 7-smooth
length when no compile-time length is close). GPU: forward and adjoint
within 2e-6 of the FP64 oracle, the same within 2e-6 of the generic kernel
(AGD_GENERIC_WAVELET), and the FP64-accumulated dot test within 1e-6.
"""
import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from tests.test_gpu_parity import (_kw, _mute_vec, _random_model, _to_oracle_layout,
                                   _to_product_layout)

HALF = [512, 768, 1024, 1536, 2048, 3072, 4096, 6144]
NFK = 224  # n_times_fk of these geometries (n_t = n_w + 150, time windows 64)


def _geometry(n_w, seed):
    p = synth.config3(seed, n_y=2, n_x=16, n_t=n_w + 150, n_w=n_w, windows=(2, 16, 64),
                      n_its=5)
    return _kw(p["geometry"])


def _lengths(kw):
    pm = L.Problem(host_only=True, **kw)
    try:
        return {(pm.window(w)["nfk"], pm.window(w)["nconv"], pm.window(w)["wM"])
                for w in range(pm.n_windows)}
    finally:
        pm.close()


def _case(half):
    """n_w such that nconv + nfk - 1 = 2 half - 12"""
    n_w = 2 * half - 2 * NFK + 2 - 12
    return n_w, _geometry(n_w, 1100 + half)


@pytest.mark.parametrize("half", HALF)
def test_host_picks_compile_time_length(half):
    n_w, kw = _case(half)
    (nfk, nconv, wm), = _lengths(kw)
    assert nfk == NFK and nconv == nfk + n_w - 1
    assert wm == 2 * half


def test_host_keeps_smooth_length_when_no_kernel_is_close():
    # nconv + nfk - 1 = 6146: the next compile-time length 8192 is more than
    # 25% above the 7-smooth 6174 = 2 * 3^2 * 7^3
    n_w = 6146 - 2 * NFK + 2
    (nfk, nconv, wm), = _lengths(_geometry(n_w, 7))
    assert nconv + nfk - 1 == 6146
    assert wm == 6174


@pytest.mark.gpu
@pytest.mark.parametrize("half", HALF)
def test_compile_time_wavelet_kernel_matches_oracle_and_generic(half, monkeypatch):
    from oracle import oracle as O
    _, kw = _case(half)
    om = O.Model(double=True, **kw)
    rng = np.random.default_rng(half)
    x = _random_model(om, rng)
    y = _mute_vec(om, rng.random(om.record_len) - 0.5)
    ref_f = om.forward(x)
    ref_a = om.adjoint_update(y, -1.0, om.zeros_model())

    def ops():
        pm = L.Problem(**kw)
        try:
            f = pm.forward(_to_product_layout(om, x).astype(np.complex64)).astype(np.float64)
            a = pm.adjoint(y.astype(np.float32)).astype(np.complex128)
            return f, _to_oracle_layout(om, a)
        finally:
            pm.close()

    f1, a1 = ops()
    monkeypatch.setenv("AGD_GENERIC_WAVELET", "1")
    f2, a2 = ops()
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    print("half length %d: forward %.2e adjoint %.2e | generic: forward %.2e adjoint %.2e" %
          (half, rel(f1, ref_f), rel(a1, ref_a), rel(f2, ref_f), rel(a2, ref_a)))
    assert rel(f1, ref_f) <= 2e-6 and rel(a1, ref_a) <= 2e-6
    assert rel(f1, f2) <= 2e-6 and rel(a1, a2) <= 2e-6
    lhs = float(np.dot(f1, y))
    rhs = float(np.sum(x.real * a1.real + x.imag * a1.imag))
    assert abs(lhs - rhs) <= 1e-6 * float(np.linalg.norm(f1) * np.linalg.norm(y))
