"""Each operator of the forward model A = mute . blend . wavelet . F passes
the adjoint (dot) test on the GPU on its own, accumulated in FP64:

    |<Op x, y> - <x, Op^H y>| <= 1e-6 * ||Op x|| * ||y||

  * F, the windowed FK operator (fk.c:307-452): model -> traces, on a
    Hermitian-consistent model (the reference's FK.DotTest,
    fk_test.cpp:547-660, uses the same restriction; the adjoint doubles the
    interior time frequencies, fk.c:436-451);
  * the wavelet convolution (wavelet.c:239-301): traces -> traces, on input
    traces that are zero beyond n_times_fk (what F delivers);
  * the blend with the mute (blend.c:862-875, 936-960, 1262-1292): traces
    -> record, adjoint taken on the muted record.

The composite A is dot-tested in test_gpu_parity. Inputs come from the same
seeded geometries (specialised and generic kernel sizes, wavelets, bad and
missing traces, several patches).
"""
import zlib

import numpy as np
import pytest

from agdeblend_amd import lib as L
from tests.test_gpu_parity import GEOS, _kw

pytestmark = pytest.mark.gpu

BAR = 1e-6


def _hermitian_model(pm, rng):
    """R2C (time) and DFT (space) of random real windows, in the product's
    [ntw][Ft][S] layout."""
    x = np.zeros(pm.model_len, np.complex128)
    for w in range(pm.n_windows):
        a = pm.window(w)
        shape = [a["ntw"]] + a["P"] + [a["Pt"]]
        z = rng.random(shape) - 0.5
        xf = np.fft.rfftn(z, axes=list(range(1, len(shape))))  # [ntw][P...][Ft]
        xf = xf.reshape(a["ntw"], a["S"], a["Ft"]).transpose(0, 2, 1).reshape(-1)
        x[a["x_off"]:a["x_off"] + xf.size] = xf
    return x


def _zero_beyond_fk(pm, t):
    t = t.copy()
    for w in range(pm.n_windows):
        a = pm.window(w)
        off = _trace_off(pm, w)
        blk = t[off:off + a["ntr"] * a["nconv"]].reshape(a["ntr"], a["nconv"])
        blk[:, a["nfk"]:] = 0.0
    return t


def _trace_off(pm, w):
    off = 0
    for v in range(w):
        a = pm.window(v)
        off += a["ntr"] * a["nconv"]
    return off


def _mute_mask(pm):
    iv = pm.intervals()
    offs = np.concatenate([[0], np.cumsum(iv[:, 2] - iv[:, 1])]).astype(np.int64)
    first = {}
    for k, (c, _, _) in enumerate(iv):
        first.setdefault(int(c), k)
    mask = np.ones(pm.record_len)
    for c, i, s, n in pm.mutes():
        o = offs[first[int(c)] + int(i)] + int(s)
        mask[o:o + int(n)] = 0.0
    return mask


def _dot_ok(name, lhs, rhs, norm):
    print("%s: <Op x, y> %.9e  <x, Op^H y> %.9e  rel %.2e" %
          (name, lhs, rhs, abs(lhs - rhs) / max(norm, 1e-300)))
    assert abs(lhs - rhs) <= BAR * norm, (lhs, rhs, norm)


@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_fk_operator_is_adjoint(name, geo):
    pm = L.Problem(**_kw(geo))
    try:
        if pm.n_windows == 0:
            pytest.skip("no windows")
        rng = np.random.default_rng(zlib.crc32(("fk" + name).encode()))
        x = _hermitian_model(pm, rng)
        y = rng.random(pm.trace_len) - 0.5
        fx = pm.operator(0, 0, x.astype(np.complex64)).astype(np.float64)
        g = pm.operator(0, 1, y.astype(np.float32)).astype(np.complex128)
        # F x is zero beyond each trace's n_times_fk
        assert np.array_equal(fx, _zero_beyond_fk(pm, fx))
        _dot_ok("FK " + name, float(np.dot(fx, y)),
                float(np.sum(x.real * g.real + x.imag * g.imag)),
                float(np.linalg.norm(fx) * np.linalg.norm(y)))
    finally:
        pm.close()


WAVELET_GEOS = [g for g in GEOS if "wavelet_lengths" in g[1]]


@pytest.mark.parametrize("name,geo", WAVELET_GEOS, ids=[g[0] for g in WAVELET_GEOS])
def test_wavelet_operator_is_adjoint(name, geo):
    pm = L.Problem(**_kw(geo))
    try:
        if pm.n_windows == 0:
            pytest.skip("no windows")
        rng = np.random.default_rng(zlib.crc32(("wv" + name).encode()))
        t = _zero_beyond_fk(pm, rng.random(pm.trace_len) - 0.5)
        s = rng.random(pm.trace_len) - 0.5
        wt = pm.operator(1, 0, t.astype(np.float32)).astype(np.float64)
        ws = pm.operator(1, 1, s.astype(np.float32)).astype(np.float64)
        _dot_ok("wavelet " + name, float(np.dot(wt, s)), float(np.dot(t, ws)),
                float(np.linalg.norm(wt) * np.linalg.norm(s)))
    finally:
        pm.close()


@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_blend_operator_is_adjoint(name, geo):
    pm = L.Problem(**_kw(geo))
    try:
        if pm.n_windows == 0 or pm.record_len == 0:
            pytest.skip("empty geometry")
        rng = np.random.default_rng(zlib.crc32(("bl" + name).encode()))
        t = rng.random(pm.trace_len) - 0.5
        r = rng.random(pm.record_len) - 0.5
        bt = pm.operator(2, 0, t.astype(np.float32)).astype(np.float64)
        br = pm.operator(2, 1, (r * _mute_mask(pm)).astype(np.float32)).astype(np.float64)
        _dot_ok("blend " + name, float(np.dot(bt, r)), float(np.dot(t, br)),
                float(np.linalg.norm(bt) * np.linalg.norm(r)))
    finally:
        pm.close()
