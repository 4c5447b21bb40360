"""CPU tests of the product library's host side.

The geometry and blend tables libagdeblend.so builds (window layout, FK
sizes, channel numbering, record intervals, mutes, per-trace blend
coordinates) must be bit-identical to the oracle's restatement of the
reference, which the known-answer tests pin. These run without a GPU
(agdx_create_host builds the tables only).
"""
import os
import re

import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _geometries():
    cases = []
    for i in (0, 1, 3):
        g = synth.SMALL[i]()["geometry"]
        cases.append(("small%d" % i, g))
    g = synth.config0()["geometry"]
    cases.append(("c0", g))
    # the model known-answer geometries
    cases.append(("kat_forward", dict(
        volumes=[0, 1, 2, 0], n_dims=[2, 3, 2, 2],
        window_shapes=[[2, 4], [2, 2, 4], [4, 4]], coords=[[0], [0, 0], [0], [1]],
        shapes=[[4, 4], [2, 3, 5], [4, 4], [4, 4]],
        shottimes=[[1, 15, 51, 21], [7, 27, 35, 43, 100, 57], [1, 2, 3, 4], [21, 65, 71, 77]],
        channels=[[1] * 4, [1, 1, 1, 1, 2, 1], [1] * 4, [1] * 4],
        trace_types=[[0] * 4, [0] * 6, [2] * 4, [0] * 4],
        wavelet_lengths=[1], wavelet_idxs=[[0] * 4, [0] * 6, [0] * 4, [0] * 4],
        wavelets=[np.array([1.0], np.float32)])))
    rng = np.random.default_rng(77)
    from tests.test_oracle_kats import _random_deblend_problem
    for k in range(12):
        prob = _random_deblend_problem(rng, wavelets=bool(k % 2))
        cases.append(("random%d" % k, prob))
    return cases


def _kw(g):
    keys = ["volumes", "n_dims", "window_shapes", "coords", "shapes", "shottimes",
            "channels", "trace_types"]
    kw = {k: g[k] for k in keys}
    for k in ("wavelet_lengths", "wavelet_idxs", "wavelets"):
        if k in g:
            kw[k] = g[k]
    return kw


@pytest.mark.parametrize("name,geo", _geometries(), ids=[c[0] for c in _geometries()])
def test_tables_match_oracle(name, geo):
    kw = _kw(geo)
    om = O.Model(double=True, **kw)
    pm = L.Problem(host_only=True, **kw)
    assert pm.n_windows == om.n_windows
    assert pm.record_len == om.record_len
    assert pm.model_len == om.model_len
    assert pm.intervals().tolist() == om.intervals().tolist()
    assert pm.mutes().tolist() == om.mutes().tolist()
    assert pm.unique_channels().tolist() == om.unique_channels().tolist()
    for w in range(om.n_windows):
        a = pm.window(w)
        b = om.windows[w]
        nsp = b["ndim"] - 1
        assert a["patch"] == b["patch"]
        assert a["ntr"] == b["ntr"] and a["ntw"] == b["ntw"]
        assert a["nfk"] == b["nfk"] and a["nconv"] == b["nconv"]
        assert a["P"] == b["P"][:nsp] and a["Pt"] == b["P"][nsp]
        assert a["W"] == b["W"][:nsp] and a["Wt"] == b["W"][nsp]
        assert a["Ft"] == b["Ft"]
        assert a["coords"] == b["coords"]
        assert a["x_off"] == b["model_off"]
        np.testing.assert_array_equal(pm.blend_coords(w), om.blend_coords(w))
    pm.close()


def _header_functions():
    names = []
    for h in ("agdeblend.h", "agdeblend_gpu.h", "agdeblend_unet.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"\b(agd\w+|agdeblend_\w+|agdx_\w+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    lib = L.load()
    names = _header_functions()
    assert "agd_deblend" in names and "agd_blend" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(L.EXPORTS)


def test_product_tables_do_not_touch_oracle():
    # the product package must not import the oracle
    for root, _, files in os.walk(os.path.join(ROOT, "agdeblend_amd")):
        for f in files:
            if f.endswith((".py", ".c", ".h", ".hip")):
                txt = open(os.path.join(root, f)).read()
                assert not re.search(r"(from|import)\s+oracle|liboracle", txt), f


@pytest.mark.parametrize("wt", [64, 128, 256, 512, 200])
def test_time_factorization_tables(wt):
    """Each specialised time class starts on the default factorization of
    its half length Pt / 2; agdx_set_time_factors accepts exactly the
    candidates with that product, and a class of the generic kernels
    (Wt = 200: Pt = 250) has none."""
    from tests.test_autotune import TIME_CAND
    from tests.test_gpu_parity import _spec_geometry
    kw = _kw(_spec_geometry((64, 2048), (32, wt), 50 + wt))
    pm = L.Problem(host_only=True, **kw)
    try:
        fac = pm.time_factors()
        assert len(fac) == 1
        M = 5 * wt // 8
        if M not in TIME_CAND:
            assert fac == [(0, 0)]
            with pytest.raises(ValueError):
                pm.set_time_factors(0, 5, 25)
            return
        assert fac == [TIME_CAND[M][0]]
        for m1, m2 in TIME_CAND[M]:
            pm.set_time_factors(0, m1, m2)
            assert pm.time_factors() == [(m1, m2)]
        for m1, m2 in ((M, 1), (2, M // 2), (7, 7)):
            with pytest.raises(ValueError):
                pm.set_time_factors(0, m1, m2)
        with pytest.raises(ValueError):
            pm.set_time_factors(1, *TIME_CAND[M][0])
    finally:
        pm.close()
