"""GPU parity of the HIP path against the oracle (run on an MI355X).

Bars (written in each test):
  * forward / adjoint operators vs the FP64 oracle: relative error <= 2e-6
  * operator adjointness in FP64 accumulation (the standard dot test):
    |<Ax,y> - <x,A^H y>| <= 1e-6 * ||Ax|| * ||y||
  * agd_deblend after a fixed number of iterations vs the FP64 oracle:
    relative L2 <= 1e-4
  * agd_blend sum / overwrite / untapered mean: bit-exact vs the float
    oracle; tapered mean: <= 2 ulp relative
  * integer tables: bit-exact (also covered on the CPU by test_host_tables)
"""
import zlib

import numpy as np
import pytest

from agdeblend_amd import agdeblend as agd
from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIVE, BAD, MISSING = 0, 1, 2


def _kw(g):
    keys = ["volumes", "n_dims", "window_shapes", "coords", "shapes", "shottimes",
            "channels", "trace_types"]
    kw = {k: g[k] for k in keys}
    for k in ("wavelet_lengths", "wavelet_idxs", "wavelets"):
        if k in g:
            kw[k] = g[k]
    return kw


def _to_product_layout(om, x):
    """oracle [ntw][P_space...][Ft] per window -> product [ntw][Ft][S]"""
    out = np.empty_like(x)
    for wi in om.windows:
        S = wi["npad_tr"]
        n = wi["ntw"] * S * wi["Ft"]
        o = wi["model_off"]
        out[o:o + n] = x[o:o + n].reshape(wi["ntw"], S, wi["Ft"]).transpose(0, 2, 1).reshape(-1)
    return out


def _to_oracle_layout(om, x):
    out = np.empty_like(x)
    for wi in om.windows:
        S = wi["npad_tr"]
        n = wi["ntw"] * S * wi["Ft"]
        o = wi["model_off"]
        out[o:o + n] = x[o:o + n].reshape(wi["ntw"], wi["Ft"], S).transpose(0, 2, 1).reshape(-1)
    return out


def _random_model(om, rng):
    """Hermitian-consistent random model (R2C of random reals per window)."""
    x = om.zeros_model().astype(np.complex128)
    for wi in om.windows:
        real = rng.random([wi["ntw"]] + wi["P"]) - 0.5
        xw = np.fft.rfftn(real, axes=list(range(1, wi["ndim"] + 1))).reshape(-1)
        x[wi["model_off"]:wi["model_off"] + len(xw)] = xw
    return x


def _spec_geometry(shape, windows, seed, interval=None):
    """A single-patch geometry whose windows hit one specialised kernel class."""
    rng = np.random.default_rng(seed)
    nd = len(shape)
    lead = shape[:-1]
    ntr = int(np.prod(lead))
    nsrc = lead[-1]
    step = interval if interval is not None else shape[-1] // 2
    src_t = np.arange(nsrc) * step + rng.integers(-step // 4, step // 4 + 1, nsrc)
    src_t -= src_t.min()
    shot = np.broadcast_to(src_t, lead).reshape(-1).astype(np.int64)
    if nd == 3:
        ch = np.broadcast_to(np.arange(lead[0])[:, None], lead).reshape(-1)
    else:
        ch = np.zeros(ntr, np.int64)
    tt = rng.choice([0, 0, 0, 0, 0, 0, 0, 0, 1, 2], ntr)
    return dict(volumes=[0], n_dims=[nd], window_shapes=[list(windows)],
                coords=[[0] * (nd - 1)], shapes=[list(shape)], shottimes=[shot],
                channels=[ch.astype(np.intc)], trace_types=[tt.astype(np.intc)])


SPEC = [("spec_20x40_320", _spec_geometry((16, 64, 512), (16, 32, 256), 11)),
        ("spec_40x40_640", _spec_geometry((32, 64, 1024), (32, 32, 512), 12)),
        ("spec_10x40_320", _spec_geometry((8, 64, 512), (8, 32, 256), 13)),
        ("spec_1x40_320", _spec_geometry((128, 1024), (32, 256), 14)),
        # specialised spatial size with a generic time size: untiled U / V
        ("spec_20x40_t250", _spec_geometry((16, 64, 500), (16, 32, 200), 15)),
        # specialised spatial size whose windows hold 8 traces: untiled
        ("spec_1x10_320", _spec_geometry((64, 512), (8, 256), 16))]


def _geometries():
    out = list(SPEC)
    out += [("c0_small", synth.SMALL[0]()["geometry"]),
           ("c1_small", synth.SMALL[1]()["geometry"]),
           ("c3_small", synth.SMALL[3]()["geometry"])]
    from tests.test_oracle_kats import _random_deblend_problem
    rng = np.random.default_rng(5)
    for k in range(6):
        out.append(("random%d" % k, _random_deblend_problem(rng, wavelets=bool(k % 2))))
    return out


GEOS = _geometries()


def _mute_vec(om, v):
    iv = om.intervals()
    offs = np.concatenate([[0], np.cumsum(iv[:, 2] - iv[:, 1])])
    first = {}
    for k, (c, a, b) in enumerate(iv):
        first.setdefault(int(c), k)
    v = v.copy()
    for c, i, s, n in om.mutes():
        o = offs[first[int(c)] + int(i)] + s
        v[o:o + n] = 0
    return v


@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_forward_adjoint_match_oracle_and_are_adjoint(name, geo):
    kw = _kw(geo)
    om = O.Model(double=True, **kw)
    if om.n_windows == 0 or om.record_len == 0:
        pytest.skip("empty geometry")
    pm = L.Problem(**kw)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    x = _random_model(om, rng)
    # forward
    ref = om.forward(x)
    got = pm.forward(_to_product_layout(om, x).astype(np.complex64)).astype(np.float64)
    scale = max(np.abs(ref).max(), 1e-30)
    assert np.abs(got - ref).max() <= 2e-6 * scale * max(1.0, np.sqrt(len(ref)) / 64)
    rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
    assert rel <= 2e-6, rel
    # adjoint
    y = _mute_vec(om, rng.random(om.record_len) - 0.5)
    gref = om.adjoint_update(y, -1.0, om.zeros_model())
    gp = _to_oracle_layout(om, pm.adjoint(y.astype(np.float32)).astype(np.complex128))
    rel_g = np.linalg.norm(gp - gref) / max(np.linalg.norm(gref), 1e-30)
    assert rel_g <= 2e-6, rel_g
    # dot test of the GPU operators, accumulated in FP64
    lhs = float(np.dot(got, y))
    rhs = float(np.sum(x.real * gp.real + x.imag * gp.imag))
    norm = float(np.linalg.norm(got) * np.linalg.norm(y))
    assert abs(lhs - rhs) <= 1e-6 * norm, (lhs, rhs, norm)
    pm.close()


def test_stepsize_matches_oracle():
    for name, geo in GEOS:
        kw = _kw(geo)
        om = O.Model(double=False, **kw)
        if om.n_windows == 0 or om.record_len == 0:
            continue
        pm = L.Problem(**kw)
        assert pm.stepsize() == om.stepsize(), name
        pm.close()


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_deblend_matches_fp64_oracle(cfg):
    prob = synth.SMALL[cfg]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    n_its = prob["n_its"]
    ref, stats = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
    out = agd.deblend(g["volumes"], g["window_shapes"], g["coords"], g["shottimes"],
                      g["channels"], g["trace_types"], 1.0, n_its, 0, blended,
                      wavelet_idxs=g.get("wavelet_idxs"), wavelets=g.get("wavelets"))
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    rel = (num / den) ** 0.5
    assert rel <= 1e-4, rel


@pytest.mark.parametrize("name,geo", SPEC, ids=[g[0] for g in SPEC])
def test_deblend_specialised_classes_match_fp64_oracle(name, geo):
    """Fused adjoint + forward kernels over several iterations."""
    kw = _kw(geo)
    rng = np.random.default_rng(7)
    truth = [rng.standard_normal(tuple(geo["shapes"][0])).astype(np.float32)]
    prob = dict(geometry=geo, truth=truth)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    n_its = 6
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
    out = agd.deblend(geo["volumes"], geo["window_shapes"], geo["coords"], geo["shottimes"],
                      geo["channels"], geo["trace_types"], 1.0, n_its, 0, blended)
    rel = np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0])
    assert rel <= 1e-4, rel


def test_deblend_with_initial_factor_below_one():
    # Q1: with f < 1 the start is the sign-flipped shrunk update
    prob = synth.SMALL[0]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    ref, _ = O.deblend(data=blended, initial_factor=0.5, n_its=8, double=True, **kw)
    out = agd.deblend(g["volumes"], g["window_shapes"], g["coords"], g["shottimes"],
                      g["channels"], g["trace_types"], 0.5, 8, 0, blended)
    rel = np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0])
    assert rel <= 1e-4, rel


def test_agd_blend_kats():
    # test.cpp:30-430
    out = agd.blend([[25], [27]], [[123], [123]], [[LIVE], [LIVE]],
                    [np.array([[1., 2, 3]]), np.array([[4., 5, 6]])], agd.AGDBlendSum)
    assert out[0].tolist() == [[1, 2, 7]] and out[1].tolist() == [[7, 5, 6]]
    for tin, tout in [(MISSING, MISSING), (MISSING, LIVE), (LIVE, MISSING)]:
        out = agd.blend([[25]], [[123]], [[tin]], [np.array([[1., 2, 3]])], agd.AGDBlendSum,
                        trace_types_out=[[tout]])
        assert out[0].tolist() == [[0, 0, 0]]
    out = agd.blend([[25, 27, 29, 31, -10]], [[123, 123, 123, 123, -789]],
                    [[MISSING, LIVE, LIVE, BAD, LIVE]],
                    [np.arange(1, 16, dtype=np.float32).reshape(5, 3)], agd.AGDBlendSum,
                    n_times_out=[2, 3], shottimes_out=[[24, 26, -10], [28, 30]],
                    channels_out=[[123, 123, -789], [123, 123]],
                    trace_types_out=[[MISSING, LIVE, LIVE], [LIVE, LIVE]])
    assert out[0].tolist() == [[0, 0], [0, 4], [13, 14]]
    assert out[1].tolist() == [[5, 13, 8], [8, 0, 0]]


@pytest.mark.parametrize("mode,taper", [(0, 0), (2, 0), (1, 0), (1, 5)])
def test_agd_blend_random_matches_oracle(mode, taper):
    rng = np.random.default_rng(100 + mode * 10 + taper)
    for _ in range(4):
        np_ = int(rng.integers(1, 4))
        ntr = [int(rng.integers(1, 40)) for _ in range(np_)]
        nt = [int(rng.integers(1, 60)) for _ in range(np_)]
        smax = max(1, sum(a * b for a, b in zip(ntr, nt)) // 3)
        shot = [rng.integers(0, smax, n) for n in ntr]
        ch = [rng.integers(0, max(1, sum(ntr) // 10), n) for n in ntr]
        tt = [rng.integers(0, 3, n) for n in ntr]
        data = [rng.standard_normal((a, b)).astype(np.float32) for a, b in zip(ntr, nt)]
        tto = [np.where(t == BAD, LIVE, t) for t in tt]
        ref = O.blend(ntr, nt, shot, ch, tt, data, mode, taper, ntr, nt, shot, ch, tto)
        got = agd.blend(shot, ch, tt, data, mode, taper_length=taper, trace_types_out=tto)
        for a, b in zip(got, ref):
            if mode == 1 and taper > 1:
                np.testing.assert_allclose(a, b, rtol=3e-7, atol=1e-7)
            else:
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode,taper", [(0, 0), (2, 0), (1, 0), (1, 5)])
def test_agd_blend_deep_overlap_matches_oracle(mode, taper):
    """600 traces of one channel within a few samples of each other: every
    256-sample record tile has several hundred blend entries, so the kernel
    stages them in several passes; the sums stay those of the serial order."""
    rng = np.random.default_rng(7 + mode + taper)
    ntr, nt = [600], [300]
    shot = [rng.integers(0, 40, ntr[0])]
    ch = [np.zeros(ntr[0], np.int64)]
    tt = [np.zeros(ntr[0], np.int64)]
    data = [rng.standard_normal((ntr[0], nt[0])).astype(np.float32)]
    ref = O.blend(ntr, nt, shot, ch, tt, data, mode, taper, ntr, nt, shot, ch, tt)
    got = agd.blend(shot, ch, tt, data, mode, taper_length=taper, trace_types_out=tt)
    for a, b in zip(got, ref):
        if mode == 1 and taper > 1:
            np.testing.assert_allclose(a, b, rtol=3e-7, atol=1e-7)
        else:
            np.testing.assert_array_equal(a, b)


def test_blend_input_record_matches_oracle():
    prob = synth.SMALL[1]()
    g = prob["geometry"]
    kw = _kw(g)
    om = O.Model(double=False, **kw)
    pm = L.Problem(**kw)
    ref = om.blend_input(prob["truth"], g["shottimes"], g["channels"], g["trace_types"])
    got = pm.set_data(prob["truth"], want_record=True)
    np.testing.assert_array_equal(got, ref)
    pm.close()


# ---------------------------------------------------------------------------
# Warm-start hook: x0 = s R2C(pad(y_w)) per window (no tapers, no x2)
# ---------------------------------------------------------------------------
WARM = [g for g in GEOS if g[0].startswith(("spec_", "c0_", "c1_", "random"))]


@pytest.mark.parametrize("name,geo", WARM, ids=[g[0] for g in WARM])
def test_warm_model_matches_oracle(name, geo):
    kw = _kw(geo)
    om = O.Model(double=True, **kw)
    if om.n_windows == 0 or om.record_len == 0:
        pytest.skip("empty geometry")
    rng = np.random.default_rng(zlib.crc32(("warm" + name).encode()))
    shapes = [tuple(int(v) for v in s) for s in geo["shapes"]]
    y = [rng.standard_normal(s) for s in shapes]
    ref = om.from_data(y, [list(s) for s in shapes])
    pm = L.Problem(**kw)
    pm.warm_model([a.astype(np.float32) for a in y])
    got = _to_oracle_layout(om, pm.model().astype(np.complex128))
    rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
    assert rel <= 2e-6, rel
    pm.close()


@pytest.mark.parametrize("name", ["spec_20x40_320", "spec_40x40_640", "spec_1x40_320"])
def test_warm_model_output_is_the_data(name):
    """Windows of nominal length: the tapers sum to one, so assembling the
    output of the warm-start model of y gives y back."""
    geo = dict(SPEC)[name]
    kw = _kw(geo)
    shapes = [tuple(int(v) for v in s) for s in geo["shapes"]]
    rng = np.random.default_rng(3)
    y = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    pm = L.Problem(**kw)
    pm.set_data(y)
    pm.warm_model(y)
    out = pm.output()
    for a, b in zip(out, y):
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max(), np.abs(a - b).max()
    pm.close()


@pytest.mark.parametrize("cfg", [0, 1])
@pytest.mark.parametrize("n_its", [1, 8])
def test_warm_deblend_matches_fp64_oracle(cfg, n_its):
    """n_its = 1: the warm-start model's output, with no threshold needed
    (the library skips the first update that computes lambda0)."""
    prob = synth.SMALL[cfg]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    # a data estimate between the blended input and the truth
    warm = [0.5 * (np.asarray(a, np.float64) + b) for a, b in zip(blended, prob["truth"])]
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, warm=warm,
                       **kw)
    pm = L.Problem(**kw)
    pm.set_data(blended)
    pm.run_warm(1.0, n_its, [a.astype(np.float32) for a in warm])
    out = pm.output()
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    rel = (num / den) ** 0.5
    assert rel <= 1e-4, rel
    pm.close()


def test_captured_iterations_match_eager(monkeypatch):
    """The iterations replayed from a captured graph give bit-identical
    results to eager launches, run after run (graph, eager, graph)."""
    prob = synth.SMALL[1]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    pm = L.Problem(**kw)
    pm.set_data(blended)
    outs = []
    for eager in (False, True, False):
        if eager:
            monkeypatch.setenv("AGD_NO_GRAPH", "1")
        else:
            monkeypatch.delenv("AGD_NO_GRAPH", raising=False)
        pm.run(1.0, 12)
        outs.append(pm.output())
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)
    pm.close()


@pytest.mark.parametrize("cfg", [1, 3])
def test_window_chunk_order_matches(monkeypatch, cfg):
    """Iterations in window-chunk order (adjoint and next forward chunk by
    chunk) give bit-identical results to the whole-volume order, eager and
    replayed from a graph."""
    prob = synth.SMALL[cfg]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    outs = []
    for chunk_mb, graph in (("0", True), ("0.5", False), ("0.5", True)):
        monkeypatch.setenv("AGD_CHUNK_MB", chunk_mb)
        if graph:
            monkeypatch.delenv("AGD_NO_GRAPH", raising=False)
        else:
            monkeypatch.setenv("AGD_NO_GRAPH", "1")
        pm = L.Problem(**kw)
        pm.set_data(blended)
        pm.run(1.0, 9)
        outs.append(pm.output())
        pm.close()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("cfg", [1, 3])
def test_window_chunk_order_matches_with_warm_start(monkeypatch, cfg):
    """A warm start hands its guard's residual to the first iteration; in
    window-chunk order that iteration starts from the same record, so the
    results are bit-identical to the whole-volume order. On the vibroseis
    workload (cfg 3) the model is the data before the wavelet, so a data
    estimate used as the model fits worse than zero: the guard falls back,
    in both orders alike."""
    prob = synth.SMALL[cfg]()
    g = prob["geometry"]
    kw = _kw(g)
    blended = synth.blend_input(prob, O.blend_same_geometry)
    warm = [(0.5 * (np.asarray(a, np.float64) + b)).astype(np.float32)
            for a, b in zip(blended, prob["truth"])]
    outs = []
    for chunk_mb in ("0", "0.5"):
        monkeypatch.setenv("AGD_CHUNK_MB", chunk_mb)
        pm = L.Problem(**kw)
        pm.set_data(blended)
        pm.run_warm(1.0, 9, warm)
        assert L.last_warm_fallback() == (cfg == 3)
        outs.append(pm.output())
        pm.close()
    for a, b in zip(outs[1], outs[0]):
        assert np.array_equal(a, b)
