"""The plain zero-start inversion as the warm start's fallback.

A warm start is kept only when its model fits the data better than the zero
model and better than the data itself as a prediction of the record,
||A x_warm - d|| < min(||d||, ||mute(sum blend(data)) - d||) (FP64 norms);
otherwise the inversion runs from the zero start exactly as agd_deblend does
(agdeblend.c:584-629). The oracle and the library apply the same rule. An
estimate between the blended input and the truth passes it; the input pushed
further from the truth (better than zero, worse than the input) and a large
random field do not.
"""
import numpy as np
import pytest

from agdeblend_amd import synth
from oracle import oracle as O

LIVE, BAD, MISSING = 0, 1, 2
N_ITS = 6


def _kw(g):
    keys = ["volumes", "n_dims", "window_shapes", "coords", "shapes", "shottimes",
            "channels", "trace_types"]
    return {k: g[k] for k in keys if k in g}


def _case(cfg):
    prob = synth.SMALL[cfg]()
    kw = _kw(prob["geometry"])
    blended = synth.blend_input(prob, O.blend_same_geometry)
    good = [0.5 * (np.asarray(a, np.float64) + b) for a, b in zip(blended, prob["truth"])]
    rng = np.random.default_rng(5)
    scale = max(float(np.abs(np.asarray(a)).max()) for a in blended)
    bad = [20.0 * scale * rng.standard_normal(np.shape(a)) for a in blended]
    poor = [1.5 * np.asarray(a, np.float64) - 0.5 * b for a, b in zip(blended, prob["truth"])]
    return kw, blended, good, poor, bad


def _rel(out, ref):
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    return (num / den) ** 0.5


@pytest.mark.parametrize("cfg", [0, 1])
def test_oracle_guard(cfg):
    """Oracle: the random field and the estimate worse than the input fall
    back and then give the plain run's output bit for bit; the intermediate
    estimate is kept."""
    kw, blended, good, poor, bad = _case(cfg)
    plain, _ = O.deblend(data=blended, n_its=N_ITS, threads=1, **kw)
    for warm in (bad, poor):
        out, st = O.deblend(data=blended, n_its=N_ITS, threads=1, warm=warm, **kw)
        assert st[2] == 1
        for a, b in zip(out, plain):
            assert np.array_equal(a, b)
    out, st = O.deblend(data=blended, n_its=N_ITS, threads=1, warm=good, **kw)
    assert st[2] == 0
    assert _rel(out, plain) > 1e-3  # the warm start changed the result


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [0, 1])
def test_library_guard_matches_oracle(cfg):
    """Library: the same decision as the oracle for both estimates; after a
    fallback the output is bit-identical to the plain GPU inversion and
    within 1e-4 of the FP64 oracle."""
    from agdeblend_amd import lib as L
    kw, blended, good, poor, bad = _case(cfg)
    pm = L.Problem(**kw)
    pm.set_data(blended)
    pm.run(1.0, N_ITS)
    plain = pm.output()
    for warm, expect in ((bad, 1), (poor, 1), (good, 0)):
        ref, st = O.deblend(data=blended, n_its=N_ITS, double=True, warm=warm, **kw)
        assert st[2] == expect
        pm.set_data(blended)
        pm.run_warm(1.0, N_ITS, [a.astype(np.float32) for a in warm])
        assert int(L.last_warm_fallback()) == expect
        out = pm.output()
        assert _rel(out, ref) <= 1e-4, _rel(out, ref)
        if expect:
            for a, b in zip(out, plain):
                assert np.array_equal(a, b)
    pm.close()


@pytest.mark.gpu
def test_deblend_warm_without_network_runs_plain():
    """warmstart.deblend_warm with no network (or patches the network cannot
    take) runs the plain zero-start inversion."""
    import torch
    from agdeblend_amd import lib as L
    from agdeblend_amd import warmstart as W
    torch.cuda.init()
    kw, blended, good, poor, bad = _case(1)
    pm = L.Problem(**kw)
    pm.set_data(blended)
    pm.run(1.0, N_ITS)
    plain = pm.output()
    pm.set_data(blended)
    assert W.deblend_warm(pm, None, 1.0, N_ITS) is None
    for a, b in zip(pm.output(), plain):
        assert np.array_equal(a, b)
    pm.close()
    assert not W.accepts([(32, 30, 256)]) and W.accepts([(32, 32, 256)])


def test_network_patch_rule():
    """The network takes 3-D patches with every size a multiple of 4 (two
    2x2x2 poolings); anything else goes to the plain inversion."""
    from agdeblend_amd import warmstart as W
    assert W.accepts([(32, 32, 256), (8, 64, 512)])
    assert not W.accepts([(32, 30, 256)])
    assert not W.accepts([(32, 32, 256), (6, 32, 256)])
    assert not W.accepts([(256, 1024)])
