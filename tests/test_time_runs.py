"""Time blocks split into runs of time windows.

A block of the specialised time kernels transforms one 16-trace tile over
a run [tw0, tw1) of its window's time windows: all of them in the
whole-volume order, at most 4 in the window-chunk order (AGD_CHUNK_MB),
AGD_TIME_TW runs when set. The forward block of a run with tw0 > 0
transforms time window tw0 - 1 again for its overlap carry, so every
record sample is the same sum as in one run (fk.c:307-452 computes each
window's traces in one pass).

CPU: the runs of every tile partition [0, ntw) into nearly equal parts,
and the generic kernels' blocks keep one run. GPU: deblending with split
runs, eager and replayed from a graph, both forward kernel variants, in
whole-volume and window-chunk order, is bit-identical to one run.
"""
import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from tests.test_gpu_parity import SPEC, _kw

GEOS = [g for g in SPEC if g[0] in ("spec_20x40_320", "spec_40x40_640", "spec_1x40_320")]
GEOS.append(("c0_small", synth.SMALL[0]()["geometry"]))


def _blocks(kw):
    pm = L.Problem(host_only=True, **kw)
    try:
        tb = pm.time_blocks()
        ntw = {w: pm.window(w)["ntw"] for w in range(pm.n_windows)}
        generic = [f == (0, 0) for f in pm.time_factors()]
        return tb, ntw, all(generic)
    finally:
        pm.close()


@pytest.mark.parametrize("run,chunk,expect", [(None, None, 0), ("3", None, 3), ("1", None, 1),
                                              (None, "0.5", 4), ("2", "0.5", 2), ("0", "0.5", 0)])
@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_runs_partition_each_tile(name, geo, run, chunk, expect, monkeypatch):
    for k, v in (("AGD_TIME_TW", run), ("AGD_CHUNK_MB", chunk)):
        if v is None:
            monkeypatch.delenv(k, raising=False)
        else:
            monkeypatch.setenv(k, v)
    tb, ntw, generic = _blocks(_kw(geo))
    assert not generic
    tiles = {}
    for w, t0, a, b in tb.tolist():
        tiles.setdefault((w, t0), []).append((a, b))
    for (w, t0), runs in tiles.items():
        n = ntw[w]
        assert runs[0][0] == 0 and runs[-1][1] == n
        assert all(r[1] == s[0] for r, s in zip(runs, runs[1:]))
        lens = [b - a for a, b in runs]
        assert max(lens) - min(lens) <= 1
        if expect == 0:
            assert runs == [(0, n)]
        else:
            assert max(lens) <= expect and len(runs) == -(-n // expect)


def test_generic_time_blocks_keep_one_run(monkeypatch):
    monkeypatch.setenv("AGD_TIME_TW", "2")
    tb, ntw, generic = _blocks(_kw(synth.SMALL[1]()["geometry"]))
    assert generic
    for w, t0, a, b in tb.tolist():
        assert (a, b) == (0, ntw[w])


@pytest.mark.gpu
@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_split_runs_deblend_bit_identical(name, geo, monkeypatch):
    kw = _kw(geo)
    rng = np.random.default_rng(17)
    truth = [rng.standard_normal(tuple(s)).astype(np.float32) for s in geo["shapes"]]
    from oracle import oracle as O
    blended = synth.blend_input(dict(geometry=geo, truth=truth), O.blend_same_geometry)

    def deblend(run, chunk, variant, graph):
        for k, v in (("AGD_TIME_TW", run), ("AGD_CHUNK_MB", chunk),
                     ("AGD_NO_GRAPH", None if graph else "1")):
            if v is None:
                monkeypatch.delenv(k, raising=False)
            else:
                monkeypatch.setenv(k, v)
        pm = L.Problem(**kw)
        try:
            for c, f in enumerate(pm.time_factors()):
                if f != (0, 0):
                    pm.set_time_variant(c, variant)
            pm.set_data(blended)
            pm.run(1.0, 9)
            return pm.output()
        finally:
            pm.close()

    for variant in (0, 1):
        ref = deblend(None, None, variant, True)
        cases = [("1", None, True), ("3", None, False), (None, "0.5", True), ("2", "0.5", False)]
        for run, chunk, graph in cases:
            out = deblend(run, chunk, variant, graph)
            print(name, "variant", variant, "run", run, "chunk", chunk, "graph", graph, flush=True)
            for a, b in zip(out, ref):
                assert np.array_equal(a, b)
