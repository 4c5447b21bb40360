"""The spatial-tile autotuner: its choice is one of the candidates, the
problem it leaves behind still matches the FP64 oracle, and re-tuning is
stable (the same classes, a valid choice)."""
import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O
from tests.test_gpu_parity import SPEC, _kw

pytestmark = pytest.mark.gpu

CAND = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32}


@pytest.mark.parametrize("name,geo", SPEC[:3], ids=[g[0] for g in SPEC[:3]])
def test_autotuned_problem_matches_fp64_oracle(name, geo):
    kw = _kw(geo)
    rng = np.random.default_rng(31)
    truth = [rng.standard_normal(tuple(geo["shapes"][0])).astype(np.float32)]
    blended = synth.blend_input(dict(geometry=geo, truth=truth), O.blend_same_geometry)
    n_its = 6
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
    pm = L.Problem(**kw)
    try:
        pm.set_data(blended)
        nk, ms = pm.autotune(reps=2)
        print(name, "nk", nk, "ms", ["%.3f" % t for t in ms])
        assert len(nk) >= 1 and all(k in CAND for k in nk if k)
        assert all(t > 0 for k, t in zip(nk, ms) if k)
        pm.run(1.0, n_its)
        out = pm.output()
        nk2, _ = pm.autotune(reps=1)
        assert len(nk2) == len(nk) and all(k in CAND for k in nk2 if k)
    finally:
        pm.close()
    rel = np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0])
    assert rel <= 1e-4, rel
