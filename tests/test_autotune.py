"""The autotuner (time-transform factorizations, then the spatial tile):
its choices are among the candidates, the problem it leaves behind still
matches the FP64 oracle, re-tuning is stable (the same classes, a valid
choice), and every factorization it may pick gives the oracle's operators."""
import zlib

import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O
from tests.test_gpu_parity import (SPEC, _kw, _mute_vec, _random_model, _spec_geometry,
                                   _to_oracle_layout, _to_product_layout)

pytestmark = pytest.mark.gpu

CAND = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32}
# (M1, M2) factorizations of the time half length M = Pt / 2 with a kernel
TIME_CAND = {40: [(5, 8), (8, 5), (4, 10)], 80: [(10, 8), (8, 10), (16, 5)],
             160: [(10, 16), (16, 10), (20, 8), (8, 20)],
             320: [(20, 16), (16, 20), (32, 10), (10, 32)]}


@pytest.mark.parametrize("wt", [64, 128, 256, 512])
def test_every_time_factorization_matches_fp64_oracle(wt):
    """Forward and adjoint operators with each candidate factorization of
    the time transform and each variant of the forward kernel: relative
    error <= 2e-6 against the FP64 oracle and the FP64 dot test to 1e-6."""
    geo = _spec_geometry((64, 2048), (32, wt), 40 + wt)
    kw = _kw(geo)
    om = O.Model(double=True, **kw)
    pm = L.Problem(**kw)
    try:
        fac = pm.time_factors()
        M = 5 * wt // 8  # Pt / 2 with Pt = 5 Wt / 4
        assert fac == [TIME_CAND[M][0]], fac
        rng = np.random.default_rng(zlib.crc32(b"tf%d" % wt))
        x = _random_model(om, rng)
        y = _mute_vec(om, rng.random(om.record_len) - 0.5)
        ref = om.forward(x)
        gref = om.adjoint_update(y, -1.0, om.zeros_model())
        for (m1, m2), var in [(f, v) for f in TIME_CAND[M] for v in (0, 1)]:
            pm.set_time_factors(0, m1, m2)
            pm.set_time_variant(0, var)
            assert pm.time_factors() == [(m1, m2)] and pm.time_variants() == [var]
            got = pm.forward(_to_product_layout(om, x).astype(np.complex64)).astype(np.float64)
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert rel <= 2e-6, (m1, m2, var, rel)
            gp = _to_oracle_layout(om, pm.adjoint(y.astype(np.float32)).astype(np.complex128))
            rel_g = np.linalg.norm(gp - gref) / np.linalg.norm(gref)
            assert rel_g <= 2e-6, (m1, m2, rel_g)
            lhs = float(np.dot(got, y))
            rhs = float(np.sum(x.real * gp.real + x.imag * gp.imag))
            assert abs(lhs - rhs) <= 1e-6 * np.linalg.norm(got) * np.linalg.norm(y)
        with pytest.raises(ValueError):
            pm.set_time_factors(0, 7, 7)
    finally:
        pm.close()


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
        tf = pm.time_factors()
        print(name, "nk", nk, "ms", ["%.3f" % t for t in ms], "time factors", tf,
              "variants", pm.time_variants())
        assert all(f in TIME_CAND[f[0] * f[1]] for f in tf if f[0])
        assert all(v in (0, 1) for v in pm.time_variants())
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
