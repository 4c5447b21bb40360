"""The reference's end-to-end deblending test (src/test.cpp:432-561) on its
own inputs (regenerated with glibc rand(), see reference_cases.py).

CPU: the oracle (float32, the reference's AGD_TYPE) meets the test's
criterion ||out - true||^2 < ||true||^2 / 10 after 800 iterations on all five
problems. GPU: agd_deblend of the HIP build meets the same criterion, and its
output stays within 1e-4 relative L2 of the float64 oracle after the full
800 iterations (the float32 oracle itself ends 1e-7 to 4e-7 from the float64
one on these problems, so ISTA does not amplify the rounding here; 1e-4 is
the bar of the shorter runs in test_gpu_parity).
"""
import numpy as np
import pytest

import reference_cases as R
from oracle import oracle as O

CASES = R.cases()


def _geo(c):
    return {k: c[k] for k in ("volumes", "n_dims", "window_shapes", "coords", "shapes",
                              "shottimes", "channels", "trace_types")}


def _blend_oracle(c):
    return O.blend_same_geometry(c["shottimes"], c["channels"], c["trace_types"], c["truth"],
                                 0)


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_meets_reference_criterion(i):
    c = CASES[i]
    blended = _blend_oracle(c)
    out, _ = O.deblend(data=blended, initial_factor=1.0, n_its=R.N_ITS, double=False,
                       **_geo(c))
    ok, ratio = R.passes(out, c["truth"])
    assert ok, ratio


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(len(CASES)))
def test_hip_meets_reference_criterion_and_oracle(i):
    from agdeblend_amd import agdeblend as agd
    c = CASES[i]
    # the wrapper shapes its output like shottimes (flat here), as the
    # reference's does: back to the patch shapes
    blended = agd.blend(c["shottimes"], c["channels"], c["trace_types"], c["truth"],
                        agd.AGDBlendSum)
    blended = [np.asarray(a).reshape(t.shape) for a, t in zip(blended, c["truth"])]
    ref_blend = _blend_oracle(c)
    for a, b in zip(blended, ref_blend):
        assert np.array_equal(a, b)
    out = agd.deblend(c["volumes"], c["window_shapes"], c["coords"], c["shottimes"],
                      c["channels"], c["trace_types"], 1.0, R.N_ITS, 0, blended)
    ok, ratio = R.passes(out, c["truth"])
    assert ok, ratio
    ref, _ = O.deblend(data=ref_blend, initial_factor=1.0, n_its=R.N_ITS, double=True,
                       **_geo(c))
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    err = (num / den) ** 0.5
    print("problem %d: %d iterations, relative L2 to the float64 oracle %.3e" % (i, R.N_ITS, err))
    assert err <= 1e-4, err
