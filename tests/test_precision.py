"""Split-precision spatial DFTs (bf16 x 3 on the matrix cores) beside the
FP32 register FFTs, both against the FP64 oracle.

Each bf16 x 3 product drops the lo.lo term of the split operands (~2^-16
relative) and rounds the lo parts to bf16, so the operators carry ~1e-5
relative error where the FP32 FFTs carry ~1e-7. Bars (written in the
tests): operators <= 2e-5 relative to the FP64 oracle, the FP64-accumulated
dot test <= 1e-6 relative (the FP32 path's bar), agd_deblend after 6
iterations <= 1e-4 (the FP32 path's bar). The measured errors are printed
(pytest -s) next to the FP32 path's."""
import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O
from tests.test_gpu_parity import (_kw, _mute_vec, _random_model, _spec_geometry,
                                   _to_oracle_layout, _to_product_layout)

pytestmark = pytest.mark.gpu

GEOS = [("square_40x40_640", _spec_geometry((32, 64, 1024), (32, 32, 512), 21)),
        ("square_20x20_320", _spec_geometry((16, 48, 512), (16, 16, 256), 22))]


def _op_errors(om, pm, rng):
    x = _random_model(om, rng)
    ref = om.forward(x)
    got = pm.forward(_to_product_layout(om, x).astype(np.complex64)).astype(np.float64)
    e_fwd = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    y = _mute_vec(om, rng.random(om.record_len) - 0.5)
    gref = om.adjoint_update(y, -1.0, om.zeros_model())
    gp = _to_oracle_layout(om, pm.adjoint(y.astype(np.float32)).astype(np.complex128))
    e_adj = np.linalg.norm(gp - gref) / np.linalg.norm(gref)
    lhs = float(np.dot(got, y))
    rhs = float(np.sum(x.real * gp.real + x.imag * gp.imag))
    e_dot = abs(lhs - rhs) / float(np.linalg.norm(got) * np.linalg.norm(y))
    return e_fwd, e_adj, e_dot


@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_bf16x3_operators_against_fp64_oracle(name, geo):
    kw = _kw(geo)
    om = O.Model(double=True, **kw)
    pm = L.Problem(**kw)
    try:
        assert pm.set_spatial_precision(1) >= 1
        e3 = _op_errors(om, pm, np.random.default_rng(5))
        assert pm.set_spatial_precision(0) == 0
        e32 = _op_errors(om, pm, np.random.default_rng(5))
    finally:
        pm.close()
    print("%s  bf16x3: forward %.2e adjoint %.2e dot %.2e | fp32: forward %.2e adjoint %.2e "
          "dot %.2e" % ((name,) + e3 + e32))
    assert e3[0] <= 2e-5 and e3[1] <= 2e-5 and e3[2] <= 1e-6, e3
    # the split precision is coarser than FP32 but far finer than bf16 (4e-3)
    assert e3[0] > e32[0] and e3[0] < 4e-3


@pytest.mark.parametrize("name,geo", GEOS, ids=[g[0] for g in GEOS])
def test_bf16x3_deblend_against_fp64_oracle(name, geo):
    kw = _kw(geo)
    rng = np.random.default_rng(8)
    truth = [rng.standard_normal(tuple(geo["shapes"][0])).astype(np.float32)]
    blended = synth.blend_input(dict(geometry=geo, truth=truth), O.blend_same_geometry)
    n_its = 6
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
    errs = []
    for mode in (1, 0):
        pm = L.Problem(**kw)
        try:
            pm.set_spatial_precision(mode)
            pm.set_data(blended)
            pm.run(1.0, n_its)
            out = pm.output()
        finally:
            pm.close()
        errs.append(np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0]))
    print("%s  deblend %d iterations: bf16x3 %.2e, fp32 %.2e" % (name, n_its, errs[0], errs[1]))
    assert errs[0] <= 1e-4 and errs[1] <= 1e-4, errs
