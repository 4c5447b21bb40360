"""Parity at the workloads' full sizes (run on an MI355X).

The oracle comparisons elsewhere use sizes the oracle finishes in seconds;
these run the full C1 and C4 geometries, where offsets exceed 32 bits in
bytes and the grids have 10^5 blocks:
  * C4: the blended input record (agd_blend's mean blend and mute of the
    patch data) is bit-exact to the oracle's;
  * C1: the forward operator agrees with the FP64 oracle to 2e-6 relative,
    and the GPU forward and adjoint pass the dot test in FP64 to 1e-6;
    agd_deblend after three iterations agrees to 1e-4;
  * C4: agd_deblend after two iterations, and the learned warm start (the
    network's estimate given to both sides), agree to 1e-4;
  * C2 with the bf16 x 3 spatial DFTs: agd_deblend after two iterations
    agrees to 1e-4;
  * C0 (all 100 iterations) and C3 (three iterations, with the wavelet):
    agd_deblend agrees to 1e-4;
  * C3 over its full 150 iterations and C1 over 20 iterations against the
    FP64 oracle's output kept in tests/golden/full_size_deblend.npz (made
    by tests/golden/make_full_size.py: the oracle needs minutes for these):
    the kept traces within 1e-4 relative L2, the whole output's energy and
    random projections within the bounds a 1e-4 relative error allows.
"""
import os

import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd import synth
from oracle import oracle as O
from tests.test_gpu_parity import _kw, _mute_vec, _random_model, _to_oracle_layout, \
    _to_product_layout

pytestmark = pytest.mark.gpu


def test_c4_blended_input_record_is_bit_exact():
    prob = synth.config4()
    g = prob["geometry"]
    kw = _kw(g)
    om = O.Model(double=False, **kw)
    ref = om.blend_input(prob["truth"], g["shottimes"], g["channels"], g["trace_types"])
    pm = L.Problem(**kw)
    try:
        got = pm.set_data(prob["truth"], want_record=True)
    finally:
        pm.close()
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_c1_operators_at_full_size():
    prob = synth.config1(interval_s=4.0, dither_s=1.0)
    kw = _kw(prob["geometry"])
    om = O.Model(double=True, **kw)
    rng = np.random.default_rng(2024)
    x = _random_model(om, rng)
    ref = om.forward(x)
    pm = L.Problem(**kw)
    try:
        got = pm.forward(_to_product_layout(om, x).astype(np.complex64)).astype(np.float64)
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel <= 2e-6, rel
        y = _mute_vec(om, rng.random(om.record_len) - 0.5)
        gp = _to_oracle_layout(om, pm.adjoint(y.astype(np.float32)).astype(np.complex128))
    finally:
        pm.close()
    lhs = float(np.dot(got, y))
    rhs = float(np.sum(x.real * gp.real + x.imag * gp.imag))
    assert abs(lhs - rhs) <= 1e-6 * np.linalg.norm(got) * np.linalg.norm(y), (lhs, rhs)


def test_c1_deblend_at_full_size():
    """agd_deblend on the full C1 volume (initial update, two thresholded
    iterations, output) within 1e-4 relative L2 of the FP64 oracle."""
    prob = synth.config1(interval_s=4.0, dither_s=1.0)
    kw = _kw(prob["geometry"])
    blended = synth.blend_input(prob, O.blend_same_geometry)
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=3, double=True, **kw)
    pm = L.Problem(**kw)
    try:
        pm.set_data(blended)
        pm.run(1.0, 3)
        out = pm.output()
    finally:
        pm.close()
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    rel = (num / den) ** 0.5
    assert rel <= 1e-4, rel


def _rel(out, ref):
    num = sum(float(np.sum((np.asarray(a, np.float64) - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(np.asarray(b, np.float64) ** 2)) for b in ref)
    return (num / den) ** 0.5


def test_c4_deblend_and_learned_warm_start_at_full_size():
    """The headline workload: agd_deblend on the whole C4 volume (initial
    update, one thresholded iteration, output) within 1e-4 of the FP64
    oracle; and the learned warm start with one iteration (the network's
    estimate handed to both the library and the oracle: warm-start model,
    guard, output) within 1e-4 of the oracle from the same estimate."""
    import torch
    from agdeblend_amd import warmstart as W
    torch.cuda.init()
    prob = synth.config4()
    kw = _kw(prob["geometry"])
    blended = synth.blend_input(prob, O.blend_same_geometry)
    pm = L.Problem(**kw)
    try:
        pm.set_data(blended)
        pm.run(1.0, 2)
        plain = pm.output()
        net = W.UNet().load(W.WEIGHTS)
        pm.set_data(blended)
        est = W.deblend_warm(pm, net, 0.03, 1)
        assert est is not None and not L.last_warm_fallback()
        warm = pm.output()
        shape = tuple(int(v) for v in pm.shapes[0])
        est_host = [est.cpu().numpy().reshape(shape)]
    finally:
        pm.close()
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=2, double=True, **kw)
    assert _rel(plain, ref) <= 1e-4, _rel(plain, ref)
    ref_w, st = O.deblend(data=blended, initial_factor=0.03, n_its=1, double=True,
                          warm=est_host, **kw)
    assert st[2] == 0
    assert _rel(warm, ref_w) <= 1e-4, _rel(warm, ref_w)


def test_c2_split_precision_deblend_at_full_size():
    """C2 with the bf16 x 3 matrix-core spatial DFTs: agd_deblend after two
    iterations within 1e-4 of the FP64 oracle (the error is printed)."""
    prob = synth.config2(interval_s=4.0, dither_s=1.0)
    kw = _kw(prob["geometry"])
    blended = synth.blend_input(prob, O.blend_same_geometry)
    pm = L.Problem(**kw)
    try:
        assert pm.set_spatial_precision(1) >= 1
        pm.set_data(blended)
        pm.run(1.0, 2)
        x3 = pm.output()
        assert pm.set_spatial_precision(0) == 0
        pm.set_data(blended)
        pm.run(1.0, 2)
        fp32 = pm.output()
    finally:
        pm.close()
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=2, double=True, **kw)
    print("C2 full size, agd_deblend 2 iterations vs FP64: bf16x3 %.2e, FP32 %.2e"
          % (_rel(x3, ref), _rel(fp32, ref)))
    assert _rel(x3, ref) <= 1e-4
    assert _rel(fp32, ref) <= 1e-4


@pytest.mark.parametrize("cfg,n_its", [(0, 100), (3, 3)])
def test_full_size_deblend_small_configs(cfg, n_its):
    """C0 (the 2-D workload, its full 100 iterations) and C3 (vibroseis,
    with the wavelet convolution and a recording gap) within 1e-4 of the
    FP64 oracle."""
    prob = synth.WORKLOADS[cfg]()
    kw = _kw(prob["geometry"])  # with the wavelets, when the workload has them
    blended = synth.blend_input(prob, O.blend_same_geometry)
    pm = L.Problem(**kw)
    try:
        pm.set_data(blended)
        pm.run(1.0, n_its)
        out = pm.output()
    finally:
        pm.close()
    ref, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
    rel = _rel(out, ref)
    print("C%d full size, agd_deblend %d iterations vs FP64: %.2e" % (cfg, n_its, rel))
    assert rel <= 1e-4, rel


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "full_size_deblend.npz")


@pytest.mark.parametrize("name", ["c3", "c1"])
def test_long_full_size_deblend_against_fp64_golden(name):
    from tests.golden import make_full_size as G
    z = np.load(GOLDEN, allow_pickle=False)
    make, n_its, stride = G.CASES[name]
    assert int(z[name + "_n_its"]) == n_its and int(z[name + "_stride"]) == stride
    prob = make()
    kw = _kw(prob["geometry"])
    blended = synth.blend_input(prob, O.blend_same_geometry)
    assert G.input_hash(blended) == str(z[name + "_hash"])  # the fixture's input, bit for bit
    print("%s full size, %d iterations" % (name, n_its), flush=True)
    pm = L.Problem(**kw)
    try:
        pm.set_data(blended)
        pm.run(1.0, n_its)
        out = pm.output()
    finally:
        pm.close()
    traces, energy, proj = G.summary(out, stride)
    ref_traces = z[name + "_traces"].astype(np.float64)
    rel = np.linalg.norm(traces - ref_traces) / np.linalg.norm(ref_traces)
    ref_energy = float(z[name + "_energy"])
    n = sum(int(np.prod(np.shape(o))) for o in out)
    print("%s, %d iterations vs FP64 oracle: kept traces %.2e, energy %.2e, projections %s"
          % (name, n_its, rel, abs(energy - ref_energy) / ref_energy,
             np.abs(proj - z[name + "_proj"]) / np.sqrt(ref_energy * n)))
    assert rel <= 1e-4
    # a relative L2 error e of the whole output moves its energy by at most
    # about 2e and each +-1 projection by at most e ||out|| sqrt(n)
    assert abs(energy - ref_energy) <= 2.0e-4 * ref_energy
    assert np.all(np.abs(proj - z[name + "_proj"]) <= 1e-4 * np.sqrt(ref_energy * n))
