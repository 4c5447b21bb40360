"""The learned warm start on data unlike its training set.

The network was trained on volumes of the 64 x 1024 x 2048 workload's kind
(20 Hz wavelet, shots every 2 s +- 0.5 s, all traces live). Here it meets
the 64 x 512 x 2048 batching workload (25 Hz wavelet, shots every 4 s +- 1 s,
5% bad and 3% missing traces): the network's own SNR, the guard's residual
ratio ||A x_warm - d|| / ||d||, and the SNR of the warm and the zero start
after the same number of iterations are printed; the warm start must not end
below the zero start at any of them (if the guard rejects the estimate, the
plain run executes and the two are equal), and a rejected estimate must come
back bit-identical to the plain run.
"""
import numpy as np
import pytest

from agdeblend_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

F = 0.03  # the warm start's initial_factor (as in the bench)


def test_warm_start_on_unseen_geometry():
    from agdeblend_amd import agdeblend as agd
    from agdeblend_amd import lib as L
    from agdeblend_amd import warmstart as W
    prob = synth.config1(interval_s=4.0, dither_s=1.0)
    blended = synth.blend_input(prob, agd.blend)
    g = prob["geometry"]
    kw = {k: g[k] for k in ("volumes", "n_dims", "window_shapes", "coords", "shapes",
                            "shottimes", "channels", "trace_types")}
    pm = L.Problem(**kw)
    try:
        net = W.UNet().load(W.WEIGHTS)
        assert W.accepts(pm.shapes)
        pm.set_data(blended)
        est = W.estimate(pm, net)
        torch.cuda.synchronize()
        snr_in = synth.snr(prob, blended)
        snr_net = synth.snr(prob, [est.cpu().numpy().reshape(prob["truth"][0].shape)])
        print("unseen geometry: input %.2f dB, network estimate %.2f dB" % (snr_in, snr_net))
        for n_its in (2, 5, 20, 60):
            pm.set_data(blended)
            pm.run(1.0, n_its)
            plain = pm.output()
            snr_plain = synth.snr(prob, plain)
            pm.set_data(blended)
            W.deblend_warm(pm, net, F, n_its)
            warm = pm.output()
            snr_warm = synth.snr(prob, warm)
            fell_back = L.last_warm_fallback()
            print("  n_its %3d: zero start %6.2f dB, warm start %6.2f dB (guard ratio %.3f%s)"
                  % (n_its, snr_plain, snr_warm, L.last_warm_ratio() or -1.0,
                     ", fell back" if fell_back else ""))
            if fell_back:
                pm.set_data(blended)
                pm.run(F, n_its)
                ref = pm.output()
                for a, b in zip(warm, ref):
                    assert np.array_equal(a, b)
            else:
                assert snr_warm >= snr_plain, (n_its, snr_warm, snr_plain)
    finally:
        pm.close()
