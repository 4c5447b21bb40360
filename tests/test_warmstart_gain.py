"""What the trained warm-start network gains over the plain inversion, on a
seeded volume of the 64 x 1024 x 2048 workload's kind that the network was
not trained on (half the channels and a quarter of the sources, same event
statistics, shots every 2 s +- 0.5 s): the estimate alone, and the same
short iteration count from the zero start and from the estimate."""
import numpy as np
import pytest

from agdeblend_amd import synth

torch = pytest.importorskip("torch")


def test_committed_weights_file():
    from agdeblend_amd import warmstart as W
    z = np.load(W.WEIGHTS, allow_pickle=False)
    assert [str(s) for s in z["layers"]] == ["%s:%d:%d" % l for l in W.LAYERS]
    n = sum(27 * ci * co + co for _, ci, co in W.LAYERS)
    assert z["w"].shape == (n,) and np.all(np.isfinite(z["w"]))
    assert float(z["scale"]) > 0


@pytest.mark.gpu
def test_learned_warm_start_beats_zero_start():
    from agdeblend_amd import agdeblend as agd
    from agdeblend_amd import lib as L
    from agdeblend_amd import warmstart as W
    prob = synth.config1(seed=777, n_ch=32, n_src=256, n_t=2048, windows=(32, 32, 256),
                         p_bad=0.0, p_missing=0.0)
    prob["blended"] = synth.blend_input(prob, agd.blend)
    g = prob["geometry"]
    kw = {k: g[k] for k in ("volumes", "n_dims", "window_shapes", "coords", "shapes",
                            "shottimes", "channels", "trace_types")}
    pm = L.Problem(**kw)
    try:
        net = W.UNet().load(W.WEIGHTS)
        snr_in = synth.snr(prob, prob["blended"])
        pm.set_data(prob["blended"])
        est = W.estimate(pm, net)
        torch.cuda.synchronize()
        snr_net = synth.snr(prob, [est.cpu().numpy().reshape(prob["truth"][0].shape)])
        n_its = 5
        pm.run(1.0, n_its)
        snr_plain = synth.snr(prob, pm.output())
        pm.set_data(prob["blended"])
        W.deblend_warm(pm, net, 0.03, n_its)
        snr_warm = synth.snr(prob, pm.output())
        print("input %.2f dB, network %.2f dB, %d iterations: zero start %.2f dB, "
              "warm start %.2f dB" % (snr_in, snr_net, n_its, snr_plain, snr_warm))
        assert snr_net > snr_in + 15.0
        assert snr_warm >= 20.0
        assert snr_warm > snr_plain + 10.0
    finally:
        pm.close()
