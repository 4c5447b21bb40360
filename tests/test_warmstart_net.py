"""The warm-start network on the GPU: its forward pass and every layer's
weight and bias gradient against the FP64 numpy restatement of the same
network (with the GPU's bf16 weights and input), and a short training run
whose loss must fall."""
import numpy as np
import pytest

from tests import unet_ref as R

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _net(seed):
    from agdeblend_amd.warmstart import UNet
    return UNet(seed=seed, head_scale=1.0)


def _bf_params(net):
    P = net.params()
    return {k: (R.bf16_round(w), b.astype(np.float64)) for k, (w, b) in P.items()}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_network_forward_and_gradients_match_fp64_reference():
    net = _net(3)
    net.scale = 2.0
    rng = np.random.default_rng(7)
    shape = (2, 4, 8, 64)
    field = rng.standard_normal(shape).astype(np.float32)
    target = (field + 0.3 * rng.standard_normal(shape)).astype(np.float32)
    f_d = torch.tensor(field, device="cuda")
    t_d = torch.tensor(target, device="cuda")
    lpart = net.train_step(f_d, t_d, lr=0.0)  # lr 0: gradients only, weights kept
    torch.cuda.synchronize()
    G = net.grads()
    B = net._buffers(shape, True)
    head = B["head"].float().cpu().numpy().astype(np.float64)
    # the reference on the same bf16 weights and bf16 input
    P = _bf_params(net)
    x = np.zeros(shape + (16,))
    x[..., 0] = R.bf16_round(field * net.scale)
    out, A = R.unet_forward(x, P)
    assert _rel(head, out) < 3e-2
    resid = (target.astype(np.float64) - field) * net.scale
    loss, dout = R.unet_loss(out, resid)
    gpu_loss = float(lpart.sum()) / field.size
    assert abs(gpu_loss - loss) <= 3e-2 * loss
    Gr = R.unet_backward(dout, P, A)
    for name, _, _ in R.LAYERS:
        assert _rel(G[name][0], Gr[name][0]) < 5e-2, name
        assert _rel(G[name][1], Gr[name][1]) < 5e-2, name


def test_prediction_is_input_plus_scaled_head():
    net = _net(4)
    net.scale = 0.5
    rng = np.random.default_rng(8)
    field = rng.standard_normal((1, 4, 4, 128)).astype(np.float32)
    f_d = torch.tensor(field, device="cuda")
    y = net.predict_unfused(f_d).cpu().numpy()
    head = net._buffers(field.shape, True)["head"].float().cpu().numpy()
    assert np.allclose(y, field + head[..., 0] / 0.5, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("shape", [(1, 8, 12, 200), (2, 4, 8, 64)])
def test_inference_pass_matches_training_pass(shape):
    """The inference kernels (first layer from the field, concatenations read
    in place, fp32 head channel) against the training pass's kernels: the
    same network, so the estimates differ only by rounding (bf16 of the
    head's output, fp32 summation order)."""
    net = _net(6)
    net.scale = 1.7
    rng = np.random.default_rng(10)
    field = rng.standard_normal(shape).astype(np.float32)
    f_d = torch.tensor(field, device="cuda")
    a = net.predict(f_d).cpu().numpy().astype(np.float64) - field
    b = net.predict_unfused(f_d).cpu().numpy().astype(np.float64) - field
    assert _rel(a, b) < 2e-2


def test_training_reduces_the_loss():
    from agdeblend_amd.warmstart import UNet
    net = UNet(seed=5)
    rng = np.random.default_rng(9)
    shape = (4, 8, 8, 64)
    # a learnable map: the target is a smoothed copy of the input
    field = rng.standard_normal(shape).astype(np.float32)
    k = np.array([0.25, 0.5, 0.25])
    target = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), -1, field)
    f_d = torch.tensor(field, device="cuda")
    t_d = torch.tensor(target.astype(np.float32), device="cuda")
    losses = []
    for _ in range(60):
        lp = net.train_step(f_d, t_d, lr=3e-3)
        losses.append(float(lp.sum()) / field.size)
    assert losses[-1] < 0.3 * losses[0], losses[::10]
