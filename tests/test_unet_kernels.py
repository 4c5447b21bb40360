"""The warm-start network's kernels (agdu_*) against an FP64 numpy
reference on small tensors. BF16 storage: inputs are rounded to bf16 before
the reference sees them; outputs are compared to within bf16 rounding of
FP32-accumulated sums."""
import ctypes

import numpy as np
import pytest

from tests import unet_ref as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _lib():
    from agdeblend_amd import lib as L
    return L.load()


def _bf(a):
    return torch.tensor(np.asarray(a, np.float32), device="cuda").to(torch.bfloat16)


def _np(t):
    return t.float().cpu().numpy().astype(np.float64)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _close(got, ref, rtol=2e-2):
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(got - ref).max() / scale
    assert err <= rtol, err


def test_mfma_operand_layout_via_1x1_conv():
    """A convolution whose only non-zero tap is the centre is a plain
    matrix product; checks the lane layout of the matrix-core operands."""
    rng = np.random.default_rng(0)
    x = R.bf16_round(rng.standard_normal((1, 1, 1, 64, 16)))
    w = np.zeros((27, 16, 16))
    w[13] = R.bf16_round(rng.standard_normal((16, 16)))
    y = torch.empty((1, 1, 1, 64, 16), dtype=torch.bfloat16, device="cuda")
    x_d = _bf(x)
    w_d = _bf(w)
    assert _lib().agdu_conv3(_p(x_d), _p(w_d), None, None, _p(y), 1, 1, 1, 64, 16, 16, 0,
                             None) == 0
    torch.cuda.synchronize()
    ref = np.einsum("ndhwi,oi->ndhwo", x, w[13])
    _close(_np(y), ref, 1e-2)


@pytest.mark.parametrize("ci,co", [(16, 16), (16, 32), (32, 64), (96, 32), (48, 16)])
def test_conv3_forward(ci, co):
    rng = np.random.default_rng(ci * 7 + co)
    shape = (2, 4, 6, 64, ci)
    x = R.bf16_round(rng.standard_normal(shape))
    w = R.bf16_round(rng.standard_normal((27, co, ci)) / np.sqrt(27 * ci))
    b = rng.standard_normal(co).astype(np.float32)
    y = torch.empty(shape[:-1] + (co,), dtype=torch.bfloat16, device="cuda")
    bt = torch.tensor(b, device="cuda")
    L = _lib()
    x_d = _bf(x)
    w_d = _bf(w)
    assert L.agdu_conv3(_p(x_d), _p(w_d), _p(bt), None, _p(y), 2, 4, 6, 64, ci, co, 3,
                        None) == 0
    torch.cuda.synchronize()
    _close(_np(y), R.conv3(x, w, b.astype(np.float64), relu=True))


def test_conv3_streaming_16_channels_across_depth_blocks():
    """The 16 -> 16 convolution streams over the input planes in blocks of 8
    output planes: several depth blocks, a partial last one, partial row and
    column tiles, against the FP64 restatement; its pooling form against the
    pooling kernel over several depth blocks."""
    rng = np.random.default_rng(1616)
    L = _lib()
    shape = (1, 19, 5, 130, 16)
    x = R.bf16_round(rng.standard_normal(shape))
    w = R.bf16_round(rng.standard_normal((27, 16, 16)) / np.sqrt(27 * 16))
    b = rng.standard_normal(16).astype(np.float32)
    y = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
    bt = torch.tensor(b, device="cuda")
    assert L.agdu_conv3(_p(_bf(x)), _p(_bf(w)), _p(bt), None, _p(y), 1, 19, 5, 130, 16, 16, 3,
                        None) == 0
    torch.cuda.synchronize()
    _close(_np(y), R.conv3(x, w, b.astype(np.float64), relu=True))
    N, D, H, W = 1, 18, 6, 68
    xp = _bf(rng.standard_normal((N, D, H, W, 16)))
    y1 = torch.empty((N, D, H, W, 16), dtype=torch.bfloat16, device="cuda")
    y2 = torch.empty_like(y1)
    p1 = torch.empty((N, D // 2, H // 2, W // 2, 16), dtype=torch.bfloat16, device="cuda")
    p2 = torch.empty_like(p1)
    arg = torch.empty((N, D // 2, H // 2, W // 2, 16), dtype=torch.uint8, device="cuda")
    wt = _bf(w)
    assert L.agdu_conv3(_p(xp), _p(wt), _p(bt), None, _p(y1), N, D, H, W, 16, 16, 3, None) == 0
    assert L.agdu_maxpool(_p(y1), _p(p1), _p(arg), N, D, H, W, 16, None) == 0
    assert L.agdu_conv3_pool(_p(xp), _p(wt), _p(bt), _p(y2), _p(p2), N, D, H, W, 16, 16, 3,
                             None) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    assert torch.equal(p1.view(torch.int16), p2.view(torch.int16))
    ref = R.conv3(_np(xp).astype(np.float64), w, b.astype(np.float64), relu=True)
    _close(_np(y1), ref)


@pytest.mark.parametrize("ci,co", [(16, 32), (32, 64), (96, 32)])
def test_conv3_input_gradient_with_relu_mask(ci, co):
    rng = np.random.default_rng(ci + 3 * co)
    shape = (1, 4, 4, 64, ci)
    dy = R.bf16_round(rng.standard_normal(shape[:-1] + (co,)))
    w = R.bf16_round(rng.standard_normal((27, co, ci)) / np.sqrt(27 * co))
    gate = R.bf16_round(rng.standard_normal(shape))
    L = _lib()
    wtt = torch.empty((27, ci, co), dtype=torch.bfloat16, device="cuda")
    w_d = _bf(w)
    assert L.agdu_flip_weights(_p(w_d), _p(wtt), ci, co, None) == 0
    dx = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
    dy_d = _bf(dy)
    gate_d = _bf(gate)
    assert L.agdu_conv3(_p(dy_d), _p(wtt), None, _p(gate_d), _p(dx), *shape[:4], co, ci,
                        4, None) == 0
    torch.cuda.synchronize()
    ref = R.conv3_input_grad(dy, w) * (gate > 0)
    _close(_np(dx), ref)


@pytest.mark.parametrize("ci,co,shape", [
    (16, 16, (2, 4, 4, 64)), (48, 16, (1, 3, 5, 40)), (32, 32, (2, 2, 3, 130)),
    (16, 32, (1, 4, 4, 64)), (32, 64, (2, 3, 2, 20)), (64, 64, (1, 2, 4, 16)),
    (96, 32, (1, 4, 3, 70))])
def test_conv3_weight_and_bias_gradients(ci, co, shape):
    rng = np.random.default_rng(5 * ci + co)
    x = R.bf16_round(rng.standard_normal(shape + (ci,)))
    dy = R.bf16_round(rng.standard_normal(shape + (co,)))
    L = _lib()
    n = L.agdu_wgrad_part_len(*shape, ci, co)
    assert n > 0
    part = torch.empty(n, dtype=torch.float32, device="cuda")
    dw = torch.empty((27, co, ci), dtype=torch.float32, device="cuda")
    db = torch.empty(co, dtype=torch.float32, device="cuda")
    x_d = _bf(x)
    dy_d = _bf(dy)
    assert L.agdu_conv3_wgrad(_p(x_d), _p(dy_d), _p(part), _p(dw), _p(db), *shape, ci, co,
                              None) == 0
    torch.cuda.synchronize()
    _close(dw.cpu().numpy().astype(np.float64), R.conv3_weight_grad(x, dy), 1e-4)
    _close(db.cpu().numpy().astype(np.float64), dy.sum(axis=(0, 1, 2, 3)), 1e-4)
    # deterministic: a second call gives the same bits
    dw2 = torch.empty_like(dw)
    assert L.agdu_conv3_wgrad(_p(x_d), _p(dy_d), _p(part), _p(dw2), None, *shape, ci, co,
                              None) == 0
    torch.cuda.synchronize()
    assert torch.equal(dw, dw2)


def test_pool_upsample_and_their_gradients():
    rng = np.random.default_rng(11)
    shape = (2, 4, 6, 8, 16)
    # distinct values within every 2x2x2 cell (integers, exact in bf16), so
    # the maximum of a cell is unique
    base = rng.integers(-100, 100, (2, 2, 3, 4, 16, 1))
    cell = rng.permuted(np.broadcast_to(np.arange(8), (2, 2, 3, 4, 16, 8)), axis=-1) + base
    x = cell.reshape(2, 2, 3, 4, 16, 2, 2, 2).transpose(0, 1, 5, 2, 6, 3, 7, 4)
    x = x.reshape(shape).astype(np.float64)
    L = _lib()
    y = torch.empty((2, 2, 3, 4, 16), dtype=torch.bfloat16, device="cuda")
    arg = torch.empty((2, 2, 3, 4, 16), dtype=torch.uint8, device="cuda")
    x_d = _bf(x)
    assert L.agdu_maxpool(_p(x_d), _p(y), _p(arg), *shape, None) == 0
    dy = R.bf16_round(rng.standard_normal((2, 2, 3, 4, 16)))
    dx = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
    dy_d = _bf(dy)
    assert L.agdu_maxpool_bwd(_p(dy_d), _p(arg), _p(dx), *shape, None) == 0
    up = torch.empty((2, 8, 12, 16, 32), dtype=torch.bfloat16, device="cuda")
    up.zero_()
    x_d = _bf(x)
    assert L.agdu_upsample(_p(x_d), _p(up), *shape, 32, 16, None) == 0
    dup = R.bf16_round(rng.standard_normal((2, 8, 12, 16, 32)))
    dxu = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
    dup_d = _bf(dup)
    assert L.agdu_upsample_bwd(_p(dup_d), _p(dxu), *shape, 32, 16, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_np(y), R.maxpool(x))
    # gradient lands on the maximum of each 2x2x2 cell
    xr = x.reshape(2, 2, 2, 3, 2, 4, 2, 16)
    is_max = xr == R.maxpool(x)[:, :, None, :, None, :, None, :]
    ref = (is_max * dy[:, :, None, :, None, :, None, :]).reshape(shape)
    assert np.array_equal(_np(dx), ref)
    uref = R.upsample(x)
    assert np.array_equal(_np(up)[..., 16:], uref) and not _np(up)[..., :16].any()
    sref = dup[..., 16:].reshape(2, 4, 2, 6, 2, 8, 2, 16).sum(axis=(2, 4, 6))
    _close(_np(dxu), sref, 1e-2)


def test_loss_and_adam():
    rng = np.random.default_rng(12)
    nvox, C = 5000, 16
    y = R.bf16_round(rng.standard_normal((nvox, C)))
    t = rng.standard_normal(nvox).astype(np.float32)
    L = _lib()
    npart = L.agdu_loss_parts(nvox)
    part = torch.empty(npart, dtype=torch.float32, device="cuda")
    dy = torch.empty((nvox, C), dtype=torch.bfloat16, device="cuda")
    tt = torch.tensor(t, device="cuda")
    y_d = _bf(y)
    assert L.agdu_l2_loss(_p(y_d), _p(tt), _p(dy), _p(part), nvox, C, 0.5, None) == 0
    torch.cuda.synchronize()
    e = y[:, 0] - t.astype(np.float64) * 0.5
    assert abs(float(part.sum()) - float(np.sum(e * e))) <= 1e-4 * float(np.sum(e * e))
    _close(_np(dy)[:, 0], 2 * e / nvox, 1e-2)
    assert not _np(dy)[:, 1:].any()
    # Adam, two steps, against the closed form
    n = 1000
    w0 = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    w = torch.tensor(w0, device="cuda")
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    gt = torch.tensor(g, device="cuda")
    wb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    for step in (1, 2):
        assert L.agdu_adam(_p(w), _p(m), _p(v), _p(gt), _p(wb), n, 1e-2, 0.9, 0.999, 1e-8,
                           step, None) == 0
    torch.cuda.synchronize()
    wr, mr, vr = w0.astype(np.float64), 0.0, 0.0
    for step in (1, 2):
        mr = 0.9 * mr + 0.1 * g
        vr = 0.999 * vr + 0.001 * g.astype(np.float64) ** 2
        wr = wr - 1e-2 * (mr / (1 - 0.9 ** step)) / (np.sqrt(vr / (1 - 0.999 ** step)) + 1e-8)
    assert np.abs(w.cpu().numpy() - wr).max() <= 1e-5
    assert np.array_equal(wb.float().cpu().numpy(), R.bf16_round(w.cpu().numpy()).astype(np.float32))


@pytest.mark.parametrize("cu,cs,co", [(32, 16, 16), (64, 32, 32)])
def test_conv3_cat_equals_formed_concatenation(cu, cs, co):
    """The convolution reading [upsample(xu), x] in place gives the same bits
    as upsampling, concatenating and convolving."""
    rng = np.random.default_rng(cu + cs)
    N, D, H, W = 1, 4, 6, 70
    xu = _bf(rng.standard_normal((N, D // 2, H // 2, W // 2, cu)))
    xs = _bf(rng.standard_normal((N, D, H, W, cs)))
    w = _bf(rng.standard_normal((27, co, cu + cs)) / np.sqrt(27 * (cu + cs)))
    b = torch.tensor(rng.standard_normal(co).astype(np.float32), device="cuda")
    L = _lib()
    cat = torch.zeros((N, D, H, W, cu + cs), dtype=torch.bfloat16, device="cuda")
    assert L.agdu_upsample(_p(xu), _p(cat), N, D // 2, H // 2, W // 2, cu, cu + cs, 0, None) == 0
    assert L.agdu_copy_channels(_p(xs), _p(cat), N * D * H * W, cs, 0, cu + cs, cu, cs, 0, None,
                                None) == 0
    y1 = torch.empty((N, D, H, W, co), dtype=torch.bfloat16, device="cuda")
    y2 = torch.empty_like(y1)
    assert L.agdu_conv3(_p(cat), _p(w), _p(b), None, _p(y1), N, D, H, W, cu + cs, co, 3,
                        None) == 0
    assert L.agdu_conv3_cat(_p(xu), _p(xs), _p(w), _p(b), None, _p(y2), N, D, H, W, cu, cs, co,
                            3, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("ci,co", [(16, 16), (32, 32)])
def test_conv3_pool_equals_conv_then_maxpool(ci, co):
    """The convolution that also writes its 2x2x2 max pooling gives the
    convolution's bits and the pooling kernel's bits (partial tiles in w,
    blocks past the last plane included)."""
    rng = np.random.default_rng(ci + co)
    N, D, H, W = 1, 6, 8, 132
    x = _bf(rng.standard_normal((N, D, H, W, ci)))
    w = _bf(rng.standard_normal((27, co, ci)) / np.sqrt(27 * ci))
    b = torch.tensor(rng.standard_normal(co).astype(np.float32), device="cuda")
    L = _lib()
    y1 = torch.empty((N, D, H, W, co), dtype=torch.bfloat16, device="cuda")
    y2 = torch.empty_like(y1)
    p1 = torch.empty((N, D // 2, H // 2, W // 2, co), dtype=torch.bfloat16, device="cuda")
    p2 = torch.empty_like(p1)
    arg = torch.empty((N, D // 2, H // 2, W // 2, co), dtype=torch.uint8, device="cuda")
    assert L.agdu_conv3(_p(x), _p(w), _p(b), None, _p(y1), N, D, H, W, ci, co, 3, None) == 0
    assert L.agdu_maxpool(_p(y1), _p(p1), _p(arg), N, D, H, W, co, None) == 0
    assert L.agdu_conv3_pool(_p(x), _p(w), _p(b), _p(y2), _p(p2), N, D, H, W, ci, co, 3,
                             None) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    assert torch.equal(p1.view(torch.int16), p2.view(torch.int16))
    # odd sizes and channel pairs without a pooling kernel (64 output
    # channels: one plane per block) are refused; the buffers match the
    # channels asked for, so a wrong acceptance could not go out of bounds
    assert L.agdu_conv3_pool(_p(x), _p(w), _p(b), _p(y2), _p(p2), N, D, H, W - 1, ci, co, 3,
                             None) == 1
    x64 = torch.zeros((N, D, H, W, 32), dtype=torch.bfloat16, device="cuda")
    w64 = torch.zeros((27, 64, 32), dtype=torch.bfloat16, device="cuda")
    b64 = torch.zeros(64, dtype=torch.float32, device="cuda")
    y64 = torch.empty((N, D, H, W, 64), dtype=torch.bfloat16, device="cuda")
    p64 = torch.empty((N, D // 2, H // 2, W // 2, 64), dtype=torch.bfloat16, device="cuda")
    assert L.agdu_conv3_pool(_p(x64), _p(w64), _p(b64), _p(y64), _p(p64), N, D, H, W, 32, 64,
                             3, None) == 1


@pytest.mark.parametrize("cu,cs,co", [(32, 16, 16), (64, 32, 32)])
def test_conv3_up_combined_taps(cu, cs, co):
    """The combined taps are bit for bit the float32 sums of the bf16
    weights rounded once; the convolution on them matches the FP64
    reference of the same arithmetic, and the formed concatenation's
    convolution up to the rounding of the combined taps."""
    rng = np.random.default_rng(3 * cu + cs)
    N, D, H, W = 1, 6, 8, 70
    xu = _bf(rng.standard_normal((N, D // 2, H // 2, W // 2, cu)))
    xs = _bf(rng.standard_normal((N, D, H, W, cs)))
    w = _bf(rng.standard_normal((27, co, cu + cs)) / np.sqrt(27 * (cu + cs)))
    b = rng.standard_normal(co).astype(np.float32)
    b_d = torch.tensor(b, device="cuda")
    L = _lib()
    wcu = torch.empty((48, co, cu), dtype=torch.bfloat16, device="cuda")
    assert L.agdu_up_weights(_p(w), _p(wcu), cu + cs, co, cu, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_np(wcu), R.up_weights(_np(w), cu))
    y = torch.empty((N, D, H, W, co), dtype=torch.bfloat16, device="cuda")
    y2 = torch.empty_like(y)
    assert L.agdu_conv3_up(_p(xu), _p(xs), _p(w), _p(wcu), _p(b_d), None, _p(y), N, D, H, W,
                           cu, cs, co, 3, None) == 0
    assert L.agdu_conv3_cat(_p(xu), _p(xs), _p(w), _p(b_d), None, _p(y2), N, D, H, W, cu, cs,
                            co, 3, None) == 0
    torch.cuda.synchronize()
    ref = R.conv3_up(_np(xu), _np(xs), _np(w), _np(wcu), b.astype(np.float64), relu=True)
    _close(_np(y), ref, 1e-2)
    rel = np.linalg.norm(_np(y) - _np(y2)) / np.linalg.norm(_np(y2))
    assert rel <= 1e-2, rel
    # no kernel for other channel counts, odd sizes refused
    assert L.agdu_conv3_up(_p(xu), _p(xs), _p(w), _p(wcu), _p(b_d), None, _p(y), N, D, H, W - 1,
                           cu, cs, co, 3, None) == 1


def test_first_layer_and_head_kernels():
    rng = np.random.default_rng(21)
    N, D, H, W = 1, 3, 5, 72  # the head needs W % 4 == 0
    field = rng.standard_normal((N, D, H, W)).astype(np.float32)
    w = R.bf16_round(rng.standard_normal((27, 16, 16)) / 6.0)
    b = rng.standard_normal(16).astype(np.float32) * 0.1
    L = _lib()
    f_d = torch.tensor(field, device="cuda")
    w_d = _bf(w)
    b_d = torch.tensor(b, device="cuda")
    y = torch.empty((N, D, H, W, 16), dtype=torch.bfloat16, device="cuda")
    assert L.agdu_conv_in(_p(f_d), _p(w_d), 16, _p(b_d), _p(y), N, D, H, W, 1.5, None) == 0
    x = np.zeros((N, D, H, W, 16))
    x[..., 0] = R.bf16_round(field * 1.5)
    ref = R.conv3(x, w, b.astype(np.float64), relu=True)
    torch.cuda.synchronize()
    _close(_np(y), ref, 1e-2)
    # head: channel 0 of a 16 -> 16 convolution, fp32, plus base
    xin = R.bf16_round(rng.standard_normal((N, D, H, W, 16)))
    xin_d = _bf(xin)
    out = torch.empty((N, D, H, W), dtype=torch.float32, device="cuda")
    assert L.agdu_conv_out(_p(xin_d), _p(w_d), 16, _p(b_d), _p(f_d), _p(out), N, D, H, W, 0.25,
                           None) == 0
    torch.cuda.synchronize()
    ref = field + 0.25 * R.conv3(xin, w, b.astype(np.float64))[..., 0]
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("with_base", [True, False])
def test_head_across_depth_blocks_and_tiles(with_base):
    """The head over 37 planes (three depth blocks of 16, the last partial),
    three row tiles and three w tiles (the last partial), with and without
    the field it adds to: every output against the fp64 convolution."""
    rng = np.random.default_rng(23)
    N, D, H, W = 1, 37, 9, 136
    w = R.bf16_round(rng.standard_normal((27, 16, 16)) / 6.0)
    b = rng.standard_normal(16).astype(np.float32) * 0.1
    L = _lib()
    xin = R.bf16_round(rng.standard_normal((N, D, H, W, 16)))
    field = rng.standard_normal((N, D, H, W)).astype(np.float32)
    out = torch.full((N, D, H, W), np.nan, dtype=torch.float32, device="cuda")
    f_d = torch.tensor(field, device="cuda")
    x_d, w_d, b_d = _bf(xin), _bf(w), torch.tensor(b, device="cuda")
    assert L.agdu_conv_out(_p(x_d), _p(w_d), 16, _p(b_d), _p(f_d) if with_base else None,
                           _p(out), N, D, H, W, 0.5, None) == 0
    torch.cuda.synchronize()
    ref = 0.5 * R.conv3(xin, w, b.astype(np.float64))[..., 0]
    if with_base:
        ref = ref + field
    got = out.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
