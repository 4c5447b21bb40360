"""Learned warm start: a 3-level 3-D U-Net that maps pseudo-deblended data
(blended record cut back into traces at each shot time) to an estimate of
the unblended data, used as the initial model of the inversion through
agdx_run_warm_dev.

Every operation of the network runs in the agdu_* HIP kernels of
libagdeblend.so (BF16 activations and weights, FP32 accumulation on the
matrix cores, FP32 master weights and Adam moments); PyTorch only provides
device memory and the stream.

Network (channels-last tensors [n][d][h][w][c], 3x3x3 convolutions):
  e1a 16->16, e1b 16->16 | 2x2x2 max pool | e2a 16->32, e2b 32->32 | pool |
  b1 32->64, b2 64->64 | upsample, concat e2b -> 96 | d2a 96->32, d2b 32->32 |
  upsample, concat e1b -> 48 | d1a 48->16, d1b 16->16 | head 16->16 (linear)
ReLU after every layer but the head. The input field sits in channel 0 of
16; the head's channel 0 is the prediction of (unblended - input) * scale,
so the estimate is input + head[0] / scale.

The hook into the inversion is the reference's initial model
(agdeblend.c:584-590): instead of the thresholded first update, the model
starts from the FK coefficients of the estimate.
"""
import ctypes
import os

import numpy as np
import torch

from . import lib as L

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_c4.npz")

LAYERS = (("e1a", 16, 16), ("e1b", 16, 16), ("e2a", 16, 32), ("e2b", 32, 32),
          ("b1", 32, 64), ("b2", 64, 64), ("d2a", 96, 32), ("d2b", 32, 32),
          ("d1a", 48, 16), ("d1b", 16, 16), ("head", 16, 16))
CH = 16  # channels of the input and output tensors


def _check(rc, what):
    if rc:
        raise RuntimeError("%s failed" % what)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class UNet:
    """Weights, optimizer state and activation buffers on one HIP device."""

    # decoder layers whose upsampled channels run on the 48 combined taps in
    # the inference pass (agdu_conv3_up: 12 instead of 27 matrix-core steps
    # per 16 of those channels, combined weights rounded once to bf16); the
    # others read the concatenation in place with all 27 taps. d2a stays on
    # 27: its combined taps (48 x 32 x 64) leave room for one block per CU
    # only, and it measures slower.
    UP_LAYERS = ("d1a",)

    def __init__(self, seed=0, device="cuda", head_scale=0.1):
        if not torch.cuda.is_available():
            raise RuntimeError(
                "the warm-start network needs a HIP device visible to torch; in a process "
                "that also calls libagdeblend, import torch before the first library call "
                "(torch's HIP runtime must initialise the device first)")
        self.L = L.load()
        self.dev = torch.device(device)
        self.offs = {}
        n = 0
        for name, ci, co in LAYERS:
            self.offs[name] = (n, n + 27 * ci * co)
            n += 27 * ci * co + co
        self.n_params = n
        rng = np.random.default_rng(seed)
        w = np.zeros(n, np.float32)
        for name, ci, co in LAYERS:
            a, b = self.offs[name]
            std = np.sqrt(2.0 / (27 * ci)) * (head_scale if name == "head" else 1.0)
            w[a:b] = rng.standard_normal(b - a) * std
        self.scale = 1.0
        self._set_params(w)
        self.m = torch.zeros(n, device=self.dev)
        self.v = torch.zeros(n, device=self.dev)
        self.g = torch.zeros(n, device=self.dev)
        self.step_count = 0
        self._bufs = {}
        self.up_layers = set(self.UP_LAYERS)
        self._wcu = {}

    # -- parameters ---------------------------------------------------------
    def _set_params(self, w):
        self.w = torch.tensor(np.asarray(w, np.float32), device=self.dev)
        self.wb = torch.empty(self.n_params, dtype=torch.bfloat16, device=self.dev)
        self.wt = torch.empty(self.n_params, dtype=torch.bfloat16, device=self.dev)
        _check(self.L.agdu_to_bf16(_vp(self.w), _vp(self.wb), self.n_params, self._s()),
               "agdu_to_bf16")

    def _s(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def _wptr(self, name, buf=None):
        a, _ = self.offs[name]
        buf = self.wb if buf is None else buf
        return ctypes.c_void_p(buf.data_ptr() + 2 * a)

    def _bptr(self, name, buf=None):
        _, b = self.offs[name]
        buf = self.w if buf is None else buf
        return ctypes.c_void_p(buf.data_ptr() + 4 * b)

    def params(self):
        """{name: (w [27][co][ci], b [co])} as float32 numpy arrays."""
        w = self.w.cpu().numpy()
        out = {}
        for name, ci, co in LAYERS:
            a, b = self.offs[name]
            out[name] = (w[a:b].reshape(27, co, ci).copy(), w[b:b + co].copy())
        return out

    def grads(self):
        g = self.g.cpu().numpy()
        out = {}
        for name, ci, co in LAYERS:
            a, b = self.offs[name]
            out[name] = (g[a:b].reshape(27, co, ci).copy(), g[b:b + co].copy())
        return out

    def save(self, path):
        np.savez(path, w=self.w.cpu().numpy(), scale=np.float32(self.scale),
                 layers=np.array([("%s:%d:%d" % l) for l in LAYERS]))

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        layers = [str(s) for s in z["layers"]]
        if layers != ["%s:%d:%d" % l for l in LAYERS] or z["w"].size != self.n_params:
            raise ValueError("%s does not hold this network's weights" % path)
        self._set_params(z["w"])
        self.scale = float(z["scale"])
        return self

    # -- buffers --------------------------------------------------------------
    def _buffers(self, shape, train):
        key = (shape, train)
        if key in self._bufs:
            return self._bufs[key]
        N, D, H, W = shape
        if D % 4 or H % 4 or W % 4:
            raise ValueError("patch sizes must be multiples of 4, got %s" % (shape,))
        bf = torch.bfloat16

        def t(lv, c, dtype=bf):
            f = 2 ** lv
            return torch.empty((N, D // f, H // f, W // f, c), dtype=dtype, device=self.dev)

        B = dict(e1a=t(0, 16), e1b=t(0, 16), p1=t(1, 16), a1=t(1, 16, torch.uint8),
                 e2a=t(1, 32), e2b=t(1, 32), p2=t(2, 32), a2=t(2, 32, torch.uint8),
                 b1=t(2, 64), b2=t(2, 64), d2a=t(1, 32), d2b=t(1, 32), d1a=t(0, 16),
                 d1b=t(0, 16))
        if train:
            # the training pass keeps the packed input, the concatenations and
            # the head's 16 channels for the backward pass
            B.update(x=t(0, 16), c2=t(1, 96), c1=t(0, 48), head=t(0, 16))
            B.update(g16a=t(0, 16), g16b=t(0, 16), g48=t(0, 48), gh32a=t(1, 32),
                     gh32b=t(1, 32), gh96=t(1, 96), gh16=t(1, 16), gq64a=t(2, 64),
                     gq64b=t(2, 64), gq32=t(2, 32))
            nvox = N * D * H * W
            # weight-gradient scratch: the largest need over the layers
            nw = 0
            for lv, ci, co in ((0, 16, 16), (0, 48, 16), (1, 16, 32), (1, 32, 32), (1, 96, 32),
                               (2, 32, 64), (2, 64, 64)):
                f = 2 ** lv
                nw = max(nw, int(self.L.agdu_wgrad_part_len(N, D // f, H // f, W // f, ci, co)))
            B["wpart"] = torch.empty(nw, dtype=torch.float32, device=self.dev)
            B["lpart"] = torch.empty(int(self.L.agdu_loss_parts(nvox)), dtype=torch.float32,
                                     device=self.dev)
        self._bufs[key] = B
        return B

    def release(self):
        self._bufs.clear()

    # -- forward ----------------------------------------------------------------
    def _conv(self, name, x, y, dims, mode):
        ci, co = x.shape[-1], y.shape[-1]
        _check(self.L.agdu_conv3(_vp(x), self._wptr(name), self._bptr(name), None, _vp(y),
                                 *dims, ci, co, mode, self._s()), "agdu_conv3 " + name)

    def forward(self, field, B):
        """field: float32 device tensor [N, D, H, W] (unscaled input data)."""
        N, D, H, W = field.shape
        _check(self.L.agdu_pack_input(_vp(field), _vp(B["x"]), field.numel(), CH, self.scale,
                                      self._s()), "agdu_pack_input")
        return self._forward_packed(B, (N, D, H, W))

    def _forward_packed(self, B, shape):
        N, D, H, W = shape
        d0, d1, d2 = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4)
        s = self._s()
        nvox = N * D * H * W
        self._conv("e1a", B["x"], B["e1a"], d0, 3)
        self._conv("e1b", B["e1a"], B["e1b"], d0, 3)
        _check(self.L.agdu_maxpool(_vp(B["e1b"]), _vp(B["p1"]), _vp(B["a1"]), *d0, 16, s),
               "agdu_maxpool")
        self._conv("e2a", B["p1"], B["e2a"], d1, 3)
        self._conv("e2b", B["e2a"], B["e2b"], d1, 3)
        _check(self.L.agdu_maxpool(_vp(B["e2b"]), _vp(B["p2"]), _vp(B["a2"]), *d1, 32, s),
               "agdu_maxpool")
        self._conv("b1", B["p2"], B["b1"], d2, 3)
        self._conv("b2", B["b1"], B["b2"], d2, 3)
        _check(self.L.agdu_upsample(_vp(B["b2"]), _vp(B["c2"]), *d2, 64, 96, 0, s),
               "agdu_upsample")
        nv1 = nvox // 8
        _check(self.L.agdu_copy_channels(_vp(B["e2b"]), _vp(B["c2"]), nv1, 32, 0, 96, 64, 32, 0,
                                         None, s), "agdu_copy_channels")
        self._conv("d2a", B["c2"], B["d2a"], d1, 3)
        self._conv("d2b", B["d2a"], B["d2b"], d1, 3)
        _check(self.L.agdu_upsample(_vp(B["d2b"]), _vp(B["c1"]), *d1, 32, 48, 0, s),
               "agdu_upsample")
        _check(self.L.agdu_copy_channels(_vp(B["e1b"]), _vp(B["c1"]), nvox, 16, 0, 48, 32, 16,
                                         0, None, s), "agdu_copy_channels")
        self._conv("d1a", B["c1"], B["d1a"], d0, 3)
        self._conv("d1b", B["d1a"], B["d1b"], d0, 3)
        self._conv("head", B["d1b"], B["head"], d0, 1)
        return B["head"]

    def predict_ptr(self, src_ptr, shape, out):
        """Estimate of the unblended data from the pseudo-deblended field at
        device address src_ptr (float32 [N, D, H, W], e.g. the patch data of
        a deblending problem, agdx_data_dev): field + head[0] / scale, into
        the float32 device tensor out. The inference pass: the first layer
        reads the field itself, each level's last encoder layer writes its
        2x2x2 max pooling as well, the decoder convolutions read the
        upsampled and skip tensors in place of their concatenation, and the
        head computes only its channel 0, in fp32, added to the field."""
        N, D, H, W = shape
        if tuple(out.shape) != tuple(shape) or out.dtype != torch.float32:
            raise ValueError("out must be float32 of shape %s" % (shape,))
        B = self._buffers(tuple(shape), False)
        d0, d1, d2 = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4)
        s = self._s()
        L = self.L
        src = ctypes.c_void_p(src_ptr)
        _check(L.agdu_conv_in(src, self._wptr("e1a"), 16, self._bptr("e1a"), _vp(B["e1a"]), *d0,
                              self.scale, s), "agdu_conv_in")
        # the encoder's last layer of each level writes its pooling too
        self._conv_pool("e1b", B["e1a"], B["e1b"], B["p1"], d0, 16, 16)
        self._conv("e2a", B["p1"], B["e2a"], d1, 3)
        self._conv_pool("e2b", B["e2a"], B["e2b"], B["p2"], d1, 32, 32)
        self._conv("b1", B["p2"], B["b1"], d2, 3)
        self._conv("b2", B["b1"], B["b2"], d2, 3)
        self._conv_cat("d2a", B["b2"], B["e2b"], B["d2a"], d1, 64, 32, 32)
        self._conv("d2b", B["d2a"], B["d2b"], d1, 3)
        self._conv_cat("d1a", B["d2b"], B["e1b"], B["d1a"], d0, 32, 16, 16)
        self._conv("d1b", B["d1a"], B["d1b"], d0, 3)
        _check(L.agdu_conv_out(_vp(B["d1b"]), self._wptr("head"), 16, self._bptr("head"), src,
                               _vp(out), *d0, 1.0 / self.scale, s), "agdu_conv_out")
        return out

    def _conv_pool(self, name, x, y, pool, dims, ci, co):
        """Layer name (bias, ReLU) into y, and its 2x2x2 max pooling into
        pool (the inference pass: no pooling arg)."""
        _check(self.L.agdu_conv3_pool(_vp(x), self._wptr(name), self._bptr(name), _vp(y),
                                      _vp(pool), *dims, ci, co, 3, self._s()),
               "agdu_conv3_pool " + name)

    def _conv_cat(self, name, xu, xs, y, dims, cu, cs, co):
        """Decoder layer on [upsample(xu), xs] (bias, ReLU), with combined
        taps for the upsampled channels when name is in up_layers (built
        from the current bf16 weights at every call: two small launches)."""
        s = self._s()
        if name in self.up_layers:
            if name not in self._wcu:
                self._wcu[name] = torch.empty(48 * co * cu, dtype=torch.bfloat16,
                                              device=self.dev)
            wcu = _vp(self._wcu[name])
            _check(self.L.agdu_up_weights(self._wptr(name), wcu, cu + cs, co, cu, s),
                   "agdu_up_weights " + name)
            _check(self.L.agdu_conv3_up(_vp(xu), _vp(xs), self._wptr(name), wcu,
                                        self._bptr(name), None, _vp(y), *dims, cu, cs, co, 3,
                                        s), "agdu_conv3_up " + name)
        else:
            _check(self.L.agdu_conv3_cat(_vp(xu), _vp(xs), self._wptr(name), self._bptr(name),
                                         None, _vp(y), *dims, cu, cs, co, 3, s),
                   "agdu_conv3_cat " + name)

    def predict(self, field, out=None):
        """predict_ptr on a float32 device tensor [N, D, H, W]."""
        if out is None:
            out = torch.empty_like(field)
        return self.predict_ptr(field.data_ptr(), tuple(field.shape), out)

    def predict_unfused(self, field):
        """The same estimate through the training pass's kernels (packed
        input, formed concatenations, 16-channel head), for comparison."""
        B = self._buffers(tuple(field.shape), True)
        head = self.forward(field, B)
        out = torch.empty_like(field)
        _check(self.L.agdu_unpack_output(_vp(head), _vp(out), _vp(field), field.numel(), CH,
                                         1.0 / self.scale, self._s()), "agdu_unpack_output")
        return out

    # -- backward ---------------------------------------------------------------
    def _layer_bwd(self, name, x, dy, dims, dx=None, gate=None):
        """weight and bias gradient of layer name (input x, pre-activation
        gradient dy); the input gradient times (gate > 0) into dx"""
        ci, co = x.shape[-1], dy.shape[-1]
        a, b = self.offs[name]
        s = self._s()
        B = self._cur
        _check(self.L.agdu_conv3_wgrad(_vp(x), _vp(dy), _vp(B["wpart"]),
                                       ctypes.c_void_p(self.g.data_ptr() + 4 * a),
                                       ctypes.c_void_p(self.g.data_ptr() + 4 * b), *dims, ci,
                                       co, s), "agdu_conv3_wgrad " + name)
        if dx is not None:
            _check(self.L.agdu_conv3(_vp(dy), self._wptr(name, self.wt), None, _vp(gate),
                                     _vp(dx), *dims, co, ci, 4, s), "agdu_conv3 bwd " + name)

    def train_step(self, field, target, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on a batch: field and target float32 device tensors
        [N, D, H, W] (pseudo-deblended input, unblended truth). Returns a
        device tensor holding the batch's summed squared error (of the
        scaled residual); the mean loss is that over N*D*H*W."""
        N, D, H, W = field.shape
        d0, d1, d2 = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4)
        B = self._cur = self._buffers(tuple(field.shape), True)
        s = self._s()
        nvox = N * D * H * W
        head = self.forward(field, B)
        # residual target (target - field), scaled: the loss kernel takes
        # t * tscale, so pass (target - field) through a scratch field
        resid = B.get("resid")
        if resid is None:
            resid = B["resid"] = torch.empty_like(field)
        torch.sub(target, field, out=resid)
        dh = B["g16a"]
        _check(self.L.agdu_l2_loss(_vp(head), _vp(resid), _vp(dh), _vp(B["lpart"]), nvox, CH,
                                   self.scale, s), "agdu_l2_loss")
        # flipped weights of every layer for the input gradients
        for name, ci, co in LAYERS:
            _check(self.L.agdu_flip_weights(self._wptr(name), self._wptr(name, self.wt), ci, co,
                                            s), "agdu_flip_weights")
        g16a, g16b, g48 = B["g16a"], B["g16b"], B["g48"]
        self._layer_bwd("head", B["d1b"], dh, d0, g16b, B["d1b"])
        self._layer_bwd("d1b", B["d1a"], g16b, d0, g16a, B["d1a"])
        self._layer_bwd("d1a", B["c1"], g16a, d0, g48, B["c1"])
        _check(self.L.agdu_upsample_bwd(_vp(g48), _vp(B["gh32a"]), *d1, 32, 48, 0, s),
               "agdu_upsample_bwd")
        self._layer_bwd("d2b", B["d2a"], B["gh32a"], d1, B["gh32b"], B["d2a"])
        self._layer_bwd("d2a", B["c2"], B["gh32b"], d1, B["gh96"], B["c2"])
        _check(self.L.agdu_upsample_bwd(_vp(B["gh96"]), _vp(B["gq64a"]), *d2, 64, 96, 0, s),
               "agdu_upsample_bwd")
        self._layer_bwd("b2", B["b1"], B["gq64a"], d2, B["gq64b"], B["b1"])
        self._layer_bwd("b1", B["p2"], B["gq64b"], d2, B["gq32"], B["p2"])
        # e2b's output gradient: pooled path plus the skip connection
        _check(self.L.agdu_maxpool_bwd(_vp(B["gq32"]), _vp(B["a2"]), _vp(B["gh32a"]), *d1, 32, s),
               "agdu_maxpool_bwd")
        _check(self.L.agdu_copy_channels(_vp(B["gh96"]), _vp(B["gh32a"]), nvox // 8, 96, 64, 32,
                                         0, 32, 1, None, s), "agdu_copy_channels")
        self._layer_bwd("e2b", B["e2a"], B["gh32a"], d1, B["gh32b"], B["e2a"])
        self._layer_bwd("e2a", B["p1"], B["gh32b"], d1, B["gh16"], B["p1"])
        _check(self.L.agdu_maxpool_bwd(_vp(B["gh16"]), _vp(B["a1"]), _vp(g16b), *d0, 16, s),
               "agdu_maxpool_bwd")
        _check(self.L.agdu_copy_channels(_vp(g48), _vp(g16b), nvox, 48, 32, 16, 0, 16, 1, None,
                                         s), "agdu_copy_channels")
        self._layer_bwd("e1b", B["e1a"], g16b, d0, g16a, B["e1a"])
        self._layer_bwd("e1a", B["x"], g16a, d0)
        self.step_count += 1
        _check(self.L.agdu_adam(_vp(self.w), _vp(self.m), _vp(self.v), _vp(self.g), _vp(self.wb),
                                self.n_params, lr, b1, b2, eps, self.step_count, s), "agdu_adam")
        return B["lpart"]


def _level(name):
    return 2 if name in ("b1", "b2") else (1 if name in ("e2a", "e2b", "d2a", "d2b") else 0)


def flops_per_voxel():
    """Multiply-adds x 2 of one forward pass per full-resolution voxel, as
    the architecture counts them (27 taps x all input x all output
    channels of every layer)."""
    return sum(2.0 * 27 * ci * co / 8 ** _level(name) for name, ci, co in LAYERS)


# upsampled input channels of the decoder layers that read [upsample(x), skip]
UP_CHANNELS = {"d2a": 64, "d1a": 32}


def executed_flops_per_voxel():
    """Matrix-core multiply-adds x 2 per full-resolution voxel that the
    inference pass (UNet.predict_ptr) executes: the first layer with its 27
    taps as the K of one 32-long product per output channel (one live input
    channel); the head's channel 0 as three 16 x 16 tap-partial products per
    voxel (one per depth offset, 16 channels by 9 taps padded to 16); the
    layers in UNet.UP_LAYERS with 12 instead of 27 steps for their upsampled
    channels; every other layer 27 taps x all channels."""
    f = 0.0
    for name, ci, co in LAYERS:
        if name == "e1a":
            macs = 32 * co
        elif name == "head":
            macs = 3 * 16 * 16
        elif name in UNet.UP_LAYERS:
            cu = UP_CHANNELS[name]
            macs = (12 * cu + 27 * (ci - cu)) * co
        else:
            macs = 27 * ci * co
        f += 2.0 * macs / 8 ** _level(name)
    return f


def estimate(pm, net, out=None):
    """Network estimate of every patch of a deblending problem whose data is
    set (lib.Problem), read from and written to device memory: returns a
    float32 device tensor in the flat layout of pm.data_dev(). Patches must
    be 3-D with every size a multiple of 4."""
    ptr, n = pm.data_dev()
    if not ptr:
        raise RuntimeError("the problem has no device data")
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=net.dev)
    off = 0
    for shape in pm.shapes:
        shape = tuple(int(v) for v in shape)
        size = int(np.prod(shape))
        if len(shape) != 3:
            raise ValueError("the warm-start network takes 3-D patches, got %s" % (shape,))
        net.predict_ptr(ptr + 4 * off, (1,) + shape, out[off:off + size].view((1,) + shape))
        off += size
    return out


def accepts(shapes):
    """Whether the network can estimate patches of these shapes: 3-D, every
    size a multiple of 4 (two 2x2x2 poolings)."""
    return all(len(s) == 3 and all(int(v) % 4 == 0 for v in s) for s in shapes)


def deblend_warm(pm, net, initial_factor, n_its, out=None):
    """The inversion of pm started from the network's estimate: estimate on
    the device, then agdx_run_warm_dev (the reference's loop from that
    initial model). The plain zero-start inversion is the fallback: it runs
    instead when net is None or cannot take pm's patches (then None is
    returned), and the library runs it when the estimate's model fits the
    data no better than the zero model (lib.last_warm_fallback() tells).
    Returns the estimate tensor."""
    if net is None or not accepts(pm.shapes):
        pm.run(initial_factor, n_its)
        return None
    est = estimate(pm, net, out)
    torch.cuda.synchronize(net.dev)
    pm.run_warm_dev(initial_factor, n_its, est.data_ptr())
    return est
