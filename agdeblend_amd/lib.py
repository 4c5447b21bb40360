"""Loader and ctypes bindings for libagdeblend.so (the HIP build).

The library is built in-tree (agdeblend_amd/lib/libagdeblend.so) by
agdeblend_amd.build. Loading fails loudly when it is missing: there is no CPU
fallback for the product path.
"""
import ctypes
import os
import sys

import numpy as np

from ._marshal import C_INT, C_LONG, PtrTable, int_array, null

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "lib", "libagdeblend.so")

_lib = None


def load():
    """Load libagdeblend.so and declare its entry points."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            "libagdeblend.so is not built (%s); run agdeblend_amd.build.build()" % LIBPATH)
    # PyTorch-ROCm ships its own HIP runtime (libamdhip64.so) beside the one
    # this library links (libamdhip64.so.7): both work in one process only
    # when torch's initialises the device first
    torch = sys.modules.get("torch")
    if torch is not None:
        torch.cuda.is_available()
    L = ctypes.CDLL(LIBPATH)
    vp, ci, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_float

    def fn(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    geo = [ci] + [vp] * 11
    fn("agd_deblend", ci, *geo, cf, ci, ci, vp)
    fn("agd_blend", ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, ci, vp, vp, vp, vp, vp, vp)
    fn("agdeblend_deblend", ci, *geo, cf, ci, ci, vp)
    fn("agdeblend_blend", ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, ci, vp, vp, vp, vp, vp, vp)
    fn("agd_blend_noalloc", ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, ci, vp, vp, vp, vp, vp, vp)
    fn("agdx_create", vp, *geo)
    fn("agdx_create_host", vp, *geo)
    fn("agdx_destroy", None, vp)
    fn("agdx_model_len", ctypes.c_long, vp)
    fn("agdx_record_len", ctypes.c_long, vp)
    fn("agdx_n_windows", ci, vp)
    fn("agdx_window_info", ci, vp, ci, vp)
    fn("agdx_window_blend", ci, vp, ci, vp)
    fn("agdx_window_taper", ci, vp, ci, vp)
    fn("agdx_window_traces", ci, vp, ci, vp)
    fn("agdx_blend_tables", ci, ci, vp, vp, vp, vp, vp, ci, ci, vp, vp, vp, vp, vp, vp, vp,
       vp, vp)
    fn("agdx_intervals", ci, vp, vp)
    fn("agdx_mutes", ci, vp, vp)
    fn("agdx_unique_channels", ci, vp, vp)
    fn("agdx_stepsize", cf, vp)
    fn("agdx_set_data", ci, vp, vp, vp)
    fn("agdx_forward", ci, vp, vp, vp)
    fn("agdx_adjoint", ci, vp, vp, vp)
    fn("agdx_run", ci, vp, cf, ci, ci, vp)
    fn("agdx_run_warm", ci, vp, cf, ci, ci, vp)
    fn("agdx_warm_model", ci, vp, vp)
    fn("agdx_set_spatial_precision", ci, vp, ci)
    fn("agdx_autotune", ci, vp, ci, vp, vp)
    fn("agdx_n_spatial_classes", ci, vp)
    fn("agdx_n_time_classes", ci, vp)
    fn("agdx_time_factors", ci, vp, vp)
    fn("agdx_time_blocks", ci, vp, vp)
    fn("agdx_set_time_factors", ci, vp, ci, ci, ci)
    fn("agdx_time_variants", ci, vp, vp)
    fn("agdx_set_time_variant", ci, vp, ci, ci)
    fn("agdx_data_dev", vp, vp)
    fn("agdx_data_len", ctypes.c_int64, vp)
    fn("agdx_warm_model_dev", ci, vp, vp)
    fn("agdx_run_warm_dev", ci, vp, cf, ci, ci, vp)
    fn("agdx_get_output", ci, vp, vp)
    fn("agdx_finish", ci, vp)
    fn("agdx_get_model", ci, vp, vp)
    fn("agdx_trace_len", ctypes.c_long, vp)
    fn("agdx_operator", ci, vp, ci, ci, vp, vp)
    fn("agdx_iterate", ci, vp, ci, cf)
    fn("agdx_last_lambda0", cf)
    fn("agdx_last_warm_fallback", ci)
    fn("agdx_last_warm_ratio", ctypes.c_double)
    fn("agdx_timing", ci, ci)
    fn("agdx_timing_get", ci, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
       ctypes.POINTER(ctypes.c_long))
    fn("agdx_bytes", ctypes.c_double, vp, ci)
    i64 = ctypes.c_int64
    fn("agdu_conv3", ci, vp, vp, vp, vp, vp, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_conv3_cat", ci, vp, vp, vp, vp, vp, vp, ci, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_conv3_pool", ci, vp, vp, vp, vp, vp, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_up_weights", ci, vp, vp, ci, ci, ci, vp)
    fn("agdu_conv3_up", ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_conv_in", ci, vp, vp, ci, vp, vp, ci, ci, ci, ci, cf, vp)
    fn("agdu_conv_out", ci, vp, vp, ci, vp, vp, vp, ci, ci, ci, ci, cf, vp)
    fn("agdu_flip_weights", ci, vp, vp, ci, ci, vp)
    fn("agdu_wgrad_part_len", i64, ci, ci, ci, ci, ci, ci)
    fn("agdu_conv3_wgrad", ci, vp, vp, vp, vp, vp, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_maxpool", ci, vp, vp, vp, ci, ci, ci, ci, ci, vp)
    fn("agdu_maxpool_bwd", ci, vp, vp, vp, ci, ci, ci, ci, ci, vp)
    fn("agdu_upsample", ci, vp, vp, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_upsample_bwd", ci, vp, vp, ci, ci, ci, ci, ci, ci, ci, vp)
    fn("agdu_copy_channels", ci, vp, vp, i64, ci, ci, ci, ci, ci, ci, vp, vp)
    fn("agdu_pack_input", ci, vp, vp, i64, ci, cf, vp)
    fn("agdu_unpack_output", ci, vp, vp, vp, i64, ci, cf, vp)
    fn("agdu_l2_loss", ci, vp, vp, vp, vp, i64, ci, cf, vp)
    fn("agdu_loss_parts", i64, i64)
    fn("agdu_adam", ci, vp, vp, vp, vp, vp, ci, cf, cf, cf, cf, ci, vp)
    fn("agdu_to_bf16", ci, vp, vp, i64, vp)
    fn("agdu_from_bf16", ci, vp, vp, i64, vp)
    _lib = L
    return L


EXPORTS = ["agd_deblend", "agd_blend", "agdeblend_deblend", "agdeblend_blend", "agd_blend_noalloc",
           "agdx_create", "agdx_create_host", "agdx_destroy", "agdx_model_len",
           "agdx_record_len", "agdx_n_windows", "agdx_window_info", "agdx_window_blend",
           "agdx_window_taper", "agdx_window_traces", "agdx_blend_tables",
           "agdx_intervals", "agdx_mutes", "agdx_unique_channels", "agdx_stepsize",
           "agdx_set_data", "agdx_forward", "agdx_adjoint", "agdx_run", "agdx_run_warm",
           "agdx_warm_model", "agdx_set_spatial_precision", "agdx_autotune",
           "agdx_n_spatial_classes", "agdx_n_time_classes", "agdx_time_factors",
           "agdx_set_time_factors", "agdx_time_blocks", "agdx_time_variants",
           "agdx_set_time_variant", "agdx_data_dev", "agdx_data_len", "agdx_warm_model_dev",
           "agdx_run_warm_dev", "agdx_get_output",
           "agdx_finish",
           "agdx_get_model", "agdx_trace_len", "agdx_operator", "agdx_iterate", "agdx_last_lambda0", "agdx_last_warm_fallback", "agdx_last_warm_ratio",
           "agdx_timing",
           "agdx_timing_get", "agdx_bytes",
           "agdu_conv3", "agdu_conv3_pool", "agdu_conv3_cat", "agdu_up_weights", "agdu_conv3_up", "agdu_conv_in", "agdu_conv_out",
           "agdu_flip_weights", "agdu_wgrad_part_len", "agdu_conv3_wgrad",
           "agdu_maxpool", "agdu_maxpool_bwd", "agdu_upsample",
           "agdu_upsample_bwd", "agdu_copy_channels", "agdu_pack_input",
           "agdu_unpack_output", "agdu_l2_loss", "agdu_loss_parts", "agdu_adam",
           "agdu_to_bf16", "agdu_from_bf16"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Geometry:
    """Marshalled geometry arguments of agd_deblend (everything but data)."""

    def __init__(self, volumes, n_dims, window_shapes, coords, shapes, shottimes,
                 channels, trace_types, wavelet_lengths=None, wavelet_idxs=None,
                 wavelets=None):
        self.n_patches = len(volumes)
        self.volumes, self.volumes_p = int_array(volumes)
        self.n_dims, self.n_dims_p = int_array(n_dims)
        self.window_shapes = PtrTable(window_shapes, C_INT)
        self.coords = PtrTable([np.asarray(c).reshape(-1) if np.size(c) else np.zeros(1)
                                for c in coords], C_INT)
        self.shapes = PtrTable(shapes, C_INT)
        self.shottimes = PtrTable([np.asarray(s).reshape(-1) for s in shottimes], C_LONG)
        self.channels = PtrTable([np.asarray(c).reshape(-1) for c in channels], C_INT)
        self.trace_types = PtrTable([np.asarray(t).reshape(-1) for t in trace_types], C_INT)
        if wavelet_idxs is not None:
            self.wl, self.wl_p = int_array(wavelet_lengths)
            self.widx = PtrTable([np.asarray(w).reshape(-1) for w in wavelet_idxs], C_INT)
            self.wav = PtrTable(wavelets, np.float32)
            self.widx_p, self.wav_p = self.widx.ptr, self.wav.ptr
        else:
            self.wl_p, self.widx_p, self.wav_p = null(), null(), null()

    def args(self):
        return (self.n_patches, self.volumes_p, self.n_dims_p, self.window_shapes.ptr,
                self.coords.ptr, self.shapes.ptr, self.shottimes.ptr, self.channels.ptr,
                self.trace_types.ptr, self.wl_p, self.widx_p, self.wav_p)


class Problem:
    """A deblending problem resident on the GPU (agdx_* interface)."""

    def __init__(self, volumes, n_dims, window_shapes, coords, shapes, shottimes,
                 channels, trace_types, wavelet_lengths=None, wavelet_idxs=None,
                 wavelets=None, host_only=False):
        self.L = load()
        self.geo = Geometry(volumes, n_dims, window_shapes, coords, shapes, shottimes,
                            channels, trace_types, wavelet_lengths, wavelet_idxs, wavelets)
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        create = self.L.agdx_create_host if host_only else self.L.agdx_create
        self.h = create(*self.geo.args())
        if not self.h:
            raise RuntimeError("agdx_create failed")
        self.host_only = host_only
        self.model_len = self.L.agdx_model_len(self.h)
        self.record_len = self.L.agdx_record_len(self.h)
        self.n_windows = self.L.agdx_n_windows(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.agdx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def window(self, w):
        info = np.zeros(25, C_INT)
        if self.L.agdx_window_info(self.h, w, _p(info)):
            raise IndexError(w)
        nsp = int(info[1])
        return dict(patch=int(info[0]), nsp=nsp, ntr=int(info[2]), ntw=int(info[3]),
                    nfk=int(info[4]), nconv=int(info[5]), Pt=int(info[6]), Ft=int(info[7]),
                    Wt=int(info[8]), S=int(info[9]),
                    x_off=int(info[10]) + (int(info[11]) << 31),
                    P=[int(v) for v in info[12:12 + nsp]],
                    W=[int(v) for v in info[16:16 + nsp]],
                    coords=[int(v) for v in info[20:20 + nsp]], wM=int(info[24]))

    def window_taper(self, w):
        """(left, right) taper flags per space dim of window w."""
        nsp = self.window(w)["nsp"]
        out = np.zeros(max(1, 2 * nsp), C_INT)
        if self.L.agdx_window_taper(self.h, w, _p(out)):
            raise IndexError(w)
        return [int(v) for v in out[:2 * nsp]]

    def window_traces(self, w):
        """patch trace index of each trace of window w"""
        out = np.zeros(max(1, self.window(w)["ntr"]), C_INT)
        if self.L.agdx_window_traces(self.h, w, _p(out)):
            raise IndexError(w)
        return out[:self.window(w)["ntr"]]

    def blend_coords(self, w):
        n = self.window(w)["ntr"]
        out = np.zeros((n, 5), np.int64)
        self.L.agdx_window_blend(self.h, w, _p(out))
        return out

    def intervals(self):
        n = self.L.agdx_intervals(self.h, null())
        out = np.zeros((max(n, 1), 3), np.int64)
        self.L.agdx_intervals(self.h, _p(out))
        return out[:n]

    def mutes(self):
        n = self.L.agdx_mutes(self.h, null())
        out = np.zeros((max(n, 1), 4), np.int64)
        self.L.agdx_mutes(self.h, _p(out))
        return out[:n]

    def unique_channels(self):
        n = self.L.agdx_unique_channels(self.h, null())
        out = np.zeros(max(n, 1), C_INT)
        self.L.agdx_unique_channels(self.h, _p(out))
        return out[:n]

    def stepsize(self):
        return float(self.L.agdx_stepsize(self.h))

    def set_data(self, data, want_record=False):
        d = PtrTable([np.asarray(a, np.float32).reshape(-1) for a in data], np.float32)
        rec = np.zeros(max(1, self.record_len), np.float32) if want_record else None
        if self.L.agdx_set_data(self.h, d.ptr, _p(rec) if want_record else null()):
            raise RuntimeError("agdx_set_data failed")
        return rec[: self.record_len] if want_record else None

    def forward(self, x):
        x = np.ascontiguousarray(x, np.complex64)
        rec = np.zeros(max(1, self.record_len), np.float32)
        if self.L.agdx_forward(self.h, _p(x), _p(rec)):
            raise RuntimeError("agdx_forward failed")
        return rec[: self.record_len]

    def adjoint(self, rec):
        rec = np.ascontiguousarray(rec, np.float32)
        g = np.zeros(max(1, self.model_len), np.complex64)
        if self.L.agdx_adjoint(self.h, _p(rec), _p(g)):
            raise RuntimeError("agdx_adjoint failed")
        return g[: self.model_len]

    @property
    def trace_len(self):
        return int(self.L.agdx_trace_len(self.h))

    def operator(self, op, adjoint, v):
        """op 0 FK (model <-> traces), 1 wavelet (traces <-> traces),
        2 blend with mute (traces -> record; adjoint without the mute)."""
        n_in, n_out, cin, cout = {
            (0, 0): (self.model_len, self.trace_len, np.complex64, np.float32),
            (0, 1): (self.trace_len, self.model_len, np.float32, np.complex64),
            (1, 0): (self.trace_len, self.trace_len, np.float32, np.float32),
            (1, 1): (self.trace_len, self.trace_len, np.float32, np.float32),
            (2, 0): (self.trace_len, self.record_len, np.float32, np.float32),
            (2, 1): (self.record_len, self.trace_len, np.float32, np.float32)}[(op, int(adjoint))]
        v = np.ascontiguousarray(v, cin)
        assert v.size == n_in
        out = np.zeros(max(1, n_out), cout)
        if self.L.agdx_operator(self.h, int(op), int(adjoint), _p(v), _p(out)):
            raise RuntimeError("agdx_operator failed")
        return out[:n_out]

    def run(self, initial_factor, n_its, print_freq=0, x0=None):
        x0p = null()
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, np.complex64)
            x0p = _p(x0)
        if self.L.agdx_run(self.h, float(initial_factor), int(n_its), int(print_freq), x0p):
            raise RuntimeError("agdx_run failed")

    def warm_model(self, y):
        """Device model set to the warm-start model of data estimate y."""
        yt = PtrTable([np.asarray(a, np.float32).reshape(-1) for a in y], np.float32)
        if self.L.agdx_warm_model(self.h, yt.ptr):
            raise RuntimeError("agdx_warm_model failed")

    def run_warm(self, initial_factor, n_its, warm, print_freq=0):
        """Inversion started from the warm-start model of data estimate warm."""
        yt = PtrTable([np.asarray(a, np.float32).reshape(-1) for a in warm], np.float32)
        if self.L.agdx_run_warm(self.h, float(initial_factor), int(n_its), int(print_freq),
                                yt.ptr):
            raise RuntimeError("agdx_run_warm failed")

    def set_spatial_precision(self, mode):
        """0: FP32 register FFTs; 1: split-precision bf16 x 3 matrix-core DFTs
        (square spatial windows, P in {20, 40}). Returns the number of window
        classes the bf16 x 3 path serves."""
        n = int(self.L.agdx_set_spatial_precision(self.h, int(mode)))
        if n < 0:
            raise RuntimeError("agdx_set_spatial_precision failed")
        return n

    def autotune(self, reps=3):
        """Pick the time transforms' factorizations, then the spatial
        kernels' planes per task, by timing; returns the spatial choice per
        spatial class (0: generic class) and the times (ms). The time
        choice is read with time_factors()."""
        n = int(self.L.agdx_n_spatial_classes(self.h))
        nk = np.zeros(max(n, 1), C_INT)
        ms = np.zeros(max(n, 1), np.float64)
        if self.L.agdx_autotune(self.h, int(reps), _p(nk), _p(ms)):
            raise RuntimeError("agdx_autotune failed")
        return nk[:n].tolist(), ms[:n].tolist()

    def time_factors(self):
        """Current (M1, M2) factorization of each time class ((0, 0): the
        generic kernels)."""
        n = int(self.L.agdx_n_time_classes(self.h))
        out = np.zeros(2 * max(n, 1), C_INT)
        self.L.agdx_time_factors(self.h, _p(out))
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]

    def set_time_factors(self, c, m1, m2):
        if self.L.agdx_set_time_factors(self.h, int(c), int(m1), int(m2)):
            raise ValueError("no time kernel %d x %d for class %d" % (m1, m2, c))

    def time_blocks(self):
        """Time blocks as rows (window, first trace, tw0, tw1): the trace
        tile and the run of its time windows one block transforms."""
        n = int(self.L.agdx_time_blocks(self.h, None))
        out = np.zeros(4 * max(n, 1), C_INT)
        self.L.agdx_time_blocks(self.h, _p(out))
        return out[:4 * n].reshape(n, 4)

    def time_variants(self):
        """Forward time kernel variant of each time class (0 register
        prefetch, 1 lean tile)."""
        n = int(self.L.agdx_n_time_classes(self.h))
        out = np.zeros(max(n, 1), C_INT)
        self.L.agdx_time_variants(self.h, _p(out))
        return [int(v) for v in out[:n]]

    def set_time_variant(self, c, var):
        if self.L.agdx_set_time_variant(self.h, int(c), int(var)):
            raise ValueError("no variant %d for time class %d" % (var, c))

    def data_dev(self):
        """Device address of the patch data set by set_data (flat, patch
        order), and its length in floats."""
        return int(self.L.agdx_data_dev(self.h) or 0), int(self.L.agdx_data_len(self.h))

    def warm_model_dev(self, y_ptr):
        if self.L.agdx_warm_model_dev(self.h, ctypes.c_void_p(y_ptr)):
            raise RuntimeError("agdx_warm_model_dev failed")

    def run_warm_dev(self, initial_factor, n_its, y_ptr, print_freq=0):
        """run_warm with the estimate already on the device (flat layout of
        data_dev)."""
        if self.L.agdx_run_warm_dev(self.h, float(initial_factor), int(n_its), int(print_freq),
                                    ctypes.c_void_p(y_ptr)):
            raise RuntimeError("agdx_run_warm_dev failed")

    def iterate(self, n, lamb):
        if self.L.agdx_iterate(self.h, int(n), float(lamb)):
            raise RuntimeError("agdx_iterate failed")

    def finish(self):
        """Output assembled on the device (no host copy)."""
        if self.L.agdx_finish(self.h):
            raise RuntimeError("agdx_finish failed")

    def output(self):
        out = PtrTable([np.zeros(s, np.float32).reshape(-1) for s in self.shapes], np.float32,
                       writable=True)
        if self.L.agdx_get_output(self.h, out.ptr):
            raise RuntimeError("agdx_get_output failed")
        return [o.reshape(s) for o, s in zip(out.np, self.shapes)]

    def model(self):
        x = np.zeros(max(1, self.model_len), np.complex64)
        if self.L.agdx_get_model(self.h, _p(x)):
            raise RuntimeError("agdx_get_model failed")
        return x[: self.model_len]

    def bytes(self):
        return dict(M=self.L.agdx_bytes(self.h, 0), T=self.L.agdx_bytes(self.h, 1),
                    R=self.L.agdx_bytes(self.h, 2))


def timing(enable):
    load().agdx_timing(1 if enable else 0)


def timing_get(name):
    ms = ctypes.c_double(0)
    n = ctypes.c_long(0)
    load().agdx_timing_get(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def last_lambda0():
    return float(load().agdx_last_lambda0())


def last_warm_ratio():
    """||A x_warm - d|| / ||d|| measured by the last warm-started run's guard
    (None when it had no warm start)."""
    v = float(load().agdx_last_warm_ratio())
    return None if v < 0 else v


def last_warm_fallback():
    """True when the last warm-started run fell back to the zero start."""
    return bool(load().agdx_last_warm_fallback())


def blend_tables(n_times, shottimes, channels, trace_types, q_ntimes=0, q_shottimes=(),
                 q_channels=(), q_trace_types=()):
    """The blend tables agd_blend builds for per-patch traces (host only):
    dict of intervals [k][3] (channel, start, stop), mutes [k][4], unique
    channels, coords of each input trace [n][5] and of the query traces
    (q_ntimes samples each) [nq][5] (channel, interval, interval_start,
    trace_start, n)."""
    Lb = load()
    shot = [np.asarray(s, np.int64).reshape(-1) for s in shottimes]
    ntr = [len(s) for s in shot]
    ntr_a, ntr_p = int_array(ntr)
    nt_a, nt_p = int_array(n_times)
    s = PtrTable(shot, C_LONG)
    c = PtrTable([np.asarray(x).reshape(-1) for x in channels], C_INT)
    t = PtrTable([np.asarray(x).reshape(-1) for x in trace_types], C_INT)
    qs = np.require(np.asarray(q_shottimes, np.int64).reshape(-1), requirements=["C"])
    qc = np.require(np.asarray(q_channels, C_INT).reshape(-1), requirements=["C"])
    qt = np.require(np.asarray(q_trace_types, C_INT).reshape(-1), requirements=["C"])
    nq = len(qs)
    counts = np.zeros(3, C_INT)
    args = (len(shot), ntr_p, nt_p, s.ptr, c.ptr, t.ptr, nq, int(q_ntimes), _p(qs), _p(qc),
            _p(qt), _p(counts))
    if Lb.agdx_blend_tables(*args, null(), null(), null(), null(), null()):
        raise RuntimeError("agdx_blend_tables failed")
    iv = np.zeros((max(1, counts[0]), 3), np.int64)
    mu = np.zeros((max(1, counts[1]), 4), np.int64)
    ch = np.zeros(max(1, counts[2]), C_INT)
    co = np.zeros((max(1, sum(ntr)), 5), np.int64)
    qo = np.zeros((max(1, nq), 5), np.int64)
    if Lb.agdx_blend_tables(*args, _p(iv), _p(mu), _p(ch), _p(co), _p(qo)):
        raise RuntimeError("agdx_blend_tables failed")
    return dict(intervals=iv[:counts[0]], mutes=mu[:counts[1]], channels=ch[:counts[2]],
                coords=co[:sum(ntr)], q_coords=qo[:nq])
