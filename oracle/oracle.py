"""ctypes access to the CPU oracle (test infrastructure only).

Loads oracle/_build/liboracle32.so (float, the reference's default
AGD_TYPE) or liboracle64.so (double, the FP64 oracle). Only tests/, the smoke
check and bench.py's cpu_baseline leg use this module; the product library
never does.
"""
import ctypes
import os
import subprocess

import numpy as np

from agdeblend_amd._marshal import C_INT, C_LONG, PtrTable, int_array, null

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_libs = {}


def build():
    """Compile both oracle flavours (gcc, -O2 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


def lib(double=False):
    key = 64 if double else 32
    if key not in _libs:
        path = os.path.join(BUILD, "liboracle%d.so" % key)
        if not os.path.exists(path):
            build()
        _libs[key] = ctypes.CDLL(path)
        _declare(_libs[key], key)
    return _libs[key]


def _declare(L, key):
    pre = "oracle%d_" % key
    real = ctypes.c_double if key == 64 else ctypes.c_float
    vp = ctypes.c_void_p
    ci = ctypes.c_int

    def fn(name, res, *args):
        f = getattr(L, pre + name)
        f.restype = res
        f.argtypes = list(args)
        setattr(L, name, f)

    fn("deblend", ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, real, ci, ci, vp, vp)
    fn("deblend_warm", ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, real, ci, ci, vp, vp,
       vp)
    fn("model_from_data", None, vp, vp, vp, vp)
    fn("blend", ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, ci, vp, vp, vp, vp, vp, vp)
    fn("model_new", vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
    fn("model_free", None, vp)
    fn("model_n_windows", ci, vp)
    fn("model_len", ctypes.c_long, vp)
    fn("record_len", ctypes.c_long, vp)
    fn("n_channels", ci, vp)
    fn("model_window", None, vp, ci, vp)
    fn("model_window_scaler", real, vp, ci)
    fn("model_window_taper", None, vp, ci, ci, vp)
    fn("model_window_blend", None, vp, ci, vp)
    fn("model_intervals", ci, vp, vp)
    fn("model_mutes", ci, vp, vp)
    fn("model_unique_channels", None, vp, vp)
    fn("model_forward", None, vp, vp, vp)
    fn("model_adjoint_update", None, vp, vp, real, vp)
    fn("model_stepsize", real, vp)
    fn("model_shrink", None, vp, real, vp)
    fn("model_blend_input", None, vp, vp, vp, vp, vp, vp)
    fn("model_output", None, vp, vp, vp, vp)
    fn("window_fk_forward", None, vp, ci, vp, vp)
    fn("window_fk_adjoint", None, vp, ci, vp, vp)
    fn("window_wavelet", None, vp, ci, real, vp)
    fn("add_intervals", ci, ci, vp, vp, vp, vp)
    fn("taper_window", None, ci, ci, ci, vp)
    fn("lamb_schedule", None, real, ci, real, vp)
    fn("fft_c2c", None, ci, ci, vp)
    fn("fft_r2c", None, ci, vp, vp)
    fn("fft_c2r", None, ci, vp, vp)
    fn("fft_c2r_nd", None, ci, vp, vp, vp)
    fn("fft_r2c_nd", None, ci, vp, vp, vp)
    fn("wavelet_spectrum", None, vp, ci, ci, vp)
    fn("model_window_traces", None, vp, ci, vp, vp)
    fn("blendcfg_new", vp, ci, vp, vp, vp, vp, vp)
    fn("blendcfg_free", None, vp)
    fn("blendcfg_n_channels", ci, vp)
    fn("blendcfg_record_len", ctypes.c_long, vp)
    fn("blendcfg_unique", None, vp, vp)
    fn("blendcfg_intervals", ci, vp, vp)
    fn("blendcfg_mutes", ci, vp, vp)
    fn("blendcfg_coords", None, vp, ci, vp)
    fn("blendcfg_coords_for", None, vp, ci, ci, vp, vp, vp, vp)
    fn("blendcfg_forward", None, vp, ci, ci, vp, vp)
    fn("blendcfg_mute", None, vp, vp)
    fn("blendcfg_adjoint", None, vp, vp, ci, vp)
    fn("set_threads", None, ci)
    fn("get_threads", ci)


def _rdt(double):
    return np.float64 if double else np.float32


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# FFT pieces (pinned against numpy.fft)
# --------------------------------------------------------------------------
def fft_c2c(x, sign=-1, double=True):
    dt = _rdt(double)
    a = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(dt))
    lib(double).fft_c2c(len(x), sign, _p(a))
    return a[..., 0] + 1j * a[..., 1]


def fft_r2c(x, double=True):
    dt = _rdt(double)
    x = np.ascontiguousarray(x, dtype=dt)
    n = len(x)
    out = np.zeros((n // 2 + 1, 2), dt)
    lib(double).fft_r2c(n, _p(x), _p(out))
    return out[:, 0] + 1j * out[:, 1]


def fft_c2r(X, n, double=True):
    dt = _rdt(double)
    a = np.ascontiguousarray(np.stack([X.real, X.imag], -1).astype(dt))
    out = np.zeros(n, dt)
    lib(double).fft_c2r(n, _p(a), _p(out))
    return out


def fft_c2r_nd(X, shape, double=True):
    dt = _rdt(double)
    a = np.ascontiguousarray(np.stack([X.real, X.imag], -1).astype(dt))
    out = np.zeros(shape, dt)
    dims, dp = int_array(shape)
    lib(double).fft_c2r_nd(len(shape), dp, _p(a), _p(out))
    return out


def fft_r2c_nd(x, double=True):
    dt = _rdt(double)
    x = np.ascontiguousarray(x, dtype=dt)
    shape = x.shape
    out = np.zeros(shape[:-1] + (shape[-1] // 2 + 1, 2), dt)
    dims, dp = int_array(shape)
    lib(double).fft_r2c_nd(len(shape), dp, _p(x), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def taper_window(n, left, right, double=False):
    out = np.zeros(n, _rdt(double))
    lib(double).taper_window(n, int(left), int(right), _p(out))
    return out


def lamb_schedule(f, n_its, lamb0, double=False):
    """Thresholds of iterations 0..n_its-1 as the oracle's deblend applies them."""
    out = np.zeros(n_its, _rdt(double))
    lib(double).lamb_schedule(f, int(n_its), lamb0, _p(out))
    return out


def add_intervals(intervals, double=False):
    starts = np.array([a for a, _ in intervals], C_LONG)
    stops = np.array([b for _, b in intervals], C_LONG)
    os_ = np.zeros(len(intervals), C_LONG)
    oe = np.zeros(len(intervals), C_LONG)
    n = lib(double).add_intervals(len(intervals), _p(starts), _p(stops), _p(os_), _p(oe))
    return [(int(os_[i]), int(oe[i])) for i in range(n)]


def wavelet_spectrum(w, n, double=True):
    dt = _rdt(double)
    w = np.ascontiguousarray(w, dtype=dt)
    out = np.zeros((n // 2 + 1, 2), dt)
    lib(double).wavelet_spectrum(_p(w), len(w), n, _p(out))
    return out[:, 0] + 1j * out[:, 1]


# --------------------------------------------------------------------------
# Whole-API calls (same argument meaning as include/agdeblend.h)
# --------------------------------------------------------------------------
class _DeblendArgs:
    """Marshalled agd_deblend geometry arguments (data excluded)."""

    def __init__(self, volumes, n_dims, window_shapes, coords, shapes,
                 shottimes, channels, trace_types, wavelet_lengths=None,
                 wavelet_idxs=None, wavelets=None, double=False):
        self.n_patches = len(volumes)
        self.volumes, self.volumes_p = int_array(volumes)
        self.n_dims, self.n_dims_p = int_array(n_dims)
        self.window_shapes = PtrTable(window_shapes, C_INT)
        self.coords = PtrTable([np.asarray(c).reshape(-1) if len(np.asarray(c).reshape(-1)) else np.zeros(1) for c in coords], C_INT)
        self.shapes = PtrTable(shapes, C_INT)
        self.shottimes = PtrTable([np.asarray(s).reshape(-1) for s in shottimes], C_LONG)
        self.channels = PtrTable([np.asarray(c).reshape(-1) for c in channels], C_INT)
        self.trace_types = PtrTable([np.asarray(t).reshape(-1) for t in trace_types], C_INT)
        if wavelet_idxs is not None:
            self.wl, self.wl_p = int_array(wavelet_lengths)
            self.widx = PtrTable([np.asarray(w).reshape(-1) for w in wavelet_idxs], C_INT)
            self.wav = PtrTable(wavelets, _rdt(double))
            self.widx_p, self.wav_p = self.widx.ptr, self.wav.ptr
        else:
            self.wl_p, self.widx_p, self.wav_p = null(), null(), null()

    def head(self):
        return (self.n_patches, self.volumes_p, self.n_dims_p, self.window_shapes.ptr,
                self.coords.ptr, self.shapes.ptr, self.shottimes.ptr,
                self.channels.ptr, self.trace_types.ptr, self.wl_p, self.widx_p,
                self.wav_p)


def deblend(volumes, n_dims, window_shapes, coords, shapes, shottimes, channels,
            trace_types, data, initial_factor=1.0, n_its=100, print_freq=0,
            wavelet_lengths=None, wavelet_idxs=None, wavelets=None,
            double=False, threads=None, warm=None):
    """Run the oracle's agd_deblend. Returns (list of output arrays, stats)
    where stats = (initial threshold lambda0, stepsize) and, with warm,
    a third entry: 1 when the warm start fell back to the zero start. warm:
    optional data estimate (patch arrays) whose model_from_data replaces the
    start model when it fits the data better than the zero model."""
    dt = _rdt(double)
    L = lib(double)
    if threads is not None:
        L.set_threads(int(threads))
    args = _DeblendArgs(volumes, n_dims, window_shapes, coords, shapes, shottimes,
                        channels, trace_types, wavelet_lengths, wavelet_idxs,
                        wavelets, double)
    d = PtrTable([np.array(x, dtype=dt, copy=True) for x in data], dt, writable=True)
    stats = np.zeros(3 if warm is not None else 2, np.float64)
    if warm is not None:
        wt = PtrTable([np.asarray(x, dtype=dt).reshape(-1) for x in warm], dt)
        rc = L.deblend_warm(*args.head(), initial_factor, n_its, print_freq, d.ptr,
                            _p(stats), wt.ptr)
    else:
        rc = L.deblend(*args.head(), initial_factor, n_its, print_freq, d.ptr, _p(stats))
    if rc != 0:
        raise RuntimeError("oracle deblend failed")
    return [a.reshape(np.shape(x)) for a, x in zip(d.np, data)], stats


def blend(n_traces, n_times, shottimes, channels, trace_types, data, blend_mode,
          taper_length, n_traces_out, n_times_out, shottimes_out, channels_out,
          trace_types_out, double=False):
    dt = _rdt(double)
    L = lib(double)
    ntr, ntr_p = int_array(n_traces)
    nt, nt_p = int_array(n_times)
    ntro, ntro_p = int_array(n_traces_out)
    nto, nto_p = int_array(n_times_out)
    s = PtrTable([np.asarray(x).reshape(-1) for x in shottimes], C_LONG)
    c = PtrTable([np.asarray(x).reshape(-1) for x in channels], C_INT)
    t = PtrTable([np.asarray(x).reshape(-1) for x in trace_types], C_INT)
    d = PtrTable([np.asarray(x).reshape(-1) for x in data], dt)
    so = PtrTable([np.asarray(x).reshape(-1) for x in shottimes_out], C_LONG)
    co = PtrTable([np.asarray(x).reshape(-1) for x in channels_out], C_INT)
    to = PtrTable([np.asarray(x).reshape(-1) for x in trace_types_out], C_INT)
    out = PtrTable([np.zeros(int(a) * int(b), dt) for a, b in zip(n_traces_out, n_times_out)],
                   dt, writable=True)
    rc = L.blend(len(n_traces), ntr_p, nt_p, s.ptr, c.ptr, t.ptr, d.ptr, int(blend_mode),
                 int(taper_length), len(n_traces_out), ntro_p, nto_p, so.ptr, co.ptr,
                 to.ptr, out.ptr)
    if rc != 0:
        raise RuntimeError("oracle blend failed")
    return [o.reshape(int(a), int(b)) for o, a, b in zip(out.np, n_traces_out, n_times_out)]


class Model:
    """Oracle model handle: geometry, tables and single operators."""

    def __init__(self, volumes, n_dims, window_shapes, coords, shapes, shottimes,
                 channels, trace_types, wavelet_lengths=None, wavelet_idxs=None,
                 wavelets=None, double=False):
        self.double = double
        self.dt = _rdt(double)
        self.L = lib(double)
        self.args = _DeblendArgs(volumes, n_dims, window_shapes, coords, shapes,
                                 shottimes, channels, trace_types, wavelet_lengths,
                                 wavelet_idxs, wavelets, double)
        self.h = self.L.model_new(*self.args.head())
        if not self.h:
            raise RuntimeError("oracle model_new failed")
        self.n_windows = self.L.model_n_windows(self.h)
        self.model_len = self.L.model_len(self.h)
        self.record_len = self.L.record_len(self.h)
        self.windows = [self.window(w) for w in range(self.n_windows)]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.model_free(h)
            self.h = None

    def window(self, w):
        info = np.zeros(40, C_INT)
        self.L.model_window(self.h, w, _p(info))
        nd = int(info[1])
        return dict(patch=int(info[0]), ndim=nd, ntr=int(info[2]), ntw=int(info[3]),
                    nfk=int(info[4]), nconv=int(info[5]), Ft=int(info[6]),
                    npad=int(info[7]), coords=list(info[8:8 + nd - 1]),
                    shape=list(info[12:12 + nd - 1]),
                    use_taper=list(info[16:16 + 2 * (nd - 1)]),
                    W=list(info[24:24 + nd]), P=list(info[29:29 + nd]),
                    model_off=int(info[34]), npad_tr=int(info[35]),
                    scaler=float(self.L.model_window_scaler(self.h, w)))

    def window_traces(self, w, patch_shape):
        out = np.zeros(self.windows[w]["ntr"], C_INT)
        ps, pp = int_array(patch_shape)
        self.L.model_window_traces(self.h, w, pp, _p(out))
        return out

    def window_size(self, w):
        wi = self.windows[w]
        return wi["ntw"] * wi["npad_tr"] * wi["Ft"]

    def taper(self, w, d):
        out = np.zeros(self.windows[w]["W"][d], self.dt)
        self.L.model_window_taper(self.h, w, d, _p(out))
        return out

    def blend_coords(self, w):
        out = np.zeros((self.windows[w]["ntr"], 5), np.int64)
        self.L.model_window_blend(self.h, w, _p(out))
        return out

    def intervals(self):
        n = self.L.model_intervals(self.h, null())
        out = np.zeros((max(n, 1), 3), np.int64)
        self.L.model_intervals(self.h, _p(out))
        return out[:n]

    def mutes(self):
        n = self.L.model_mutes(self.h, null())
        out = np.zeros((max(n, 1), 4), np.int64)
        self.L.model_mutes(self.h, _p(out))
        return out[:n]

    def unique_channels(self):
        out = np.zeros(max(1, self.L.n_channels(self.h)), C_INT)
        self.L.model_unique_channels(self.h, _p(out))
        return out[: self.L.n_channels(self.h)]

    def zeros_model(self):
        return np.zeros(self.model_len, np.complex128 if self.double else np.complex64)

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128 if self.double else np.complex64)
        rec = np.zeros(self.record_len, self.dt)
        self.L.model_forward(self.h, _p(x), _p(rec))
        return rec

    def adjoint_update(self, rec, step, x):
        x = np.array(x, dtype=np.complex128 if self.double else np.complex64, copy=True)
        rec = np.ascontiguousarray(rec, dtype=self.dt)
        self.L.model_adjoint_update(self.h, _p(rec), step, _p(x))
        return x

    def stepsize(self):
        return float(self.L.model_stepsize(self.h))

    def shrink(self, lamb, x):
        x = np.array(x, dtype=np.complex128 if self.double else np.complex64, copy=True)
        self.L.model_shrink(self.h, lamb, _p(x))
        return x

    def blend_input(self, data, shottimes, channels, trace_types):
        d = PtrTable([np.asarray(a, self.dt).reshape(-1) for a in data], self.dt)
        s = PtrTable([np.asarray(a).reshape(-1) for a in shottimes], C_LONG)
        c = PtrTable([np.asarray(a).reshape(-1) for a in channels], C_INT)
        t = PtrTable([np.asarray(a).reshape(-1) for a in trace_types], C_INT)
        rec = np.zeros(self.record_len, self.dt)
        self.L.model_blend_input(self.h, d.ptr, s.ptr, c.ptr, t.ptr, _p(rec))
        return rec

    def output(self, x, shapes):
        x = np.ascontiguousarray(x, dtype=np.complex128 if self.double else np.complex64)
        sh = PtrTable(shapes, C_INT)
        out = PtrTable([np.zeros(int(np.prod(s)), self.dt) for s in shapes], self.dt,
                       writable=True)
        self.L.model_output(self.h, _p(x), sh.ptr, out.ptr)
        return [o.reshape(s) for o, s in zip(out.np, shapes)]

    def from_data(self, y, shapes):
        """Warm-start model s R2C(pad(y_w)) of patch data y."""
        yt = PtrTable([np.asarray(a, self.dt).reshape(-1) for a in y], self.dt)
        sh = PtrTable(shapes, C_INT)
        x = self.zeros_model()
        self.L.model_from_data(self.h, yt.ptr, sh.ptr, _p(x))
        return x

    def window_fk_forward(self, w, xw):
        wi = self.windows[w]
        xw = np.ascontiguousarray(xw, dtype=np.complex128 if self.double else np.complex64)
        out = np.zeros((wi["ntr"], wi["nconv"]), self.dt)
        self.L.window_fk_forward(self.h, w, _p(xw), _p(out))
        return out

    def window_fk_adjoint(self, w, tr):
        wi = self.windows[w]
        tr = np.ascontiguousarray(tr, dtype=self.dt)
        g = np.zeros(self.window_size(w), np.complex128 if self.double else np.complex64)
        self.L.window_fk_adjoint(self.h, w, _p(tr), _p(g))
        return g

    def window_wavelet(self, w, buf, direction):
        b = np.array(buf, dtype=self.dt, copy=True)
        self.L.window_wavelet(self.h, w, direction, _p(b))
        return b


class BlendConfig:
    """Oracle blend configuration built from patches of traces
    (blend.c:629-760)."""

    def __init__(self, n_traces, n_times, shottimes, channels, trace_types, double=False):
        self.double = double
        self.dt = _rdt(double)
        self.L = lib(double)
        self.n_traces = [int(n) for n in n_traces]
        self.n_times = [int(n) for n in n_times]
        self._ntr, ntr_p = int_array(n_traces)
        self._nt, nt_p = int_array(n_times)
        self._s = PtrTable([np.asarray(x).reshape(-1) for x in shottimes], C_LONG)
        self._c = PtrTable([np.asarray(x).reshape(-1) for x in channels], C_INT)
        self._t = PtrTable([np.asarray(x).reshape(-1) for x in trace_types], C_INT)
        self.h = self.L.blendcfg_new(len(n_traces), ntr_p, nt_p, self._s.ptr, self._c.ptr,
                                     self._t.ptr)
        self.n_channels = self.L.blendcfg_n_channels(self.h)
        self.record_len = self.L.blendcfg_record_len(self.h)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.blendcfg_free(h)
            self.h = None

    def unique_channels(self):
        out = np.zeros(max(1, self.n_channels), C_INT)
        self.L.blendcfg_unique(self.h, _p(out))
        return [int(v) for v in out[: self.n_channels]]

    def intervals(self):
        n = self.L.blendcfg_intervals(self.h, null())
        out = np.zeros((max(n, 1), 3), np.int64)
        self.L.blendcfg_intervals(self.h, _p(out))
        return out[:n]

    def mutes(self):
        n = self.L.blendcfg_mutes(self.h, null())
        out = np.zeros((max(n, 1), 4), np.int64)
        self.L.blendcfg_mutes(self.h, _p(out))
        return out[:n]

    def coords(self, p):
        out = np.zeros((max(1, self.n_traces[p]), 5), np.int64)
        self.L.blendcfg_coords(self.h, p, _p(out))
        return out[: self.n_traces[p]]

    def coords_for(self, n_times, shottimes, channels, trace_types):
        s = np.ascontiguousarray(shottimes, C_LONG)
        c = np.ascontiguousarray(channels, C_INT)
        t = np.ascontiguousarray(trace_types, C_INT)
        out = np.zeros((max(1, len(s)), 5), np.int64)
        self.L.blendcfg_coords_for(self.h, len(s), int(n_times), _p(s), _p(c), _p(t), _p(out))
        return out[: len(s)]

    def forward(self, data, mode=0, taper_length=0):
        d = PtrTable([np.asarray(a, self.dt).reshape(-1) for a in data], self.dt)
        rec = np.zeros(max(1, self.record_len), self.dt)
        self.L.blendcfg_forward(self.h, int(mode), int(taper_length), d.ptr, _p(rec))
        return rec[: self.record_len]

    def mute(self, rec):
        r = np.array(rec, dtype=self.dt, copy=True)
        if len(r):
            self.L.blendcfg_mute(self.h, _p(r))
        return r

    def adjoint(self, rec, p):
        r = np.ascontiguousarray(rec, dtype=self.dt)
        out = np.zeros(max(1, self.n_traces[p] * self.n_times[p]), self.dt)
        self.L.blendcfg_adjoint(self.h, _p(r), p, _p(out))
        return out[: self.n_traces[p] * self.n_times[p]].reshape(self.n_traces[p], self.n_times[p])


def blend_same_geometry(shottimes, channels, trace_types, data, blend_mode, taper_length=0,
                        trace_types_out=None, double=False):
    """agd_blend with the output geometry equal to the input, shaped like the
    Python wrapper's blend(shottimes, channels, trace_types, data, mode)."""
    shapes = [np.shape(d) for d in data]
    n_traces = [int(np.prod(s[:-1])) for s in shapes]
    n_times = [int(s[-1]) for s in shapes]
    if trace_types_out is None:
        trace_types_out = [np.where(np.asarray(t) == 1, 0, np.asarray(t)) for t in trace_types]
    out = blend(n_traces, n_times, shottimes, channels, trace_types,
                [np.asarray(d).reshape(-1) for d in data], blend_mode, taper_length,
                n_traces, n_times, shottimes, channels, trace_types_out, double=double)
    return [o.reshape(s) for o, s in zip(out, shapes)]
