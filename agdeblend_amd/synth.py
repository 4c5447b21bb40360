"""Seeded synthetic blended-acquisition volumes for the five workloads.

Each generator returns a dict with
  geometry  : agd_deblend geometry arguments (volumes, n_dims, window_shapes,
              coords, shapes, shottimes, channels, trace_types and, for the
              vibroseis workload, wavelet_lengths / wavelet_idxs / wavelets)
  truth     : list of unblended patches (float32, [space..., time])
  n_its     : iteration count of the workload
  name      : workload name
``blend_input(prob, blend_fn)`` produces the pseudo-deblended input
(agd_blend Sum with the output geometry equal to the input), and ``snr``
the deblending quality in dB over live traces.

Randomness comes from numpy's PCG64 generator seeded per workload, so the
same volumes are produced on every machine with this numpy.
"""
import numpy as np

LIVE, BAD, MISSING = 0, 1, 2


def ricker(t, f):
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _add_events(data, taus, amps, dt, f, half_width=None):
    """data[..., t] += sum_e amps[e] * ricker(t*dt - taus[e][...], f),
    evaluated only within the wavelet's support."""
    nt = data.shape[-1]
    flat = data.reshape(-1, nt)
    if half_width is None:
        half_width = int(np.ceil(1.6 / (f * dt))) + 1
    offs = np.arange(-half_width, half_width + 1)
    for tau, amp in zip(taus, amps):
        tau = np.asarray(tau, np.float64).reshape(-1)
        c = np.round(tau / dt).astype(np.int64)
        idx = c[:, None] + offs[None, :]
        ok = (idx >= 0) & (idx < nt)
        val = amp * ricker(idx * dt - tau[:, None], f)
        rows = np.broadcast_to(np.arange(flat.shape[0])[:, None], idx.shape)
        np.add.at(flat, (rows[ok], idx[ok]), val[ok].astype(flat.dtype))
    return data


def _shot_times(rng, n, interval_s, dither_s, dt):
    t = np.arange(n) * interval_s + rng.uniform(-dither_s, dither_s, n)
    t -= t.min()
    return np.round(t / dt).astype(np.int64)


def config0(seed=1000, n_src=256, n_t=1024):
    """2-D (SourceX, Time): 256 sources 25 m apart, one channel, 1024
    samples at 4 ms; 6 linear and 4 hyperbolic events, 20 Hz Ricker; shots
    every 2 s with +-0.5 s dither. Windows (32, 256), 100 iterations."""
    rng = np.random.default_rng(seed)
    dt, dx, f = 0.004, 25.0, 20.0
    x = np.arange(n_src) * dx
    data = np.zeros((n_src, n_t), np.float64)
    tmax = (n_t - 1) * dt
    taus, amps = [], []
    for _ in range(6):
        t0 = rng.uniform(0.1, 0.9) * tmax
        v = rng.uniform(1500.0, 4000.0) * rng.choice([-1.0, 1.0])
        x0 = rng.uniform(0, x[-1])
        taus.append(t0 + (x - x0) / v)
        amps.append(rng.uniform(0.2, 1.0))
    for _ in range(4):
        t0 = rng.uniform(0.1, 0.8) * tmax
        v = rng.uniform(1500.0, 4000.0)
        x0 = rng.uniform(0, x[-1])
        taus.append(np.sqrt(t0 ** 2 + ((x - x0) / v) ** 2))
        amps.append(rng.uniform(0.2, 1.0))
    _add_events(data, taus, amps, dt, f)
    shot = _shot_times(rng, n_src, 2.0, 0.5, dt)
    geo = dict(volumes=[0], n_dims=[2], window_shapes=[[32, 256]], coords=[[0]],
               shapes=[[n_src, n_t]], shottimes=[shot], channels=[np.zeros(n_src, np.intc)],
               trace_types=[np.zeros(n_src, np.intc)])
    return dict(name="C0 2D (256 src x 1024 t), windows (32,256)", geometry=geo,
                truth=[data.astype(np.float32)], n_its=100)


def config0_small(seed=1001):
    """A smaller 2-D case of the same kind for fast tests: 64 sources x 512
    samples, windows (16, 128)."""
    p = config0(seed, n_src=64, n_t=512)
    p["geometry"]["window_shapes"] = [[16, 128]]
    p["name"] = "C0-mini 2D (64 src x 512 t), windows (16,128)"
    p["n_its"] = 30
    return p


def _events3d(rng, n_ch, n_src, tmax, dx=25.0, n_lin=6, n_hyp=4):
    """Draw the events of a (ChannelX, SourceX) grid of n_ch x n_src points
    dx apart: n_lin planes and n_hyp hyperbolas. Returns (tau(xr, xs), amp)
    pairs, tau evaluated at receiver / source coordinates in metres."""
    xs_mean = (n_src - 1) * dx / 2.0
    xr_mean = (n_ch - 1) * dx / 2.0
    xs_max = (n_src - 1) * dx
    ev = []
    for _ in range(n_lin):
        t0 = rng.uniform(0.1, 0.9) * tmax
        ps = rng.uniform(-1.0, 1.0) / rng.uniform(1500.0, 4000.0)
        pr = rng.uniform(-1.0, 1.0) / rng.uniform(1500.0, 4000.0)
        ev.append((lambda xr, xs, t0=t0, ps=ps, pr=pr:
                   t0 + ps * (xs - xs_mean) + pr * (xr - xr_mean), rng.uniform(0.2, 1.0)))
    for _ in range(n_hyp):
        t0 = rng.uniform(0.1, 0.8) * tmax
        v = rng.uniform(1500.0, 4000.0)
        shift = rng.uniform(-0.5, 0.5) * xs_max
        ev.append((lambda xr, xs, t0=t0, v=v, shift=shift:
                   np.sqrt(t0 ** 2 + ((xs - xr - shift) / v) ** 2), rng.uniform(0.2, 1.0)))
    return ev


def _eval_events(ev, xr, xs, n_t, dt, f):
    data = np.zeros((xr.shape[0], xs.shape[1], n_t), np.float32)
    shape = (xr.shape[0], xs.shape[1])
    _add_events(data, [np.broadcast_to(tau(xr, xs), shape) for tau, _ in ev],
                [a for _, a in ev], dt, f)
    return data


def _volume3d(rng, n_ch, n_src, n_t, dt, f, dx=25.0, n_lin=6, n_hyp=4):
    ev = _events3d(rng, n_ch, n_src, (n_t - 1) * dt, dx, n_lin, n_hyp)
    xr = np.arange(n_ch)[:, None] * dx
    xs = np.arange(n_src)[None, :] * dx
    return _eval_events(ev, xr, xs, n_t, dt, f)


def training_truth(seed, n_ch=32, n_src=64, n_t=2048, full=(64, 1024)):
    """Unblended half of a training pair of the learned warm start: an event
    set drawn as for the full (ChannelX, SourceX) grid of the 64 x 1024 x
    2048 workload (25 Hz Ricker, 4 ms), evaluated on a random n_ch x n_src
    sub-block, and the shot times of its sources (every 2 s with +-0.5 s
    dither). Returns (truth float32 [n_ch][n_src][n_t], shot times int64
    [n_ch][n_src])."""
    rng = np.random.default_rng(seed)
    dt, f, dx = 0.004, 25.0, 25.0
    ev = _events3d(rng, full[0], full[1], (n_t - 1) * dt, dx)
    c0 = int(rng.integers(0, full[0] - n_ch + 1))
    s0 = int(rng.integers(0, full[1] - n_src + 1))
    xr = (c0 + np.arange(n_ch))[:, None] * dx
    xs = (s0 + np.arange(n_src))[None, :] * dx
    truth = _eval_events(ev, xr, xs, n_t, dt, f)
    shot1 = _shot_times(rng, n_src, 2.0, 0.5, dt)
    shot = np.broadcast_to(shot1[None, :], (n_ch, n_src)).astype(np.int64)
    return truth, shot


def training_blend(truth, shot, blend_fn):
    """Pseudo-deblended input of a training pair: the truth blended into one
    continuous record per channel and cut back into traces at the shot
    times by blend_fn (agd_blend Sum)."""
    n_ch, n_src = truth.shape[:2]
    channels = np.broadcast_to(np.arange(n_ch, dtype=np.intc)[:, None], (n_ch, n_src)).copy()
    tt = np.full((n_ch, n_src), LIVE, np.intc)
    out = blend_fn([shot], [channels], [tt], [truth], 0, trace_types_out=[tt])
    return np.asarray(out[0], np.float32).reshape(truth.shape)


def config1(seed=1001, n_ch=64, n_src=512, n_t=2048, windows=(16, 32, 256), n_its=200,
            p_bad=0.05, p_missing=0.03, interval_s=2.0, dither_s=0.5):
    """3-D (ChannelX, SourceX, Time) 64 x 512 x 2048 at 4 ms, 25 Hz Ricker
    dipping and hyperbolic events; channel id = ChannelX index; shots every
    2 s with +-0.5 s dither (the same shot time on every channel); 5% of
    traces bad and 3% missing at random. Windows (16, 32, 256), 200
    iterations."""
    rng = np.random.default_rng(seed)
    dt, f = 0.004, 25.0
    data = _volume3d(rng, n_ch, n_src, n_t, dt, f)
    shot1 = _shot_times(rng, n_src, interval_s, dither_s, dt)
    shot = np.broadcast_to(shot1[None, :], (n_ch, n_src)).astype(np.int64)
    channels = np.broadcast_to(np.arange(n_ch, dtype=np.intc)[:, None], (n_ch, n_src)).copy()
    u = rng.random((n_ch, n_src))
    tt = np.full((n_ch, n_src), LIVE, np.intc)
    tt[u < p_bad] = BAD
    tt[(u >= p_bad) & (u < p_bad + p_missing)] = MISSING
    geo = dict(volumes=[0], n_dims=[3], window_shapes=[list(windows)], coords=[[0, 0]],
               shapes=[[n_ch, n_src, n_t]], shottimes=[shot], channels=[channels],
               trace_types=[tt])
    return dict(name="C1 3D (64 ch x 512 src x 2048 t), windows %s" % (tuple(windows),),
                geometry=geo, truth=[data], n_its=n_its)


def config1_small(seed=1011):
    """3-D case of the same kind sized for CPU reference runs: 8 x 64 x 512,
    windows (4, 16, 128)."""
    p = config1(seed, n_ch=8, n_src=64, n_t=512, windows=(4, 16, 128), n_its=20)
    p["name"] = "C1-mini 3D (8 ch x 64 src x 512 t), windows (4,16,128)"
    return p


def config2(seed=1001, **kw):
    """Same volume as config 1 (kw: its shot timing) with windows
    (32, 32, 512)."""
    p = config1(seed, windows=(32, 32, 512), n_its=200, **kw)
    p["name"] = "C2 3D (64 ch x 512 src x 2048 t), windows (32,32,512)"
    return p


def sweep(n, dt, f0, f1, taper=0.05):
    t = np.arange(n) * dt
    T = n * dt
    phase = 2 * np.pi * (f0 * t + 0.5 * (f1 - f0) / T * t ** 2)
    w = np.sin(phase)
    m = int(taper * n)
    if m > 0:
        ramp = 0.5 * (1 - np.cos(np.pi * np.arange(m) / m))
        w[:m] *= ramp
        w[-m:] *= ramp[::-1]
    return w


def config3(seed=1003, n_y=8, n_x=128, n_t=4096, n_w=4000, windows=(8, 32, 512), n_its=150):
    """Vibroseis, two source fleets in disjoint volumes (SourceY 8, SourceX
    128, Time 4096) at 2 ms; each fleet has its own seeded linear 8-80 Hz
    sweep of 4000 samples; the recorded traces are reflectivity convolved
    with the sweep. Shots every 3 s with +-1 s dither; one 10 s recording
    gap per record. Windows (8, 32, 512), 150 iterations. One shared
    channel."""
    rng = np.random.default_rng(seed)
    dt = 0.002
    n_pre = n_t - n_w + 1
    patches, shots, wavs, widx = [], [], [], []
    n_src = n_y * n_x
    for fleet in range(2):
        w = sweep(n_w, dt, 8.0 + rng.uniform(-1, 1), 80.0 + rng.uniform(-5, 5))
        wavs.append(w.astype(np.float32))
        refl = np.zeros((n_y, n_x, n_pre), np.float64)
        xs = np.arange(n_x)[None, :] * 25.0
        ys = np.arange(n_y)[:, None] * 25.0
        taus, amps = [], []
        for _ in range(4):
            t0 = rng.uniform(0.1, 0.9) * (n_pre - 1) * dt
            p = rng.uniform(-1, 1) * 2e-5
            q = rng.uniform(-1, 1) * 2e-5
            taus.append(t0 + p * (xs - xs.mean()) + q * (ys - ys.mean()))
            amps.append(rng.uniform(0.2, 1.0))
        _add_events(refl, [np.broadcast_to(t, (n_y, n_x)) for t in taus], amps, dt, 60.0,
                    half_width=6)
        data = np.zeros((n_y, n_x, n_t), np.float64)
        flat_r = refl.reshape(-1, n_pre)
        nf = 1 << int(np.ceil(np.log2(n_t + n_w)))
        W = np.fft.rfft(w, nf)
        conv = np.fft.irfft(np.fft.rfft(flat_r, nf, axis=1) * W[None, :], nf, axis=1)
        data.reshape(-1, n_t)[:] = conv[:, :n_t]
        patches.append(data.astype(np.float32))
        t = (np.arange(n_src) * 2 + fleet) * 1.5 + rng.uniform(-1.0, 1.0, n_src)
        t[n_src // 2:] += 10.0 + n_t * dt  # the recording gap
        shots.append(np.round(t / dt).astype(np.int64).reshape(n_y, n_x))
        widx.append(np.full((n_y, n_x), fleet, np.intc))
    base = min(s.min() for s in shots)
    shots = [s - base for s in shots]
    geo = dict(volumes=[0, 1], n_dims=[3, 3], window_shapes=[list(windows)] * 2,
               coords=[[0, 0], [0, 0]], shapes=[[n_y, n_x, n_t]] * 2, shottimes=shots,
               channels=[np.zeros((n_y, n_x), np.intc)] * 2,
               trace_types=[np.zeros((n_y, n_x), np.intc)] * 2,
               wavelet_lengths=[n_w, n_w], wavelet_idxs=widx, wavelets=wavs)
    return dict(name="C3 vibroseis 2 x (8 y x 128 x x 4096 t), sweep 4000, windows %s"
                % (tuple(windows),), geometry=geo, truth=patches, n_its=n_its)


def config3_small(seed=1013):
    """Vibroseis case sized for CPU reference runs: 2 x (2 x 32 x 600), sweep of
    400 samples, windows (2, 16, 64)."""
    p = config3(seed, n_y=2, n_x=32, n_t=600, n_w=400, windows=(2, 16, 64), n_its=20)
    p["name"] = "C3-mini vibroseis 2 x (2 x 32 x 600), sweep 400, windows (2,16,64)"
    return p


def config4(seed=40000, n_ch=64, n_src=1024, n_t=2048):
    """(ChannelX 64, SourceX 1024, Time 2048) volume of config-0-style
    events; deblend windows (32, 32, 256)."""
    p = config1(seed, n_ch=n_ch, n_src=n_src, n_t=n_t, windows=(32, 32, 256), n_its=200,
                p_bad=0.0, p_missing=0.0)
    p["name"] = "C4 3D (64 ch x 1024 src x 2048 t), windows (32,32,256)"
    return p


WORKLOADS = {0: config0, 1: config1, 2: config2, 3: config3, 4: config4}
SMALL = {0: config0_small, 1: config1_small, 3: config3_small}


def make_config(i, small=False, blend_fn=None):
    """Build workload i; the pseudo-deblended input is added as 'blended'
    using blend_fn (default: the oracle-free product blend is not assumed;
    pass agdeblend_amd.agdeblend.blend or the oracle's blend)."""
    prob = (SMALL[i] if small else WORKLOADS[i])()
    if blend_fn is not None:
        prob["blended"] = blend_input(prob, blend_fn)
    else:
        prob["blended"] = None
    return prob


def blend_input(prob, blend_fn):
    """Pseudo-deblended input: blend (Sum) of the truth, extracted back at the
    input geometry. Output traces may not be Bad (agdeblend.c:712), so bad
    traces are extracted as live ones (their samples are muted anyway)."""
    g = prob["geometry"]
    tt_out = [np.where(np.asarray(t) == BAD, LIVE, np.asarray(t)).astype(np.intc)
              for t in g["trace_types"]]
    return blend_fn(g["shottimes"], g["channels"], g["trace_types"], prob["truth"], 0,
                    trace_types_out=tt_out)


def snr(prob, out):
    """10 log10(|truth|^2 / |out - truth|^2) over live traces, in dB."""
    num = den = 0.0
    for tr, o, tt in zip(prob["truth"], out, prob["geometry"]["trace_types"]):
        live = np.asarray(tt).reshape(-1) == LIVE
        t = np.asarray(tr, np.float64).reshape(len(live), -1)[live]
        e = np.asarray(o, np.float64).reshape(len(live), -1)[live] - t
        num += float(np.sum(t * t))
        den += float(np.sum(e * e))
    return 10.0 * np.log10(num / max(den, 1e-300))
