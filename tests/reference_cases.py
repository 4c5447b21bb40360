"""Inputs of the reference's end-to-end deblending test, regenerated.

src/test.cpp:432-561 (TEST AGDeblend.Deblend) draws five random problems
from glibc's rand() after srand(1): 1-3 patches of 2-4 dimensions, every
dimension 8-16 long, windows 8..dim rounded down to a multiple of 4 (always
for time), random shot times and channels, all traces live, and one random
value per time sample repeated over every trace. The data are blended with
AGDBlendSum, deblended with initial_factor 1 for 800 iterations, and the test
requires ||out - true||^2 < ||true||^2 / 10.

This module replays the same rand() sequence through the C library itself
(glibc, as on the reference's test machines), so the inputs are the
reference's bit for bit; expected results are the test's own criterion.
"""
import ctypes
import ctypes.util

import numpy as np

N_ITS = 800
RAND_MAX = 2147483647


def _libc():
    return ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")


def cases():
    """The five problems, in the test's order: list of dicts with the
    agd_deblend geometry, the true data and the blended data (the latter
    is filled by the caller's blend, as the test blends with agd_blend)."""
    c = _libc()
    c.srand(1)
    rand = c.rand
    out = []
    for _ in range(5):
        n_patches = 1 + rand() % 3
        n_dims = [2 + rand() % 3 for _ in range(n_patches)]
        shapes, windows = [], []
        for p in range(n_patches):
            ds, ws = [], []
            for d in range(n_dims[p]):
                s = 8 + rand() % 9
                w = 8 + rand() % (s - 8 + 1)
                if w != s or d == n_dims[p] - 1:
                    w = (w // 4) * 4
                ds.append(s)
                ws.append(w)
            shapes.append(ds)
            windows.append(ws)
        n_traces = [int(np.prod(s[:-1])) for s in shapes]
        n_times = [s[-1] for s in shapes]
        n_times_total = sum(a * b for a, b in zip(n_traces, n_times))
        shottime_max = n_times_total // 3
        channel_max = 1 + sum(n_traces) // 10
        shottimes, channels = [], []
        for p in range(n_patches):
            shottimes.append(np.array([rand() % shottime_max for _ in range(n_traces[p])],
                                      np.int64))
            channels.append(np.array([rand() % channel_max for _ in range(n_traces[p])],
                                     np.intc))
        truth = []
        for p in range(n_patches):
            t = np.empty((n_traces[p], n_times[p]), np.float32)
            for i in range(n_times[p]):
                v = np.float32(np.float32(rand()) / np.float32(RAND_MAX)) - np.float32(0.5)
                t[:, i] = v
            truth.append(t.reshape(shapes[p]))
        out.append(dict(
            volumes=list(range(n_patches)), n_dims=n_dims, window_shapes=windows,
            coords=[np.zeros(max(1, nd - 1), np.intc) for nd in n_dims],
            shapes=shapes, shottimes=shottimes, channels=channels,
            trace_types=[np.zeros(n, np.intc) for n in n_traces], truth=truth,
            n_traces=n_traces, n_times=n_times))
    return out


def passes(out, truth):
    """The test's criterion: ||out - true||^2 < ||true||^2 / 10 (float32
    accumulation like AGD_TYPE, reduced here in float64)."""
    e = sum(float(np.sum((np.asarray(o, np.float64) - t) ** 2)) for o, t in zip(out, truth))
    s = sum(float(np.sum(np.asarray(t, np.float64) ** 2)) for t in truth)
    return e < s / 10, (e / s) if s else 0.0
