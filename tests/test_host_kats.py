"""The product library's own host tables against the reference's known
answers (CPU; agdx_create_host and agdx_blend_tables build the tables
without a device).

test_host_tables compares the product's tables with the oracle's; here the
expected values are the reference tests' own, transcribed as data with
file:line, so the product is pinned directly and not through the oracle.
"""
import numpy as np
import pytest

from agdeblend_amd import lib as L

LIVE, BAD, MISSING = 0, 1, 2


def _problem(volumes, n_dims, coords, shapes, wshapes, shot=None):
    """A deblend geometry around the reference's window KAT inputs: all
    traces live on one channel, shots far apart (the window tests give only
    space geometry)."""
    ntr = [int(np.prod(s[:-1])) for s in shapes]
    if shot is None:
        shot = [np.arange(n, dtype=np.int64) * 1000 + 100000 * p for p, n in enumerate(ntr)]
    return L.Problem(host_only=True, volumes=volumes, n_dims=n_dims, window_shapes=wshapes,
                     coords=coords, shapes=shapes, shottimes=shot,
                     channels=[np.zeros(n, np.intc) for n in ntr],
                     trace_types=[np.zeros(n, np.intc) for n in ntr])


# model_test.cpp:19-270 (Model.SetSpaceWindows): per window (patch_idx,
# n_traces, coords, space_shape, use_taper)
SET_SPACE_WINDOWS = [
    (0, 2, [0], [2], [0, 1]), (0, 2, [1], [2], [1, 1]), (0, 2, [2], [2], [1, 0]),
    (1, 6, [0, 0], [3, 2], [0, 0, 0, 1]), (1, 6, [0, 1], [3, 2], [0, 0, 1, 1]),
    (1, 6, [0, 2], [3, 2], [0, 0, 1, 1]), (1, 6, [0, 3], [3, 2], [0, 0, 1, 0]),
    (2, 10, [0, 0], [2, 5], [0, 1, 0, 1]), (2, 8, [0, 3], [2, 4], [0, 1, 1, 0]),
    (2, 10, [1, 0], [2, 5], [1, 0, 0, 1]), (2, 8, [1, 3], [2, 4], [1, 0, 1, 0]),
    (3, 8, [0, 0], [4, 2], [0, 1, 0, 1]), (4, 8, [0, 0], [4, 2], [1, 1, 1, 0]),
    (4, 8, [2, 0], [4, 2], [1, 0, 1, 0]),
]


def test_set_space_windows():
    pm = _problem([0, 1, 2, 3, 3], [2, 3, 3, 3], [[0], [0, 0], [0, 0], [0, 0], [1, 1]],
                  [[4, 8], [3, 5, 7], [3, 7, 7], [4, 2, 4], [6, 2, 4]],
                  [[2, 8], [3, 2, 6], [2, 4, 6], [4, 2, 4]])
    try:
        assert pm.n_windows == 3 + 4 + 4 + 3
        for w, (patch, ntr, coords, shape, ut) in enumerate(SET_SPACE_WINDOWS):
            a = pm.window(w)
            got = (a["patch"], a["ntr"], a["coords"], a["W"], pm.window_taper(w))
            assert got == (patch, ntr, coords, shape, ut), w
    finally:
        pm.close()


def test_set_space_windows_larger():
    # model_test.cpp:272-336: 78 traces, windows of 32 -> 37, 37, 36
    pm = _problem([0], [2], [[0]], [[78, 8]], [[32, 8]])
    try:
        assert pm.n_windows == 3
        got = [(pm.window(w)["ntr"], pm.window(w)["coords"], pm.window(w)["W"],
                pm.window_taper(w)) for w in range(3)]
        assert got == [(37, [0], [37], [0, 1]), (37, [21], [37], [1, 1]),
                       (36, [42], [36], [1, 0])]
    finally:
        pm.close()


def test_window_arrays():
    # model_test.cpp:338-434: patch (3, 7, 7), windows (2, 4, 6), shottime of
    # trace i = i: each window's traces in order
    pm = _problem([0], [3], [[0, 0]], [[3, 7, 7]], [[2, 4, 6]],
                  shot=[np.arange(21, dtype=np.int64)])
    expect = [[0, 1, 2, 3, 4, 7, 8, 9, 10, 11], [3, 4, 5, 6, 10, 11, 12, 13],
              [7, 8, 9, 10, 11, 14, 15, 16, 17, 18], [10, 11, 12, 13, 17, 18, 19, 20]]
    try:
        assert pm.n_windows == 4
        for w in range(4):
            assert pm.window_traces(w).tolist() == expect[w], w
    finally:
        pm.close()


def _intervals(ivs):
    """AddInterval through agd_blend's tables: one live trace on channel 0
    per interval [a, b), each in its own patch of b - a samples."""
    t = L.blend_tables([b - a for a, b in ivs], [[a] for a, _ in ivs], [[0]] * len(ivs),
                       [[LIVE]] * len(ivs))
    assert (t["intervals"][:, 0] == 0).all()
    return [tuple(int(v) for v in r[1:]) for r in t["intervals"]]


def test_add_interval():
    # blend_test.cpp:19-99
    assert _intervals([(100, 200)]) == [(100, 200)]
    assert _intervals([(100, 200), (300, 400)]) == [(100, 200), (300, 400)]
    assert _intervals([(100, 200), (150, 250)]) == [(100, 250)]
    assert _intervals([(50, 75), (100, 200), (250, 350), (400, 500), (175, 275)]) == \
        [(50, 75), (100, 350), (400, 500)]


def test_unique_channels():
    # blend_test.cpp:101-159: first-appearance order, Missing ignored
    t = L.blend_tables([10, 10], [np.zeros(4, np.int64), np.zeros(2, np.int64)],
                       [[123, 456, 0, 123], [123, -789]],
                       [[LIVE, BAD, MISSING, LIVE], [LIVE, LIVE]])
    assert t["channels"].tolist() == [123, 456, -789]
    assert t["coords"][:, 0].tolist() == [0, 1, -1, 0, 0, 2]


def test_blend_config_two_channels():
    # blend_test.cpp:161-210
    t = L.blend_tables([100, 50], [[25], [35]], [[1], [0]], [[LIVE], [LIVE]])
    assert t["channels"].tolist() == [1, 0]
    assert t["intervals"].tolist() == [[0, 25, 125], [1, 35, 85]]


def test_blend_params():
    # blend_test.cpp:212-252: a record [100, 250) on channel 0, two traces of
    # 50 samples at 100 and 200
    t = L.blend_tables([150], [[100]], [[0]], [[LIVE]], q_ntimes=50, q_shottimes=[100, 200],
                       q_channels=[0, 0], q_trace_types=[LIVE, LIVE])
    assert t["q_coords"].tolist() == [[0, 0, 0, 0, 50], [0, 0, 100, 0, 50]]
    # blend_test.cpp:254-289: partial overlap of [50, 150) with [100, 200)
    t = L.blend_tables([100], [[50]], [[0]], [[LIVE]], q_ntimes=100, q_shottimes=[100],
                       q_channels=[0], q_trace_types=[LIVE])
    assert t["q_coords"].tolist() == [[0, 0, 50, 0, 50]]
    # blend_test.cpp:291-309: a missing trace is not blended
    t = L.blend_tables([100], [[50]], [[0]], [[LIVE]], q_ntimes=50, q_shottimes=[-1],
                       q_channels=[-1], q_trace_types=[MISSING])
    assert t["q_coords"][0, 0] == -1


def test_mute_intervals():
    # blend.c:585-627: a bad trace's range, intersected with its record
    # interval, is muted; expected values worked by hand: record [0, 30) on
    # channel 7 from live [0, 20) and bad [10, 30); a second record [100,
    # 120) holds only a live trace
    t = L.blend_tables([20, 20, 20], [[0], [10], [100]], [[7], [7], [7]],
                       [[LIVE], [BAD], [LIVE]])
    assert t["intervals"].tolist() == [[0, 0, 30], [0, 100, 120]]
    assert t["mutes"].tolist() == [[0, 0, 10, 20]]
