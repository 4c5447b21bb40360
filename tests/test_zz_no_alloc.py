"""agd_blend_noalloc (the scratch-free agd_blend) against agd_blend.

Both calls go through the C ABI (ctypes). agd_blend_noalloc takes its device
buffers from the library's fixed scratch array, which keeps the previous
call's contents: two calls in a row, the second with a smaller input, must
each equal agd_blend bit for bit (the same kernels in the same order). A call
whose buffers exceed the array is refused before any device work, and the
next call still works.
"""
import ctypes

import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd._marshal import C_INT, C_LONG, PtrTable, int_array

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _call(fn, shot, ch, tt, data, mode, taper):
    """fn(...) on one-patch-per-array inputs; output traces = input traces."""
    ntr = [int(np.prod(d.shape[:-1])) for d in data]
    nt = [int(d.shape[-1]) for d in data]
    a_ntr, p_ntr = int_array(ntr)
    a_nt, p_nt = int_array(nt)
    s = PtrTable([np.asarray(x).reshape(-1) for x in shot], C_LONG)
    c = PtrTable([np.asarray(x).reshape(-1) for x in ch], C_INT)
    t = PtrTable([np.asarray(x).reshape(-1) for x in tt], C_INT)
    # output traces at the input positions; AGDBad is not allowed there
    to = PtrTable([np.where(np.asarray(x) == 1, 0, x).reshape(-1) for x in tt], C_INT)
    d = PtrTable([np.asarray(x, np.float32).reshape(-1) for x in data], np.float32)
    out = PtrTable([np.full(int(np.prod(x.shape)), np.nan, np.float32) for x in data],
                   np.float32, writable=True)
    rc = fn(len(data), p_ntr, p_nt, s.ptr, c.ptr, t.ptr, d.ptr, int(mode), int(taper),
            len(data), p_ntr, p_nt, s.ptr, c.ptr, to.ptr, out.ptr)
    return rc, [o.reshape(x.shape) for o, x in zip(out.np, data)]


def _problem(seed, n_ch, n_src, n_t, interval):
    rng = np.random.default_rng(seed)
    shot = np.arange(n_src) * interval + rng.integers(-interval // 4, interval // 4 + 1, n_src)
    shot = np.broadcast_to(shot - shot.min(), (n_ch, n_src)).astype(np.int64)
    ch = np.broadcast_to(np.arange(n_ch)[:, None], (n_ch, n_src)).astype(np.intc)
    tt = rng.choice([0] * 18 + [1, 2], (n_ch, n_src)).astype(np.intc)
    data = rng.standard_normal((n_ch, n_src, n_t)).astype(np.float32)
    return [shot], [ch], [tt], [data]


@pytest.mark.parametrize("mode,taper", [(0, 0), (1, 0), (1, 24), (2, 0)])
def test_noalloc_matches_agd_blend_twice(mode, taper):
    lib = L.load()
    big = _problem(1, 16, 256, 1024, 600)     # ~16 MB in + 16 MB out
    small = _problem(2, 4, 48, 700, 450)      # a smaller second call
    for name, prob in (("first (larger)", big), ("second (smaller)", small)):
        print("agd_blend_noalloc %s call: mode %d taper %d, %s" %
              (name, mode, taper, prob[3][0].shape), flush=True)
        rc_ref, ref = _call(lib.agd_blend, *prob, mode, taper)
        rc, got = _call(lib.agd_blend_noalloc, *prob, mode, taper)
        assert rc_ref == 0 and rc == 0
        for a, b in zip(got, ref):
            assert not np.isnan(a).any()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_refuses_beyond_its_scratch(capfd):
    lib = L.load()
    # 17 M samples: input and output copies alone exceed the 64 MiB array
    prob = _problem(3, 16, 1040, 1024, 600)
    assert prob[3][0].nbytes * 2 > SCRATCH
    print("agd_blend_noalloc oversize call", flush=True)
    rc, _ = _call(lib.agd_blend_noalloc, *prob, 0, 0)
    ctypes.CDLL(None).fflush(None)
    err = capfd.readouterr().err
    assert rc == 1 and "more than its fixed" in err
    # the library still serves the next call
    small = _problem(4, 2, 32, 300, 200)
    rc_ref, ref = _call(lib.agd_blend, *small, 0, 0)
    rc, got = _call(lib.agd_blend_noalloc, *small, 0, 0)
    assert rc_ref == 0 and rc == 0
    assert np.array_equal(got[0].view(np.uint32), ref[0].view(np.uint32))
