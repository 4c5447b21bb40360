"""Error behaviour of the C ABI against the oracle: invalid inputs are
rejected before any device work (so this runs on the CPU), with return value
1 and the reference's first "ERROR: ..." line on stderr
(src/agdeblend.c:247-509 for agd_deblend, :651-732 for agd_blend)."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

from agdeblend_amd import lib as L
from agdeblend_amd._marshal import PtrTable
from oracle import oracle as O


def _capture(fn):
    """(return value, stderr text) of fn(), capturing the C-level stderr."""
    libc = ctypes.CDLL(None)
    with tempfile.TemporaryFile() as tmp:
        libc.fflush(None)
        saved = os.dup(2)
        os.dup2(tmp.fileno(), 2)
        try:
            rc = fn()
            libc.fflush(None)
        finally:
            os.dup2(saved, 2)
            os.close(saved)
        tmp.seek(0)
        return rc, tmp.read().decode()


def _first_error(text):
    for line in text.splitlines():
        if line.startswith("ERROR: "):
            return line
    return None


def _base():
    ntr, nt = 16, 64
    return dict(volumes=[0], n_dims=[2], window_shapes=[[8, 32]], coords=[[0]],
                shapes=[[ntr, nt]], shottimes=[np.arange(ntr) * 40], channels=[np.zeros(ntr)],
                trace_types=[np.zeros(ntr)], wavelet_lengths=None, wavelet_idxs=None,
                wavelets=None, f=1.0, n_its=5, data=[np.zeros((ntr, nt), np.float32)])


def _deblend_both(c):
    geo = L.Geometry(c["volumes"], c["n_dims"], c["window_shapes"], c["coords"], c["shapes"],
                     c["shottimes"], c["channels"], c["trace_types"], c["wavelet_lengths"],
                     c["wavelet_idxs"], c["wavelets"])
    out = []
    for fn in (lambda d: L.load().agd_deblend(*geo.args(), float(c["f"]), int(c["n_its"]), 0, d),
               lambda d: O.lib(False).deblend(*geo.args(), float(c["f"]), int(c["n_its"]), 0,
                                              d, None)):
        data = PtrTable([np.array(a, np.float32).reshape(-1) for a in c["data"]], np.float32,
                        writable=True)
        out.append(_capture(lambda: fn(data.ptr)))
    return out


CASES = {
    "n_its": dict(n_its=0),
    "initial_factor_zero": dict(f=0.0),
    "initial_factor_above_one": dict(f=1.5),
    "n_dims": dict(n_dims=[1], window_shapes=[[8]], shapes=[[64]]),
    "window_too_small": dict(window_shapes=[[1, 32]]),
    "window_odd": dict(window_shapes=[[7, 32]]),
    "window_bigger_than_patch": dict(window_shapes=[[32, 32]]),
    "wavelet_length": dict(wavelet_lengths=[0], wavelet_idxs=[np.zeros(16)],
                           wavelets=[np.ones(1, np.float32)]),
    "wavelet_idx_negative": dict(wavelet_lengths=[4], wavelet_idxs=[-np.ones(16)],
                                 wavelets=[np.ones(4, np.float32)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_deblend_rejects_like_the_oracle(name):
    c = _base()
    c.update(CASES[name])
    (rc_p, err_p), (rc_o, err_o) = _deblend_both(c)
    assert rc_p == 1 and rc_o == 1
    assert _first_error(err_p) is not None
    assert _first_error(err_p) == _first_error(err_o), (err_p, err_o)
    assert "ERROR in agd_deblend" in err_p


BLEND_CASES = {
    "bad_output_trace": dict(tt_out=1),
    "negative_taper": dict(mode=1, taper=-1),
    "n_patches": dict(np_=0),
}


@pytest.mark.parametrize("name", sorted(BLEND_CASES))
def test_blend_rejects_like_the_oracle(name):
    c = dict(tt_out=0, mode=0, taper=0, np_=1)
    c.update(BLEND_CASES[name])
    ntr, nt = 4, 16
    ntra = np.array([ntr], np.intc)
    nta = np.array([nt], np.intc)
    shot = PtrTable([np.arange(ntr) * 8], np.int64)
    ch = PtrTable([np.zeros(ntr)], np.intc)
    tt = PtrTable([np.zeros(ntr)], np.intc)
    tto = PtrTable([np.full(ntr, c["tt_out"])], np.intc)
    res = []
    for blend in (L.load().agd_blend, O.lib(False).blend):
        d = PtrTable([np.ones(ntr * nt, np.float32)], np.float32)
        do = PtrTable([np.zeros(ntr * nt, np.float32)], np.float32, writable=True)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        res.append(_capture(lambda: blend(c["np_"], p(ntra), p(nta), shot.ptr, ch.ptr, tt.ptr,
                                          d.ptr, c["mode"], c["taper"], c["np_"], p(ntra),
                                          p(nta), shot.ptr, ch.ptr, tto.ptr, do.ptr)))
    (rc_p, err_p), (rc_o, err_o) = res
    assert rc_p == 1 and rc_o == 1
    assert _first_error(err_p) is not None
    assert _first_error(err_p) == _first_error(err_o), (err_p, err_o)
