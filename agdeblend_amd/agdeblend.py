"""Python interface to the MI355X deblending library.

Mirrors the reference Python wrapper's ``deblend`` and ``blend``
(wrappers/python/agdeblend/agdeblend.py:13-147, 150-323): same argument
names, order, defaults, shape conventions and return values, with ctypes in
place of cffi. The MPI ``comm`` argument is accepted only as None.

Data types follow the reference wrapper's rule (``_get_fdtype``,
wrappers/python/agdeblend/agdeblend.py:326-336): the first data patch must be
float32 or float64 and the results come back in that type. The reference
serves float64 from its double-precision library (``_get_agd_lib``,
agdeblend.py:339-387); this library computes in float32 on the GPU only, so
float64 patches are computed in float32 and converted back to float64, with
a warning.
"""
import warnings

import numpy as np

from . import lib as _lib
from ._marshal import C_INT, C_LONG, PtrTable, int_array, null

AGDLive = 0
AGDBad = 1
AGDMissing = 2
AGDBlendSum = 0
AGDBlendMean = 1
AGDBlendOverwrite = 2


def _no_mpi(comm):
    if comm is not None:
        raise RuntimeError("MPI is not supported by the MI355X library")


def _get_fdtype(patch):
    """The results' dtype: that of the first data patch (float32 or float64)."""
    fdtype = np.asarray(patch).dtype
    if fdtype not in (np.float32, np.float64):
        raise RuntimeError("Data must be float32 or float64")
    if fdtype == np.float64:
        warnings.warn("libagdeblend (MI355X) computes in float32: float64 data is "
                      "converted to float32 and the results back to float64",
                      RuntimeWarning, stacklevel=3)
    return fdtype


def _check_n_patches(n_patches, var, name):
    if len(var) != n_patches:
        raise RuntimeError(name + " must have the same length as data")


def _check_n_traces(n_traces, var, name):
    if [int(np.prod(np.shape(v))) for v in var] != [int(n) for n in n_traces]:
        raise RuntimeError(name + " has an incorrect number of traces")


def deblend(volumes, window_shapes, coords, shottimes, channels, trace_types,
            initial_factor, n_its, print_freq, data, wavelet_idxs=None,
            wavelets=None, comm=None):
    """Deblends seismic data (agd_deblend). Returns the deblended patches."""
    _no_mpi(comm)
    L = _lib.load()
    n_patches = len(data)
    fdtype = _get_fdtype(data[0]) if n_patches > 0 else np.dtype(np.float32)
    data_np = [np.array(d, dtype=np.float32, copy=True, order="C") for d in data]
    if len(volumes) != n_patches:
        raise RuntimeError("volumes must have the same length as data")
    _check_n_patches(n_patches, coords, "coords")
    _check_n_patches(n_patches, shottimes, "shottimes")
    _check_n_patches(n_patches, channels, "channels")
    _check_n_patches(n_patches, trace_types, "trace_types")
    n_traces = [int(np.prod(d.shape[:-1])) for d in data_np]
    _check_n_traces(n_traces, shottimes, "shottimes")
    _check_n_traces(n_traces, channels, "channels")
    _check_n_traces(n_traces, trace_types, "trace_types")
    if (wavelet_idxs is None) != (wavelets is None):
        raise RuntimeError("If wavelet_idxs or wavelets is not None, they both must not be")
    wavelet_lengths = None
    if wavelet_idxs is not None:
        _check_n_patches(n_patches, wavelet_idxs, "wavelet_idxs")
        _check_n_traces(n_traces, wavelet_idxs, "wavelet_idxs")
        wavelet_lengths = [len(w) for w in wavelets]
        wavelets = [np.asarray(w, np.float32) for w in wavelets]
    n_volumes = int(np.max(volumes)) + 1
    n_dims = np.zeros(n_volumes, np.intc)
    for v, d in zip(volumes, data_np):
        n_dims[v] = d.ndim
    shapes = [np.array(d.shape, np.intc) for d in data_np]
    geo = _lib.Geometry(volumes, n_dims, window_shapes, coords, shapes, shottimes,
                        channels, trace_types, wavelet_lengths, wavelet_idxs, wavelets)
    dt = PtrTable(data_np, np.float32, writable=True)
    if L.agd_deblend(*geo.args(), float(initial_factor), int(n_its), int(print_freq),
                     dt.ptr):
        raise RuntimeError("Error in call to agd_deblend")
    return [a.astype(fdtype) for a in dt.np]


def blend(shottimes, channels, trace_types, data, blend_mode, taper_length=0,
          n_times_out=None, shottimes_out=None, channels_out=None,
          trace_types_out=None, comm=None, data_out=None):
    """Blends seismic data or modifies already blended data (agd_blend)."""
    _no_mpi(comm)
    L = _lib.load()
    if blend_mode not in (AGDBlendSum, AGDBlendMean, AGDBlendOverwrite):
        raise RuntimeError("Invalid blend_mode")
    fdtype = _get_fdtype(data[0]) if len(data) > 0 else np.dtype(np.float32)
    data_np = [np.require(d, np.float32, ["C", "A"]) for d in data]
    shottimes_out = shottimes if shottimes_out is None else shottimes_out
    channels_out = channels if channels_out is None else channels_out
    trace_types_out = trace_types if trace_types_out is None else trace_types_out
    n_patches = len(data_np)
    n_traces = [int(np.prod(d.shape[:-1])) for d in data_np]
    n_times = [int(d.shape[-1]) for d in data_np]
    if data_out is not None:
        n_patches_out = len(data_out)
        n_traces_out = [int(np.prod(d.shape[:-1])) for d in data_out]
        if n_times_out is None:
            n_times_out = [int(d.shape[-1]) for d in data_out]
        elif list(n_times_out) != [int(d.shape[-1]) for d in data_out]:
            raise RuntimeError("data_out does not match n_times_out")
        out_shapes = [d.shape for d in data_out]
    else:
        n_patches_out = len(shottimes_out)
        n_traces_out = [int(np.prod(np.shape(s))) for s in shottimes_out]
        if n_times_out is None:
            n_times_out = n_times
        out_shapes = [tuple(np.shape(s)) + (int(nt),) for s, nt in zip(shottimes_out, n_times_out)]
    if len(n_times_out) != n_patches_out:
        raise RuntimeError("n_times_out is not of the expected length")
    _check_n_patches(n_patches, shottimes, "shottimes")
    _check_n_patches(n_patches, channels, "channels")
    _check_n_patches(n_patches, trace_types, "trace_types")
    _check_n_patches(n_patches_out, shottimes_out, "shottimes_out")
    _check_n_patches(n_patches_out, channels_out, "channels_out")
    _check_n_patches(n_patches_out, trace_types_out, "trace_types_out")
    _check_n_traces(n_traces, shottimes, "shottimes")
    _check_n_traces(n_traces, channels, "channels")
    _check_n_traces(n_traces, trace_types, "trace_types")
    _check_n_traces(n_traces_out, shottimes_out, "shottimes_out")
    _check_n_traces(n_traces_out, channels_out, "channels_out")
    _check_n_traces(n_traces_out, trace_types_out, "trace_types_out")
    ntr, ntr_p = int_array(n_traces)
    nt, nt_p = int_array(n_times)
    ntro, ntro_p = int_array(n_traces_out)
    nto, nto_p = int_array(n_times_out)
    s = PtrTable([np.asarray(x).reshape(-1) for x in shottimes], C_LONG)
    c = PtrTable([np.asarray(x).reshape(-1) for x in channels], C_INT)
    t = PtrTable([np.asarray(x).reshape(-1) for x in trace_types], C_INT)
    d = PtrTable([x.reshape(-1) for x in data_np], np.float32)
    so = PtrTable([np.asarray(x).reshape(-1) for x in shottimes_out], C_LONG)
    co = PtrTable([np.asarray(x).reshape(-1) for x in channels_out], C_INT)
    to = PtrTable([np.asarray(x).reshape(-1) for x in trace_types_out], C_INT)
    out = PtrTable([np.zeros(int(np.prod(sh)), np.float32) for sh in out_shapes],
                   np.float32, writable=True)
    if L.agd_blend(n_patches, ntr_p, nt_p, s.ptr, c.ptr, t.ptr, d.ptr, int(blend_mode),
                   int(taper_length), n_patches_out, ntro_p, nto_p, so.ptr, co.ptr,
                   to.ptr, out.ptr):
        raise RuntimeError("Error in call to agd_blend")
    result = [o.reshape(sh).astype(fdtype, copy=False) for o, sh in zip(out.np, out_shapes)]
    if data_out is not None:
        for dst, src in zip(data_out, result):
            dst[...] = src
        return data_out
    return result
