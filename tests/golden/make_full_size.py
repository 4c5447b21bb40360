"""Writes tests/golden/full_size_deblend.npz: the FP64 oracle's agd_deblend
of two full-size workloads at their long iteration counts, for the GPU test
test_full_size.py::test_long_full_size_deblend_against_fp64_golden.

  C3  vibroseis (2 x 8 x 128 x 4096, wavelet 4000, windows 8 x 32 x 512),
      its full 150 iterations
  C1  (64 x 512 x 2048, windows 16 x 32 x 256, shots every 4 s +- 1 s),
      20 iterations

The oracle needs minutes for these (2.6 s per C3 and 24 s per C1 iteration
in FP64 on 8 cores), too long for every GPU test run, so its output is kept
here in compact form, per workload: a SHA-256 of the blended input (the GPU
side regenerates it from the same seeds and must match bit for bit), the
output of every k-th trace in full (float32), the FP64 energy of the whole
output and its projections on four seeded random +-1 vectors.

Run from the repository root: python tests/golden/make_full_size.py
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from agdeblend_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_size_deblend.npz")
CASES = {"c3": (lambda: synth.config3(), 150, 32), "c1": (lambda: synth.config1(
    interval_s=4.0, dither_s=1.0), 20, 512)}


def geometry_kw(g):
    keys = ["volumes", "n_dims", "window_shapes", "coords", "shapes", "shottimes", "channels",
            "trace_types", "wavelet_lengths", "wavelet_idxs", "wavelets"]
    return {k: g[k] for k in keys if k in g}


def input_hash(blended):
    h = hashlib.sha256()
    for b in blended:
        h.update(np.ascontiguousarray(b, np.float32).tobytes())
    return h.hexdigest()


def summary(out, stride):
    """every stride-th trace of the concatenated patches, energy, projections"""
    flat = np.concatenate([np.asarray(o, np.float64).reshape(-1, o.shape[-1]) for o in out])
    rng = np.random.default_rng(7)
    proj = np.array([float(np.sum(flat * rng.choice([-1.0, 1.0], flat.shape)))
                     for _ in range(4)])
    return flat[::stride].astype(np.float32), float(np.sum(flat * flat)), proj


def main():
    data = {}
    for name, (make, n_its, stride) in CASES.items():
        prob = make()
        kw = geometry_kw(prob["geometry"])
        blended = synth.blend_input(prob, O.blend_same_geometry)
        t0 = time.time()
        out, _ = O.deblend(data=blended, initial_factor=1.0, n_its=n_its, double=True, **kw)
        traces, energy, proj = summary(out, stride)
        print("%s: %d iterations in %.0f s, %d traces kept" % (name, n_its, time.time() - t0,
                                                              len(traces)), flush=True)
        data[name + "_hash"] = np.array(input_hash(blended))
        data[name + "_n_its"] = np.array(n_its)
        data[name + "_stride"] = np.array(stride)
        data[name + "_traces"] = traces
        data[name + "_energy"] = np.array(energy)
        data[name + "_proj"] = proj
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
