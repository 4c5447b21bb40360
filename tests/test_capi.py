"""The C boundary as a C caller sees it: tests/capi_check.c compiled with gcc
against include/agdeblend.h and linked with -lagdeblend, as the reference's C
examples link the reference library (examples/Makefile:244-245). The CPU
tests compile it (and check that the AGD_DOUBLE and AGD_MPI configurations,
which this library does not provide, stop at compile time); the GPU test runs
the reference API's agd_blend known answers (test.cpp:30-430) and a small
agd_deblend whose output is compared with the FP64 oracle (relative L2 <=
1e-4)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "agdeblend_amd", "lib")
SRC = os.path.join(ROOT, "tests", "capi_check.c")


def _compile(out, *defs):
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", INC, *defs, SRC, "-L",
           LIBDIR, "-lagdeblend", "-Wl,-rpath," + LIBDIR, "-o", out]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libagdeblend.so")):
        pytest.skip("libagdeblend.so is not built")
    out = str(tmp_path_factory.mktemp("capi") / "capi_check")
    r = _compile(out)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_c_caller_compiles_and_links(exe):
    assert os.access(exe, os.X_OK)


@pytest.mark.parametrize("macro,text", [("-DAGD_DOUBLE", "AGD_DOUBLE interface"),
                                        ("-DAGD_MPI", "AGD_MPI interface")])
def test_unprovided_configurations_stop_at_compile_time(tmp_path, macro, text):
    r = _compile(str(tmp_path / "x"), macro)
    assert r.returncode != 0 and text in r.stderr, r.stderr


@pytest.mark.gpu
def test_c_caller_on_gpu(exe, tmp_path):
    from agdeblend_amd import synth
    from agdeblend_amd import agdeblend as agd
    from oracle import oracle as O
    prob = synth.make_config(0, small=True, blend_fn=agd.blend)
    g = prob["geometry"]
    assert len(g["shapes"]) == 1 and len(g["shapes"][0]) == 2
    assert not np.any(g["channels"][0]) and not np.any(g["trace_types"][0])
    ntr, nt = g["shapes"][0]
    wtr, wt = g["window_shapes"][0]
    n_its = 8
    shot = np.ascontiguousarray(g["shottimes"][0], np.int64)
    data = np.ascontiguousarray(prob["blended"][0], np.float32)
    shot.tofile(tmp_path / "shot.i64")
    data.tofile(tmp_path / "data.f32")
    print("capi_check: C0-mini %d x %d, windows %d x %d, %d its" % (ntr, nt, wtr, wt, n_its),
          flush=True)
    r = subprocess.run([exe, str(ntr), str(nt), str(wtr), str(wt), str(n_its),
                        str(tmp_path / "shot.i64"), str(tmp_path / "data.f32"),
                        str(tmp_path / "out.f32")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
    out = np.fromfile(tmp_path / "out.f32", np.float32).reshape(ntr, nt)
    ref, _ = O.deblend(g["volumes"], g["n_dims"], g["window_shapes"], g["coords"], g["shapes"],
                       g["shottimes"], g["channels"], g["trace_types"], prob["blended"], 1.0,
                       n_its, 0, double=True)
    rel = np.linalg.norm(out - ref[0]) / np.linalg.norm(ref[0])
    print("C caller agd_deblend vs FP64 oracle: %.3e" % rel)
    assert rel <= 1e-4
