"""Python packaging: the wheel installs ``agdeblend`` (the reference wrapper's
import name, wrappers/python/agdeblend/__init__.py:1-10) and ``agdeblend_amd``
with the built library, and the installed wrapper applies the reference's
dtype rule (wrappers/python/agdeblend/agdeblend.py:326-336). CPU only: the
library is loaded, not called on a device."""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "agdeblend_amd", "lib", "libagdeblend.so")


@pytest.fixture(scope="module")
def installed(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("libagdeblend.so is not built")
    tmp = tmp_path_factory.mktemp("pkg")
    src = tmp / "src"
    src.mkdir()
    for f in ("setup.py", "pyproject.toml"):
        shutil.copy(os.path.join(ROOT, f), src / f)
    shutil.copytree(os.path.join(ROOT, "agdeblend_amd"), src / "agdeblend_amd",
                    ignore=shutil.ignore_patterns("__pycache__", "_build"))
    shutil.copytree(os.path.join(ROOT, "include"), src / "include")
    dist = tmp / "dist"
    r = subprocess.run([sys.executable, "-m", "pip", "wheel", str(src), "--no-deps",
                        "--no-build-isolation", "-w", str(dist), "-q"],
                       capture_output=True, text=True, cwd=str(tmp))
    assert r.returncode == 0, r.stdout + r.stderr
    wheels = [w for w in os.listdir(dist) if w.endswith(".whl")]
    assert len(wheels) == 1, wheels
    site = tmp / "site"
    r = subprocess.run([sys.executable, "-m", "pip", "install", "--no-deps", "--no-index",
                        "--target", str(site), "-q", str(dist / wheels[0])],
                       capture_output=True, text=True, cwd=str(tmp))
    assert r.returncode == 0, r.stdout + r.stderr
    return site


def run_in(site, code):
    env = dict(os.environ, PYTHONPATH=str(site))
    return subprocess.run([sys.executable, "-c", textwrap.dedent(code)], capture_output=True,
                          text=True, env=env, cwd=str(site.parent))


def test_wheel_contents(installed):
    assert (installed / "agdeblend" / "__init__.py").exists()
    assert (installed / "agdeblend_amd" / "lib" / "libagdeblend.so").exists()
    assert (installed / "agdeblend_amd" / "weights" / "unet_c4.npz").exists()
    assert not (installed / "oracle").exists()


def test_reference_import_names(installed):
    r = run_in(installed, """
        import agdeblend
        from agdeblend import (deblend, blend, AGDLive, AGDBad, AGDMissing,
                               AGDBlendSum, AGDBlendMean, AGDBlendOverwrite)
        assert (AGDLive, AGDBad, AGDMissing) == (0, 1, 2)
        assert (AGDBlendSum, AGDBlendMean, AGDBlendOverwrite) == (0, 1, 2)
        import agdeblend_amd, agdeblend_amd.lib as L
        assert agdeblend_amd.__file__.startswith(%r)
        lib = L.load()
        assert L.LIBPATH.startswith(%r)
        print("ok")
    """ % (str(installed), str(installed)))
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_dtype_rule(installed):
    # float16 data is refused before any device work, with the reference's message
    r = run_in(installed, """
        import numpy as np
        from agdeblend import deblend, blend
        for f in (lambda d: blend([np.zeros(4, np.int64)], [np.zeros(4, np.intc)],
                                  [np.zeros(4, np.intc)], d, 0),
                  lambda d: deblend([0], [[2, 4]], [[0]], [np.zeros(4, np.int64)],
                                    [np.zeros(4, np.intc)], [np.zeros(4, np.intc)],
                                    1.0, 1, 0, d)):
            try:
                f([np.zeros((4, 8), np.float16)])
            except RuntimeError as e:
                assert "Data must be float32 or float64" in str(e), e
            else:
                raise AssertionError("float16 accepted")
        print("ok")
    """)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
