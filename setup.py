"""Packaging of the MI355X AGDeblend library's Python interface.

Installs two import names: ``agdeblend``, the reference wrapper's
(wrappers/python/setup.py:19-38), so existing Python callers work unchanged,
and ``agdeblend_amd``, the package with libagdeblend.so (the HIP build) and
the learned warm start's weights. The library is built in-tree with make
(hipcc for gfx950) when it is not built yet; there is no CPU build of it.

    python -m pip wheel . --no-deps --no-build-isolation -w dist
"""
import os
import subprocess

from setuptools import setup
from setuptools.command.build_py import build_py

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ROOT, "agdeblend_amd", "lib", "libagdeblend.so")


class BuildWithLibrary(build_py):
    """build_py that first builds libagdeblend.so when it is missing."""

    def run(self):
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", os.path.join(ROOT, "agdeblend_amd", "csrc")],
                           check=True)
        super().run()


setup(
    name="agdeblend-mi355x",
    version="1.0.1",
    description="AGDeblend deblending and blending on AMD MI355X (HIP kernels behind the "
                "reference C API)",
    python_requires=">=3.8",
    packages=["agdeblend", "agdeblend_amd"],
    package_dir={"agdeblend": "agdeblend_amd/pyagdeblend"},
    package_data={"agdeblend_amd": ["lib/libagdeblend.so", "weights/*.npz"]},
    install_requires=["numpy"],
    extras_require={"warmstart": ["torch"]},
    cmdclass={"build_py": BuildWithLibrary},
    zip_safe=False,
)
