"""The reference wrapper's import name, ``agdeblend``.

Installed as the top-level package ``agdeblend`` (setup.py maps it here), so
callers of the reference Python wrapper (``from agdeblend import deblend``,
examples/example_4.py:3) run unchanged. It exports the names of the
reference package (wrappers/python/agdeblend/__init__.py:1-10), served by
the MI355X library.
"""
from agdeblend_amd.agdeblend import (  # noqa: F401
    AGDBad,
    AGDBlendMean,
    AGDBlendOverwrite,
    AGDBlendSum,
    AGDLive,
    AGDMissing,
    blend,
    deblend,
)
