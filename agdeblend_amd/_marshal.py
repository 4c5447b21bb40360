"""ctypes marshalling of AGDeblend's pointer-of-pointer arguments.

The C interface (include/agdeblend.h) takes per-patch arrays as
``T const *const *``; these helpers turn Python lists of array-likes into
contiguous NumPy arrays plus a ctypes pointer table that keeps them alive.
This plays the role of the reference Python wrapper's ``_convert_to_numpy``
(wrappers/python/agdeblend/agdeblend.py:390-408), with ctypes in place of
cffi.
"""
import ctypes

import numpy as np

C_INT = np.dtype(np.intc)
C_LONG = np.dtype(np.int_)  # C long (LP64: 64-bit)


class PtrTable:
    """A list of C-contiguous arrays and a ``T**`` table pointing at them."""

    def __init__(self, arrays, dtype, writable=False):
        reqs = ["C", "A"] + (["W"] if writable else [])
        self.np = [np.require(a, dtype=dtype, requirements=reqs) for a in arrays]
        self._table = (ctypes.c_void_p * max(1, len(self.np)))()
        for i, a in enumerate(self.np):
            self._table[i] = a.ctypes.data
        self.ptr = ctypes.cast(self._table, ctypes.c_void_p)

    def __len__(self):
        return len(self.np)


def int_array(values):
    a = np.require(np.asarray(values), dtype=C_INT, requirements=["C", "A"])
    return a, a.ctypes.data_as(ctypes.c_void_p)


def null():
    return ctypes.c_void_p(0)
