"""Host array type returned by the API.

The reference returns immutable JAX arrays, and its callers update them with
the functional ``x.at[mask].set(v)`` form (images.py:89-91,116-118 and the
demo notebooks).  ``MkArray`` is a NumPy ndarray that also offers ``.at``:
``x.at[idx].set(v)`` / ``.add(v)`` / ``.multiply(v)`` / ``.get()`` return new
arrays and leave ``x`` unchanged, as JAX does.
"""
from __future__ import annotations

import numpy as np


class _AtIndexer:
    __slots__ = ("_arr", "_idx")

    def __init__(self, arr, idx):
        self._arr = arr
        self._idx = idx

    def _apply(self, fn):
        out = np.array(self._arr, copy=True).view(MkArray)
        fn(out, self._idx)
        return out

    def set(self, v):
        def f(o, i):
            o[i] = v
        return self._apply(f)

    def add(self, v):
        def f(o, i):
            o[i] += v
        return self._apply(f)

    def multiply(self, v):
        def f(o, i):
            o[i] *= v
        return self._apply(f)

    def get(self):
        return np.asarray(self._arr)[self._idx]


class _At:
    __slots__ = ("_arr",)

    def __init__(self, arr):
        self._arr = arr

    def __getitem__(self, idx):
        if isinstance(idx, MkArray):
            idx = np.asarray(idx)
        elif isinstance(idx, tuple):
            idx = tuple(np.asarray(i) if isinstance(i, MkArray) else i for i in idx)
        return _AtIndexer(self._arr, idx)


class MkArray(np.ndarray):
    """ndarray with a JAX-style functional ``.at`` updater."""

    def __new__(cls, data):
        return np.asarray(data).view(cls)

    @property
    def at(self):
        return _At(self)


def wrap(a):
    return np.asarray(a).view(MkArray)
