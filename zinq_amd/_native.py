"""Loader for the in-tree native libraries of zinq_amd.

``libzinq_ints.so`` is the host integral engine (C++), ``libzinq_hip.so`` the HIP library with the
C-ABI of include/zinq_hip.h.  Both are built in-tree by ``zinq_amd.build`` (``__graft_entry__.build``).
There is no fallback: a missing library raises, so nothing can silently run a CPU substitute.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_cache = {}


def lib_path(name):
    return os.path.join(_HERE, name)


def load(name):
    """Load ``zinq_amd/<name>`` once; raise if it has not been built."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    _cache[name] = lib
    return lib


def ptr(a):
    """ctypes void pointer to the data of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)
