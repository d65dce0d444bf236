"""ctypes loader for ``oracle/port.c`` -- TEST INFRASTRUCTURE ONLY (the oracle).

``build()`` compiles ``libport.so`` with gcc/OpenMP; every wrapper takes and
returns numpy arrays.  See ``port.c`` for the reference lines each routine
restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libport.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(os.path.join(_HERE, "port.c")):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "libport.so"], check=True)
    return _SO


def _load():
    global _lib
    if _lib is None:
        build()  # compiles when libport.so is missing or older than port.c
        _lib = ctypes.CDLL(_SO)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nthreads_default() -> int:
    return len(os.sched_getaffinity(0))


def ks_test(rmli: np.ndarray, az: int, r: int, return_dist=False, return_k=False, nthreads=None,
            per_pair=False):
    """Port of _sort_numba + _ks_test_numba for float32 stacks.  per_pair: evaluate
    the p series for every pixel pair as the reference does (same values, the
    reference's cost) instead of once per step count k."""
    lib = _load()
    rmli = np.ascontiguousarray(rmli, dtype=np.float32)
    nl, w, n = rmli.shape
    shape = (nl, w, 2 * az + 1, 2 * r + 1)
    p = np.empty(shape, np.float32)
    d = np.empty(shape, np.float32) if return_dist else None
    k = np.empty(shape, np.int32) if return_k else None
    lib.port_ks_test_f32(_p(rmli), nl, w, n, az, r, _p(p), _p(d) if d is not None else None,
                         _p(k) if k is not None else None, nthreads or nthreads_default(), int(per_pair))
    out = [p]
    if return_dist:
        out.insert(0, d)
    if return_k:
        out.append(k)
    return tuple(out) if len(out) > 1 else p


def emperical_co_pc(rslc, idx, pc_is_shp, image_pairs, nthreads=None):
    lib = _load()
    rslc = np.ascontiguousarray(rslc, dtype=np.complex64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    shp = np.ascontiguousarray(pc_is_shp, dtype=np.uint8)
    pairs = np.ascontiguousarray(image_pairs, dtype=np.int32)
    nl, w, n = rslc.shape
    aw, rw = shp.shape[1:]
    out = np.empty((idx.shape[0], pairs.shape[0]), np.complex64)
    lib.port_emperical_co_pc(_p(rslc), nl, w, n, _p(idx), _p(shp), idx.shape[0], aw, rw,
                             _p(pairs), pairs.shape[0], _p(out), nthreads or nthreads_default())
    return out


def ds_temp_coh(coh, ph, image_pairs, nthreads=None):
    lib = _load()
    coh = np.ascontiguousarray(coh, dtype=np.complex64)
    ph = np.ascontiguousarray(ph, dtype=np.complex64)
    pairs = np.ascontiguousarray(image_pairs, dtype=np.int32)
    out = np.empty(ph.shape[0], np.float32)
    lib.port_ds_temp_coh(_p(coh), _p(ph), ph.shape[0], ph.shape[1], _p(pairs), pairs.shape[0],
                         _p(out), nthreads or nthreads_default())
    return out


def temp_coh(intf, rslc, image_pairs, nthreads=None):
    lib = _load()
    intf = np.ascontiguousarray(intf, dtype=np.complex64)
    rslc = np.ascontiguousarray(rslc, dtype=np.complex64)
    pairs = np.ascontiguousarray(image_pairs, dtype=np.int32)
    out = np.empty(rslc.shape[0], np.float32)
    lib.port_temp_coh(_p(intf), _p(rslc), rslc.shape[0], rslc.shape[1], _p(pairs), pairs.shape[0],
                      _p(out), nthreads or nthreads_default())
    return out
