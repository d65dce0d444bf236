"""ctypes binding of libmoraine_hip.so (the C ABI in include/moraine_hip.h).

The library is loaded lazily on first use.  There is no fallback: if the
shared object is missing or fails to load, every entry point raises
``RuntimeError`` -- the hot path only runs on the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(_HERE, "libmoraine_hip.so")

_c = ctypes
_vp, _i, _ll, _d, _sz = _c.c_void_p, _c.c_int, _c.c_longlong, _c.c_double, _c.c_size_t

# name -> (restype, argtypes); mirrors include/moraine_hip.h
SIGNATURES = {
    "mr_last_error_string": (_c.c_char_p, []),
    "mr_abi_version": (_i, []),
    "mr_ks_workspace_bytes": (_sz, [_i, _i, _i, _i, _i, _i]),
    "mr_ks_test_f32": (_i, [_vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_ks_test_f64": (_i, [_vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_ks_test_f32_noalloc": (_i, [_vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "mr_select_shp_f32": (_i, [_vp, _ll, _i, _d, _vp, _vp, _vp]),
    "mr_rslc2amp": (_i, [_vp, _ll, _vp, _vp]),
    "mr_amp_disp": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "mr_ks_test_rslc_f32": (_i, [_vp, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_ks_tables_f32": (_i, [_i, _d, _vp, _vp, _vp]),
    "mr_emperical_co": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "mr_emperical_co_pc": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp]),
    "mr_emi_workspace_bytes": (_sz, [_ll, _i]),
    "mr_emi": (_i, [_vp, _ll, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "mr_uncompress_coh": (_i, [_vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_ad_intf_pc": (_i, [_vp, _vp, _i, _i, _vp, _vp, _ll, _i, _i, _vp, _vp]),
    "mr_regularize_spectral": (_i, [_vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_ds_temp_coh": (_i, [_vp, _vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_emi_ds_temp_coh": (_i, [_vp, _ll, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_co_emi_ds_temp_coh_pc_workspace_bytes": (_sz, [_ll, _i]),
    "mr_co_emi_ds_temp_coh_pc": (_i, [_vp, _i, _i, _i, _vp, _vp, _ll, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_temp_coh": (_i, [_vp, _vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_emi_refined_workspace_bytes": (_sz, [_ll, _i]),
    "mr_emi_refined": (_i, [_vp, _ll, _i, _i, _d, _vp, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mr_select_shp_f64": (_i, [_vp, _ll, _i, _d, _vp, _vp, _vp]),
    "mr_emperical_co_c128": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "mr_emperical_co_pc_c128": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp]),
    "mr_uncompress_coh_c128": (_i, [_vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_emi_c128_workspace_bytes": (_sz, [_ll, _i]),
    "mr_emi_c128": (_i, [_vp, _ll, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "mr_ad_intf_pc_c128": (_i, [_vp, _vp, _i, _i, _vp, _vp, _ll, _i, _i, _vp, _vp]),
    "mr_regularize_spectral_c128": (_i, [_vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_ds_temp_coh_c128": (_i, [_vp, _vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_temp_coh_c128": (_i, [_vp, _vp, _ll, _i, _vp, _i, _vp, _vp]),
    "mr_pd_workspace_bytes": (_sz, [_ll, _i]),
    "mr_is_pd_f32": (_i, [_vp, _ll, _i, _vp, _vp, _sz, _vp]),
    "mr_is_pd_f64": (_i, [_vp, _ll, _i, _vp, _vp, _sz, _vp]),
    "mr_nearest_pd_f32": (_i, [_vp, _ll, _i, _vp, _vp, _sz, _vp]),
    "mr_nearest_pd_f64": (_i, [_vp, _ll, _i, _vp, _vp, _sz, _vp]),
    "mr_ds_batch_workspace_bytes": (_sz, [_ll, _i]),
    "mr_ds_batch_fused_workspace_bytes": (_sz, [_ll, _i, _i]),
    "mr_ds_batch": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _ll, _i, _vp, _sz, _vp]),
}

_lib = None


def load():
    """Load the shared library (raises RuntimeError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO_PATH):
        raise RuntimeError(
            f"{SO_PATH} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()' or python moraine_amd/_build.py)")
    lib = ctypes.CDLL(SO_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HipError(RuntimeError):
    pass


def call(name: str, *args):
    """Call an ``int``-returning entry point and raise on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.mr_last_error_string().decode(errors="replace")
        kind = {1: ValueError, 2: HipError, 3: NotImplementedError}.get(rc, RuntimeError)
        raise kind(f"{name}: {msg}")
    return rc


def query(name: str, *args):
    return getattr(load(), name)(*args)
