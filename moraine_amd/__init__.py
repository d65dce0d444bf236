"""moraine_amd: MI355X (gfx950) implementation of moraine's distributed-scatterer hot path.

Same call signatures as moraine (v0.8.5) for: ks_test, select_shp,
emperical_co, emperical_co_pc, uncompress_coh, emi, ds_temp_coh, temp_coh,
emperical_co_emi_temp_coh_pc, TempNet, nimage_from_npair.  Every array
function runs in hand-written HIP kernels (libmoraine_hip.so, C ABI in
include/moraine_hip.h); there is no CPU fallback.
"""
from .calamp import rslc2amp
from .co import ad_intf_pc, emperical_co, emperical_co_pc, isPD, nearestPD, regularize_spectral, uncompress_coh
from .pl import ds_temp_coh, emi, emperical_co_emi_temp_coh_pc
from .pqm import temp_coh
from .ps import amp_disp
from .shp import ks_test, ks_test_select_shp, select_shp
from .tnet import TempNet, nimage_from_npair

__version__ = "0.1.0"

__all__ = [
    "ks_test", "ks_test_select_shp", "select_shp", "emperical_co", "emperical_co_pc", "uncompress_coh", "emi",
    "ds_temp_coh", "temp_coh", "emperical_co_emi_temp_coh_pc", "TempNet", "nimage_from_npair", "rslc2amp",
    "amp_disp", "ad_intf_pc", "regularize_spectral", "isPD", "nearestPD",
]
