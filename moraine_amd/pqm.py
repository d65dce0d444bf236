"""Pixel quality: temporal coherence (drop-in for moraine.temp_coh, moraine/pqm.py:124-159).

Accepts 2-D point-cloud or 3-D raster input like the reference; unlike the
reference's numba path it never normalises the caller's arrays in place.
"""
from __future__ import annotations

from .pl import _tcoh


def temp_coh(intf, rslc, image_pairs=None, block_size: int = 128):
    assert intf.ndim == rslc.ndim
    return _tcoh("mr_temp_coh", intf, rslc, image_pairs)
