"""GPU: chunk streaming of a host stack equals the whole-tile drop-in chain.

ds_pipeline_streamed (bands with halo rows, pinned double buffers, a copy
stream) against ks_test_select_shp on the whole raster followed by
emperical_co_emi_temp_coh_pc on all its pixels: bit for bit, for band heights
that do and do not divide the raster and N on the one-wave and workgroup EMI
paths.
"""
import sys

import numpy as np
import pytest

from moraine_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nl,w,band", [(30, 70, 48, 16), (60, 64, 40, 64), (60, 53, 37, 20), (90, 36, 30, 12)])
def test_streamed_equals_whole_tile(n, nl, w, band):
    import moraine_amd as mr
    from moraine_amd.stream import ds_pipeline_streamed

    print(f"[gpu case] streamed n={n} {nl}x{w} band={band}", flush=True)
    sys.stdout.flush()
    z = synth.ds_stack(nl, w, n, seed=n + nl, ps_frac=0.1)
    az, r = 4, 6
    ph, q, t, cnt = ds_pipeline_streamed(z, az, r, 0.05, band_rows=band, batch=500)
    _, shp, cnt0 = mr.ks_test_select_shp(z, az, r, 0.05, return_p=False)  # KS on the amplitudes, as streamed
    idx = np.stack(np.nonzero(np.ones((nl, w), bool)), -1).astype(np.int32)
    ph0, q0, t0 = mr.emperical_co_emi_temp_coh_pc(z, idx, shp.reshape(-1, 2 * az + 1, 2 * r + 1))
    np.testing.assert_array_equal(cnt, cnt0)
    np.testing.assert_array_equal(ph.reshape(-1, n).view(np.uint32), ph0.view(np.uint32))
    np.testing.assert_array_equal(q.reshape(-1).view(np.uint32), q0.view(np.uint32))
    np.testing.assert_array_equal(t.reshape(-1).view(np.uint32), t0.view(np.uint32))
