"""mr_ks_test_f32_noalloc (KS + SHP selection with the library's static device
scratch instead of a caller workspace) against mr_ks_test_f32 on the same
stacks: p, dist, mask and SHP counts bit for bit, on two calls in a row, the
second with a smaller stack (the scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest
import torch

from moraine_amd import _dev, _lib, synth

pytestmark = pytest.mark.gpu


def _outputs(nl, w, az, r, dev):
    shape = (nl, w, 2 * az + 1, 2 * r + 1)
    return (torch.full(shape, -1.0, dtype=torch.float32, device=dev),
            torch.full(shape, -1.0, dtype=torch.float32, device=dev),
            torch.zeros(shape, dtype=torch.uint8, device=dev),
            torch.full((nl, w), -7, dtype=torch.int32, device=dev))


def _ks(x, az, r, p_max, noalloc):
    nl, w, n = x.shape
    p, d, m, c = _outputs(nl, w, az, r, x.device)
    if noalloc:
        _lib.call("mr_ks_test_f32_noalloc", _dev.ptr(x), nl, w, n, az, r, p_max, _dev.ptr(p), _dev.ptr(d),
                  _dev.ptr(m), _dev.ptr(c), _dev.stream())
    else:
        ws = _dev.workspace(_lib.query("mr_ks_workspace_bytes", nl, w, n, az, r, 4), x.device)
        _lib.call("mr_ks_test_f32", _dev.ptr(x), nl, w, n, az, r, p_max, _dev.ptr(p), _dev.ptr(d), _dev.ptr(m),
                  _dev.ptr(c), _dev.ptr(ws), ws.numel(), _dev.stream())
    return p, d, m, c


def _same(a, b):
    for x, y in zip(a, b):
        xa, ya = x.cpu().numpy(), y.cpu().numpy()
        if xa.dtype == np.float32:
            xa, ya = xa.view(np.uint32), ya.view(np.uint32)
        np.testing.assert_array_equal(xa, ya)


def test_noalloc_ks_equals_workspace_ks_two_calls():
    torch.cuda.set_device(0)
    cases = [  # (nl, w, n, az, r, seed): the second call is smaller than the first
        (96, 120, 60, 7, 10, 11),
        (40, 52, 17, 5, 7, 12),
    ]
    for nl, w, n, az, r, seed in cases:
        print(f"noalloc KS case nl={nl} w={w} n={n} az={az} r={r} seed={seed}", flush=True)
        z = synth.ds_stack(nl, w, n, seed=seed, ps_frac=0.1)
        x = torch.from_numpy((np.abs(z) ** 2).astype(np.float32)).cuda()
        x[3, 5, 2] = float("nan")  # a NaN stack exercises the invalid-pixel rule
        ref = _ks(x, az, r, 0.05, noalloc=False)
        got = _ks(x, az, r, 0.05, noalloc=True)
        torch.cuda.synchronize()
        _same(got, ref)
        assert int(ref[3].sum()) > nl * w  # the SHP counts are non-trivial


def test_noalloc_ks_rejects_stack_larger_than_scratch():
    # 2048 x 2048 x 60 float32 keys need 960 MiB, above the 128 MiB scratch
    with pytest.raises(NotImplementedError, match="static scratch"):
        _lib.call("mr_ks_test_f32_noalloc", None, 2048, 2048, 60, 7, 10, 0.05, None, None, None, None, None)
