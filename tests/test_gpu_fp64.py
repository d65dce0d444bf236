"""FP64 inputs keep their precision (the reference's dtype rules): complex128
stacks and coherence run the FP64 kernels and return complex128 (quality and
temporal coherence float32, as the reference returns them); float64 p is
compared with p_max in float64.  Checked against the FP64 oracle."""
import sys

import numpy as np
import pytest
import torch

from moraine_amd import synth
from oracle import moraine_ref as R

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _mr():
    import moraine_amd as mr

    return mr


def _c128_stack(nl, w, n, seed):
    # complex128 samples that are not complex64-representable (a downcast would show)
    z = synth.ds_stack(nl, w, n, seed=seed, ps_frac=0.1).astype(np.complex128)
    rng = np.random.default_rng(seed)
    return z * (1 + 1e-9 * rng.standard_normal(z.shape))


def test_select_shp_float64_compares_in_float64():
    mr = _mr()
    p = np.full((2, 3, 3, 5), 0.2)
    p[0, 0, 0, :] = np.nextafter(0.05, 0.0)  # rounds to 0.05 (or above) in float32
    p[1, 2, 1, 1] = np.nan
    _case("select_shp float64")
    m, c = mr.select_shp(p, 0.05)
    m_ref, c_ref = R.select_shp(p, 0.05)
    np.testing.assert_array_equal(m, m_ref)
    np.testing.assert_array_equal(c, c_ref)
    assert m[0, 0, 0].all() and c[0, 0] == 5
    m32, _ = mr.select_shp(p.astype(np.float32), 0.05)
    assert not m32[0, 0, 0].any()  # float32 input keeps float32 semantics


@pytest.mark.parametrize("n", [17, 60, 90])
def test_c128_coherence_emi_tcoh_vs_fp64(n):
    mr = _mr()
    az, r = 3, 5
    nl, w = 10, 12
    z = _c128_stack(nl, w, n, 500 + n)
    rng = np.random.default_rng(n)
    shp = rng.random((nl, w, 2 * az + 1, 2 * r + 1)) < 0.6
    idx = np.stack(np.nonzero(np.ones((nl, w), bool)), -1).astype(np.int32)
    pc = shp.reshape(-1, 2 * az + 1, 2 * r + 1)
    _case(f"complex128 emperical_co_pc / emi / ds_temp_coh n={n}")
    c = mr.emperical_co_pc(z, idx, pc)
    assert c.dtype == np.complex128
    c64 = R.emperical_co_pc64(z, idx, pc)
    # FP64 accumulation: 1e-12 per entry (|coh| <= 1)
    assert np.nanmax(np.abs(c - c64)) < 1e-12
    ph, q = mr.emi(c)
    assert ph.dtype == np.complex128 and q.dtype == np.float32
    ph64, q64 = R.emi64(c)
    err = np.abs(np.angle(ph * np.conj(ph64)))
    # FP64 eigen-solve on the same coherence: 1e-7 rad per point RMS (points with K < N
    # and a near-singular |Gamma| are where the FP32 path loses accuracy; FP64 does not)
    rms = np.sqrt(np.nanmean(err ** 2, axis=1))
    print(f"  phase RMS vs FP64: max {np.nanmax(rms):.2e}")
    assert np.nanmax(rms) < 1e-7
    np.testing.assert_allclose(q, q64.astype(np.float32), rtol=1e-6, atol=1e-6)
    t = mr.ds_temp_coh(c, ph)
    np.testing.assert_allclose(t, R.ds_temp_coh64(c, ph), rtol=1e-6, atol=1e-7)
    t2 = mr.temp_coh(c, ph)
    np.testing.assert_allclose(t2, R.temp_coh64(c, ph), rtol=1e-6, atol=1e-7)
    # the fused driver on complex128 equals the composition
    pf, qf, tf = mr.emperical_co_emi_temp_coh_pc(z, idx, pc)
    assert pf.dtype == np.complex128
    np.testing.assert_array_equal(pf, ph)
    np.testing.assert_array_equal(qf, q)
    np.testing.assert_array_equal(tf, t)


@pytest.mark.parametrize("n", [17, 70])
def test_c128_raster_uncompress_neighbours(n):
    mr = _mr()
    az, r = 2, 3
    nl, w = 7, 9
    z = _c128_stack(nl, w, n, 700 + n)
    shp = np.random.default_rng(n).random((nl, w, 2 * az + 1, 2 * r + 1)) < 0.7
    _case(f"complex128 emperical_co raster n={n}")
    cov, coh = mr.emperical_co(z, shp)
    assert cov.dtype == np.complex128 and coh.dtype == np.complex128
    cov64, coh64 = R.emperical_co64(z, shp)
    assert np.nanmax(np.abs(coh - coh64)) < 1e-12
    dg = np.real(np.diagonal(cov64, axis1=2, axis2=3))
    assert np.nanmax(np.abs(cov - cov64) / np.sqrt(dg[..., :, None] * dg[..., None, :])) < 1e-12
    # uncompress_coh keeps complex128 and equals the oracle exactly
    iu = np.triu_indices(n, 1)
    cc = coh[..., iu[0], iu[1]].reshape(-1, iu[0].size)
    np.testing.assert_array_equal(mr.uncompress_coh(cc), R.uncompress_coh(cc))
    # regularize_spectral: FP64 arithmetic equals the complex128 oracle bit for bit
    full = coh.reshape(-1, n, n)
    beta = np.random.default_rng(1).random(full.shape[0])
    np.testing.assert_array_equal(mr.regularize_spectral(full, beta), R.regularize_spectral(full, beta))
    # ad_intf_pc on complex128 images
    idx = np.array([[0, 0], [3, 4], [6, 8]], np.int32)
    m = np.random.default_rng(2).random((3, 5, 7)) < 0.6
    a = mr.ad_intf_pc(z[..., 0], z[..., 1], idx, m)
    assert a.dtype == np.complex128
    np.testing.assert_allclose(a, R.ad_intf_pc64(z[..., 0], z[..., 1], idx, m), rtol=0, atol=1e-13)


def test_c128_torch_tensors_stay_complex128():
    mr = _mr()
    z = torch.from_numpy(_c128_stack(6, 8, 20, 3)).cuda()
    idx = torch.tensor([[2, 3], [4, 5]], dtype=torch.int32, device="cuda")
    shp = torch.ones((2, 3, 5), dtype=torch.bool, device="cuda")
    _case("complex128 torch in/out")
    c = mr.emperical_co_pc(z, idx, shp)
    assert isinstance(c, torch.Tensor) and c.dtype == torch.complex128 and c.is_cuda
    ph, q = mr.emi(c)
    assert ph.dtype == torch.complex128 and q.dtype == torch.float32
