"""isPD / nearestPD on the GPU (moraine/co.py:409-463) against the oracle's
restatement, whose own pins are Higham's properties (tests/test_oracle.py):
the reference's nearestPD calls an undefined wherePD, so no reference output
exists; parity is against the mathematics."""
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


def _coh_mag(n, k, npts, seed):
    """|Gamma| of DS sample coherence with k samples (k < n: rank-deficient Gamma,
    |Gamma| usually indefinite)."""
    z = synth.ds_stack(1, npts * k, n, seed=seed).reshape(npts, k, n).astype(np.complex128)
    S = np.swapaxes(z, 1, 2) @ z.conj()
    d = np.sqrt(np.real(np.diagonal(S, axis1=1, axis2=2)))
    return np.abs(S / (d[:, :, None] * d[:, None, :]))


@pytest.mark.parametrize("n", [5, 17, 60, 121])
def test_is_pd_and_nearest_pd_vs_oracle(n):
    mr = _mr()
    rng = np.random.default_rng(n)
    X = rng.standard_normal((6, n, n))
    sym = (X + np.swapaxes(X, -1, -2)) / 2
    pd = X @ np.swapaxes(X, -1, -2) + n * np.eye(n)
    mag = _coh_mag(n, max(2, n // 2), 6, seed=n)
    stack = np.concatenate([sym, pd, mag, X])
    _case(f"isPD / nearestPD n={n}, {stack.shape[0]} matrices")
    np.testing.assert_array_equal(mr.isPD(stack), R.isPD(stack))
    out = mr.nearestPD(stack)
    ref = R.nearestPD(stack)
    assert out.dtype == np.float64
    nrm = np.linalg.norm(stack, axis=(-2, -1))[:, None, None]
    err = np.abs(out - ref) / nrm
    print(f"  max |GPU - oracle| / ||A||_F = {err.max():.2e}")
    # FP64 Jacobi SVD vs LAPACK SVD, and the shift loop: 1e-10 of ||A||_F per entry
    assert err.max() < 1e-10
    # every output is positive definite: by the library's own Cholesky test, and by
    # LAPACK's eigenvalues to rounding (the shift loop stops at a smallest eigenvalue
    # of the order of the spacing, where two Cholesky orderings may disagree)
    assert mr.isPD(out).all()
    assert (np.linalg.eigvalsh(out).min(axis=-1) >= -1e-14 * nrm[:, 0, 0]).all()
    np.testing.assert_array_equal(out, np.swapaxes(out, -1, -2))  # and exactly symmetric
    # PD input comes back unchanged (to the SVD's rounding)
    np.testing.assert_allclose(out[6:12], pd, rtol=0, atol=1e-12 * np.abs(pd).max())


def test_nearest_pd_float32_and_torch():
    mr = _mr()
    n = 30
    mag = _coh_mag(n, 12, 8, seed=3).astype(np.float32)
    assert not R.isPD(mag).all()
    _case("nearestPD float32 numpy and torch")
    out = mr.nearestPD(mag)
    assert out.dtype == np.float32 and mr.isPD(out).all()
    ref = R.nearestPD(mag)
    assert np.abs(out - ref).max() < 1e-5 * np.abs(mag).max()
    t = mr.nearestPD(torch.from_numpy(mag).cuda())
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32
    np.testing.assert_array_equal(t.cpu().numpy(), out)
    m = mr.isPD(torch.from_numpy(mag).cuda())
    assert m.dtype == torch.bool and m.shape == (8,)
    with pytest.raises(TypeError):
        mr.nearestPD(mag.astype(np.complex64))
