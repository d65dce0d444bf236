"""Regenerate the seeded golden fixtures in this folder from the oracle.

Run from the repository root:  python tests/golden/make_golden.py
Inputs are seeded synthetic stacks (moraine_amd.synth); expected outputs come
from the oracle (numpy restatement, cross-checked against the C port and the
literal pure-python loops before writing).  printed.json holds the values the
reference notebooks print and is maintained by hand.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from moraine_amd import synth  # noqa: E402
from oracle import moraine_ref as R  # noqa: E402
from oracle import port as P  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def ks_case():
    z = synth.ds_stack(32, 40, 17, seed=11, ps_frac=0.1)
    rmli = (np.abs(z) ** 2).astype(np.float32)
    rmli[3, 5, 2] = np.nan
    rmli[20, 33, 16] = np.nan
    rmli[10, 10:14] = 2.5  # tie-heavy constant stacks
    k = R.ks_k_field(rmli, 5, 7)
    P.build()
    _, _, kp = P.ks_test(rmli, 5, 7, return_dist=True, return_k=True)
    assert np.array_equal(k, kp), "numpy and C restatements disagree"
    np.savez_compressed(os.path.join(OUT, "ks_cfg1.npz"), rmli=rmli, k=k.astype(np.int8), az=5, r=7)


def co_case():
    rng = np.random.default_rng(5)
    rslc = (rng.random((5, 10, 17), dtype=np.float32) + 1j * rng.random((5, 10, 17), dtype=np.float32)).astype(np.complex64)
    az, r = 1, 2
    is_shp = rng.choice([False, True], size=(5, 10, 3, 5), p=[0.3, 0.7])
    for i in range(5):
        for j in range(10):
            for k in range(3):
                for l in range(5):
                    if k == az and l == r:
                        is_shp[i, j, k, l] = True
                    if not (0 <= i + k - az < 5 and 0 <= j + l - r < 10):
                        is_shp[i, j, k, l] = False
    cov, coh = R.emperical_co64(rslc, is_shp)
    idx = np.stack(np.nonzero(np.ones((5, 10), bool)), -1).astype(np.int32)
    coh_pc = R.emperical_co_pc(rslc, idx, is_shp.reshape(-1, 3, 5))
    np.savez_compressed(os.path.join(OUT, "co_small.npz"), rslc=rslc, is_shp=is_shp, cov64=cov, coh64=coh,
                        idx=idx, coh_pc=coh_pc)


def emi_case():
    out = {}
    for n in (17, 30, 60):
        coh, exp = synth.closure_free_coh(8, n, seed=n)
        out[f"cf_coh_{n}"] = coh
        out[f"cf_ph_{n}"] = exp.astype(np.complex64)
        z = synth.ds_stack(4, 4, n, seed=100 + n, regions=False)
        idx = np.stack(np.nonzero(np.ones((4, 4), bool)), -1).astype(np.int32)
        shp = np.ones((16, 5, 5), bool)
        c = R.emperical_co_pc(z, idx, shp)
        ph, q = R.emi64(c)
        out[f"ds_coh_{n}"] = c
        out[f"ds_ph64_{n}"] = ph.astype(np.complex64)
        out[f"ds_q64_{n}"] = q
    np.savez_compressed(os.path.join(OUT, "emi_small.npz"), **out)


def tcoh_case():
    z = synth.ds_stack(6, 6, 17, seed=7)
    idx = np.stack(np.nonzero(np.ones((6, 6), bool)), -1).astype(np.int32)
    shp = np.ones((36, 3, 3), bool)
    c = R.emperical_co_pc(z, idx, shp)
    ph, _ = R.emi_ref32(c)
    t_ds = R.ds_temp_coh(c, ph)
    t_pq = R.temp_coh(c, ph)
    np.savez_compressed(os.path.join(OUT, "tcoh_small.npz"), coh=c, ph=ph, t_ds=t_ds, t_pq=t_pq)


if __name__ == "__main__":
    ks_case()
    co_case()
    emi_case()
    tcoh_case()
    print("golden fixtures written to", OUT)
