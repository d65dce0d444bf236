"""Seeded synthetic SLC stacks for the DS hot path (workload generator).

Distributed-scatterer model: per pixel z = s * (L w) * exp(j*phi) with
w ~ CN(0, I_N), L = chol(Gamma), Gamma_ik = 0.6*exp(-|t_i - t_k| / 24 d) + 0.1
(i != k), t_i = 12 d * i, phi_i = (4*pi/lambda) * v * t_i, v ~ U[0, 10] mm/yr,
lambda = 5.55 cm (C band).  Amplitudes are therefore Rayleigh distributed with
scale s.  Brightness regions: background scale 1 plus three rectangles with
scales 2, 0.5 and 3.  PS-like pixels (optional fraction) have a constant
amplitude over time and phase noise with coherence 0.95.

``ds_stack`` returns numpy arrays (small, CPU-checkable cases);
``ds_stack_torch`` generates the same model directly on a torch device (the
full-size tiles used for throughput measurement).  The two are statistically
identical, not bitwise identical.
"""
from __future__ import annotations

import math

import numpy as np

WAVELENGTH_M = 0.0555
DT_DAYS = 12.0


def coherence_model(n: int, g0: float = 0.6, tau: float = 24.0, floor: float = 0.1, dt: float = DT_DAYS):
    t = dt * np.arange(n)
    g = g0 * np.exp(-np.abs(t[:, None] - t[None, :]) / tau) + floor
    np.fill_diagonal(g, 1.0)
    return g


def region_scale(nl: int, w: int) -> np.ndarray:
    s = np.ones((nl, w), np.float32)
    s[: nl // 2, w // 2:] = 2.0
    s[nl // 2:, : w // 3] = 0.5
    s[nl // 3: (2 * nl) // 3, w // 3: (2 * w) // 3] = 3.0
    return s


def _phase(v_mm_yr, n):
    t = DT_DAYS * np.arange(n)
    v_m_day = v_mm_yr / 1000.0 / 365.25
    return (4 * np.pi / WAVELENGTH_M) * v_m_day[..., None] * t


def ds_stack(nl: int, w: int, n: int, seed: int = 0, ps_frac: float = 0.0, regions: bool = True,
             g0: float = 0.6, tau: float = 24.0) -> np.ndarray:
    """(nl, w, n) complex64 stack from the DS/PS model above (numpy)."""
    rng = np.random.default_rng(seed)
    L = np.linalg.cholesky(coherence_model(n, g0, tau))
    wv = (rng.standard_normal((nl, w, n)) + 1j * rng.standard_normal((nl, w, n))) / math.sqrt(2.0)
    z = wv @ L.T
    v = rng.uniform(0.0, 10.0, (nl, w))
    ph = _phase(v, n)
    s = region_scale(nl, w) if regions else np.ones((nl, w), np.float32)
    z = z * np.exp(1j * ph) * s[..., None]
    if ps_frac > 0:
        is_ps = rng.random((nl, w)) < ps_frac
        sigma = math.sqrt(-2.0 * math.log(0.95))
        amp = s * rng.uniform(3.0, 6.0, (nl, w))
        eps = rng.standard_normal((nl, w, n))
        zps = amp[..., None] * np.exp(1j * (ph + sigma * eps))
        z = np.where(is_ps[..., None], zps, z)
    return z.astype(np.complex64)


def ds_stack_torch(nl: int, w: int, n: int, seed: int = 0, ps_frac: float = 0.0, device="cuda",
                   regions: bool = True, row_block: int = 256):
    """Same model generated on a torch device in row blocks; returns complex64 (nl, w, n)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    L = torch.from_numpy(np.linalg.cholesky(coherence_model(n)).astype(np.float32)).to(device)
    s_all = torch.from_numpy(region_scale(nl, w) if regions else np.ones((nl, w), np.float32)).to(device)
    t = torch.arange(n, device=device, dtype=torch.float32) * DT_DAYS
    out = torch.empty((nl, w, n), dtype=torch.complex64, device=device)
    sigma = math.sqrt(-2.0 * math.log(0.95))
    for r0 in range(0, nl, row_block):
        r1 = min(nl, r0 + row_block)
        shp = (r1 - r0, w, n)
        re = torch.randn(shp, generator=g, device=device) @ L.T
        im = torch.randn(shp, generator=g, device=device) @ L.T
        v = torch.rand((r1 - r0, w), generator=g, device=device) * 10.0
        ph = (4 * math.pi / WAVELENGTH_M) * (v / 1000.0 / 365.25)[..., None] * t
        s = s_all[r0:r1, :, None]
        amp_re = re / math.sqrt(2.0) * s
        amp_im = im / math.sqrt(2.0) * s
        c, sn = torch.cos(ph), torch.sin(ph)
        zr = amp_re * c - amp_im * sn
        zi = amp_re * sn + amp_im * c
        if ps_frac > 0:
            is_ps = (torch.rand((r1 - r0, w), generator=g, device=device) < ps_frac)[..., None]
            amp = s * (3.0 + 3.0 * torch.rand((r1 - r0, w, 1), generator=g, device=device))
            pp = ph + sigma * torch.randn(shp, generator=g, device=device)
            zr = torch.where(is_ps, amp * torch.cos(pp), zr)
            zi = torch.where(is_ps, amp * torch.sin(pp), zi)
        out[r0:r1] = torch.complex(zr, zi)
    return out


def closure_free_coh(n_pts: int, n: int, seed: int = 0, ref: int = 0):
    """Compressed coherence of closure-free Gamma = |Gamma| o theta theta^H.

    For such inputs EMI has the closed-form answer ph = theta*conj(theta_ref)/|.|
    with quality 1 (the Schur product |Gamma|^-1 o |Gamma| has smallest
    eigenvalue 1 with eigenvector 1).  Returns (coh (n_pts, P) complex64,
    theta_ref (n_pts, n) complex128 expected phase history).
    """
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    coh = np.empty((n_pts, iu[0].size), np.complex128)
    exp = np.empty((n_pts, n), np.complex128)
    for i in range(n_pts):
        g0 = rng.uniform(0.3, 0.9)
        tau = rng.uniform(12.0, 96.0)
        A = coherence_model(n, g0, tau)
        th = np.exp(1j * rng.uniform(-np.pi, np.pi, n))
        G = A * np.outer(th, th.conj())
        coh[i] = G[iu]
        e = th * np.conj(th[ref])
        exp[i] = e / np.abs(e)
    return coh.astype(np.complex64), exp


def ds_samples(n_px: int, n_choices=(30, 60, 90, 150), k_range=(20, 315), g0_range=(0.3, 0.9),
               tau_range=(12.0, 96.0), seed: int = 0):
    """Per-pixel DS sample matrices for the batched form (numpy, small cases).

    Pixel i: N_i drawn from ``n_choices``, K_i ~ U[k_range] (inclusive), and K_i
    independent samples z = (L w) exp(j phi) of the exponential-decay model
    Gamma_ik = g0 exp(-|t_i - t_k| / tau) + 0.1 with g0 ~ U[g0_range],
    tau ~ U[tau_range] days.  Returns a list of (K_i, N_i) complex64 arrays.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_px):
        n = int(rng.choice(n_choices))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        L = np.linalg.cholesky(coherence_model(n, rng.uniform(*g0_range), rng.uniform(*tau_range)))
        wv = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))) / math.sqrt(2.0)
        ph = _phase(np.array(rng.uniform(0.0, 10.0)), n)
        out.append(((wv @ L.T) * np.exp(1j * ph)).astype(np.complex64))
    return out


def ds_samples_torch(n_px: int, n_choices=(30, 60, 90, 150), k_range=(20, 315), g0_range=(0.3, 0.9),
                     tau_range=(12.0, 96.0), seed: int = 0, device="cuda", chunk: int = 8192):
    """The ``ds_samples`` model generated on a torch device in ragged form.

    Returns (samples, row_off, n_images, n_samples): samples is a flat complex64
    tensor holding pixel i's (K_i, N_i) row-major matrix at complex element
    row_off[i] (int64 tensor); n_images is a host int32 array, n_samples an
    int32 tensor.  Pixels are stored in index order.
    """
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    rng = np.random.default_rng(seed)
    n_img = rng.choice(np.asarray(n_choices, np.int32), n_px).astype(np.int32)
    n_smp = rng.integers(k_range[0], k_range[1] + 1, n_px).astype(np.int32)
    sizes = n_img.astype(np.int64) * n_smp
    row_off = np.zeros(n_px, np.int64)
    row_off[1:] = np.cumsum(sizes)[:-1]
    samples = torch.empty(int(sizes.sum()), dtype=torch.complex64, device=device)
    for n in sorted(set(n_img.tolist())):
        ids = np.nonzero(n_img == n)[0]
        t = torch.arange(n, device=device, dtype=torch.float64) * DT_DAYS
        dt = (t[:, None] - t[None, :]).abs()
        for c0 in range(0, ids.size, chunk):
            sel = ids[c0:c0 + chunk]
            m = sel.size
            kmax = int(n_smp[sel].max())
            g0 = torch.rand(m, generator=g, device=device, dtype=torch.float64) * (g0_range[1] - g0_range[0]) + g0_range[0]
            tau = torch.rand(m, generator=g, device=device, dtype=torch.float64) * (tau_range[1] - tau_range[0]) + tau_range[0]
            G = g0[:, None, None] * torch.exp(-dt[None] / tau[:, None, None]) + 0.1
            G.diagonal(dim1=1, dim2=2).fill_(1.0)
            L = torch.linalg.cholesky(G).to(torch.float32)
            wr = torch.randn((m, kmax, n), generator=g, device=device) / math.sqrt(2.0)
            wi = torch.randn((m, kmax, n), generator=g, device=device) / math.sqrt(2.0)
            zr = wr @ L.transpose(1, 2)
            zi = wi @ L.transpose(1, 2)
            v = torch.rand(m, generator=g, device=device) * 10.0
            ph = (4 * math.pi / WAVELENGTH_M) * (v / 1000.0 / 365.25)[:, None] * t.float()
            c, s = torch.cos(ph)[:, None, :], torch.sin(ph)[:, None, :]
            z = torch.complex(zr * c - zi * s, zr * s + zi * c)
            keep = torch.arange(kmax, device=device)[None, :] < torch.from_numpy(n_smp[sel]).to(device)[:, None]
            rows = z[keep]  # (sum K, n) in pixel order
            # pixels of one N group are not contiguous in `samples`: scatter row blocks
            dst = torch.from_numpy(np.repeat(row_off[sel], n_smp[sel] * n)
                                   + _ragged_arange(n_smp[sel] * n)).to(device)
            samples[dst] = rows.reshape(-1)
    return samples, torch.from_numpy(row_off).to(device), n_img, torch.from_numpy(n_smp).to(device)


def _ragged_arange(lengths: np.ndarray) -> np.ndarray:
    """Concatenation of arange(l) for each l in lengths (int64)."""
    lengths = np.asarray(lengths, np.int64)
    tot = int(lengths.sum())
    if tot == 0:
        return np.zeros(0, np.int64)
    starts = np.zeros(lengths.size, np.int64)
    starts[1:] = np.cumsum(lengths)[:-1]
    return np.arange(tot, dtype=np.int64) - np.repeat(starts, lengths)


def ds_stack_mixed(nl: int, w: int, n: int, seed: int = 0, g0_range=(0.2, 0.9), near_singular_frac: float = 0.02,
                   tau: float = 24.0):
    """(nl, w, n) complex64 stack with per-pixel coherence levels (precision form).

    Pixel coherence Gamma_ik = g0 exp(-|t_i - t_k| / tau) + 0.1 with g0 ~ U[g0_range];
    a fraction ``near_singular_frac`` of pixels is near-singular instead:
    Gamma_ik = g0 with g0 ~ U[0.98, 0.999] for all i != k (almost rank one).
    Returns (stack, near_singular bool (nl, w)).
    """
    rng = np.random.default_rng(seed)
    ns = rng.random((nl, w)) < near_singular_frac
    z = np.empty((nl, w, n), np.complex128)
    v = rng.uniform(0.0, 10.0, (nl, w))
    ph = _phase(v, n)
    for i in range(nl):
        for j in range(w):
            if ns[i, j]:
                G = np.full((n, n), rng.uniform(0.98, 0.999))
                np.fill_diagonal(G, 1.0)
            else:
                G = coherence_model(n, rng.uniform(*g0_range), tau)
            L = np.linalg.cholesky(G)
            wv = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / math.sqrt(2.0)
            z[i, j] = (L @ wv) * np.exp(1j * ph[i, j])
    return z.astype(np.complex64), ns


def ds_stack_mixed_torch(nl: int, w: int, n: int, seed: int = 0, g0_range=(0.2, 0.9), near_singular_frac: float = 0.02,
                         tau: float = 24.0, levels: int = 64, device="cuda", row_block: int = 128):
    """The ``ds_stack_mixed`` model generated on a torch device (large tiles).

    g0 ~ U[g0_range] per pixel, quantised to ``levels`` values (one Cholesky
    factor per level); near-singular pixels take Gamma_ik = g0, g0 ~ U[0.98,
    0.999], in 16 levels.  Linear deformation phase as ``ds_stack``.  Returns
    (complex64 (nl, w, n) tensor, near-singular bool (nl, w) tensor).
    """
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    lv = np.linspace(g0_range[0], g0_range[1], levels)
    Ls = [np.linalg.cholesky(coherence_model(n, float(a), tau)) for a in lv]
    ns_lv = np.linspace(0.98, 0.999, 16)
    for a in ns_lv:
        G = np.full((n, n), a)
        np.fill_diagonal(G, 1.0)
        Ls.append(np.linalg.cholesky(G))
    L = torch.from_numpy(np.stack(Ls).astype(np.float32)).to(device)  # (levels + 16, n, n)
    t = torch.arange(n, device=device, dtype=torch.float32) * DT_DAYS
    out = torch.empty((nl, w, n), dtype=torch.complex64, device=device)
    ns_all = torch.empty((nl, w), dtype=torch.bool, device=device)
    for r0 in range(0, nl, row_block):
        rows = min(row_block, nl - r0)
        m = rows * w
        ns = torch.rand(m, generator=g, device=device) < near_singular_frac
        li = torch.randint(0, levels, (m,), generator=g, device=device)
        li = torch.where(ns, levels + torch.randint(0, 16, (m,), generator=g, device=device), li)
        wr = torch.randn((m, n), generator=g, device=device) / math.sqrt(2.0)
        wi = torch.randn((m, n), generator=g, device=device) / math.sqrt(2.0)
        zr = torch.empty_like(wr)
        zi = torch.empty_like(wi)
        for lvl in torch.unique(li).tolist():  # one factor per level: z = L w for its pixels
            sel = (li == lvl).nonzero()[:, 0]
            Lt = L[lvl].T.contiguous()
            zr[sel] = wr[sel] @ Lt
            zi[sel] = wi[sel] @ Lt
        v = torch.rand(m, generator=g, device=device) * 10.0
        ph = (4 * math.pi / WAVELENGTH_M) * (v / 1000.0 / 365.25)[:, None] * t
        c, s = torch.cos(ph), torch.sin(ph)
        out[r0:r0 + rows] = torch.complex(zr * c - zi * s, zr * s + zi * c).reshape(rows, w, n)
        ns_all[r0:r0 + rows] = ns.reshape(rows, w)
    return out, ns_all
