"""Temporal network helpers (image-pair layout of compressed coherence).

Mirrors moraine/tnet.py: ``TempNet`` (:52-120, without the zarr save/load) and
``nimage_from_npair`` (:123-130).  Pairs are all (i, j), i < j, in
``np.triu_indices`` order, optionally limited to j - i <= bandwidth (:15-20).
"""
from __future__ import annotations

import math

import numpy as np


def _imagepair_from_bandwidth(nimages: int, bandwidth: int | None = None) -> np.ndarray:
    if bandwidth is None:
        bandwidth = nimages
    assert nimages >= bandwidth
    ref, sec = np.triu_indices(nimages, 1)
    keep = (sec - ref) <= bandwidth
    return np.stack((ref[keep], sec[keep]), axis=-1).astype(np.int32)


def _edges_sorted(e: np.ndarray) -> bool:
    key = e[:, 0].astype(np.int64) * (int(e.max()) + 1) + e[:, 1]
    return bool(np.all(np.diff(key) >= 0))


def _edges_directed(e: np.ndarray) -> bool:
    return bool(np.all(e[:, 0] < e[:, 1]))


def _edges_connected(e: np.ndarray) -> bool:
    # every new reference image must first appear paired with its successor
    first = np.ones(e.shape[0], bool)
    first[1:] = e[1:, 0] != e[:-1, 0]
    return bool(np.all(e[first, 0] == e[first, 1] - 1))


class TempNet:
    def __init__(self, image_pairs, check_if_valid: bool = True):
        image_pairs = np.asarray(image_pairs)
        if check_if_valid:
            if not _edges_sorted(image_pairs):
                raise ValueError("input image pairs are not sorted.")
            if not _edges_directed(image_pairs):
                raise ValueError("input image pairs are not directed (reference index larger than or equal "
                                 "to secondary index).")
            if not _edges_connected(image_pairs):
                raise ValueError("input image pairs are not connected.")
        self._image_pairs = image_pairs.astype(np.int32)
        self._n_images = int(self._image_pairs[-1, -1]) + 1
        self._image_pairs_1d = self._image_pairs[:, 0].astype(np.int64) * self._n_images + self._image_pairs[:, 1]

    @classmethod
    def from_bandwidth(cls, nimages: int, bandwidth: int | None = None) -> "TempNet":
        return cls(_imagepair_from_bandwidth(nimages, bandwidth), check_if_valid=False)

    @property
    def image_pairs(self) -> np.ndarray:
        return self._image_pairs

    def image_pairs_idx(self, image_pairs=None, ref=None, sec=None):
        single = False
        if image_pairs is not None:
            ref = image_pairs[:, 0]
            sec = image_pairs[:, 1]
        else:
            if isinstance(ref, int):
                assert isinstance(sec, int)
                single = True
                ref, sec = np.array([ref]), np.array([sec])
            else:
                ref, sec = np.array(ref), np.array(sec)
            assert len(ref) == len(sec)
        q = np.asarray(ref, np.int64) * self._n_images + np.asarray(sec, np.int64)
        miss = np.isin(q, self._image_pairs_1d, assume_unique=True, invert=True)
        out = np.searchsorted(self._image_pairs_1d, q)
        out[miss] = -1
        return out[0] if single else out


def nimage_from_npair(npair: int) -> int:
    """Number of images n with n(n-1)/2 == npair (raises ValueError otherwise).

    n is the positive root of n^2 - n - 2*npair = 0, (1 + sqrt(1 + 8*npair)) / 2,
    taken in exact integer arithmetic.  As in the reference, 0 pairs gives 0.
    """
    npair = int(npair)
    if npair == 0:
        return 0
    root = math.isqrt(1 + 8 * npair) if npair > 0 else -1
    n = (1 + root) // 2
    if root < 0 or n * (n - 1) != 2 * npair:
        raise ValueError("input is not correct value of number of all pairs")
    return n
