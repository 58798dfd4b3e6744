"""Philox4x32-10 on numpy arrays (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11): the host side of the counter-based streams the kernels draw from, for host decisions (the NeuralCE
exchange) that must be reproducible from a counter alone.  Same rounds and constants as the device function in
csrc/neuralce_hip.hip; the two are compared in the tests."""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_LO = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32-valued array, key: two uint32 -> (..., 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint64) & _LO
    c0, c1, c2, c3 = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        c0, c1, c2, c3 = n0, p1 & _LO, n2, p0 & _LO
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return np.stack([c0, c1, c2, c3], -1).astype(np.uint32)


def uniform53(w):
    """(..., 4) Philox output -> uniform in (0, 1) from the top 53 bits of its first two words."""
    x = (w[..., 0].astype(np.uint64) | (w[..., 1].astype(np.uint64) << np.uint64(32))) >> np.uint64(11)
    return (x.astype(np.float64) + 0.5) * 2.0 ** -53
