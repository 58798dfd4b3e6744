"""Synthetic benchmark workloads (seeded): the PtNi-like fcc binary cluster expansion and cells.

fcc, a = 3.924 Angstrom, species Ni / Pt, cutoffs 8 / 6 / 5 Angstrom (pairs / triplets / quadruplets): 37 orbits + the
zerolet, the same orbit list as the reference's shipped icet_c8_6_5 model.  ECIs: J_orbit ~ N(0, (0.020 eV / r^2)^2)
with r the orbit's largest pair distance in Angstrom (r = 1 for the singlet), zerolet 0.  They are per-atom ECIs:
the total energy of an N-site cell is N * sum_orbit J_orbit Phi_orbit (energy_scale = N with parameters = J).
"""
from __future__ import annotations

import numpy as np

from .ce_io import ClusterExpansion
from .clusterspace import ClusterSpace
from .lattice import bulk_fcc

A_FCC = 3.924
_CS_CACHE = {}


def fcc_cluster_space(cutoffs=(8.0, 6.0, 5.0), a=A_FCC):
    key = (tuple(cutoffs), a)
    if key not in _CS_CACHE:
        _CS_CACHE[key] = ClusterSpace(bulk_fcc("Ni", a, cubic=False), list(cutoffs), [["Ni", "Pt"]])
    return _CS_CACHE[key]


def synthetic_ce(seed=0, cutoffs=(8.0, 6.0, 5.0)):
    cs = fcc_cluster_space(cutoffs)
    rng = np.random.default_rng(seed)
    params = np.zeros(len(cs))
    for i, o in enumerate(cs.orbits):
        r = max(o.distances) if o.distances else 1.0
        params[i + 1] = rng.normal(0.0, 0.020 / r ** 2)
    return ClusterExpansion(cs, params)


def fcc_cell(n, x=0.5, seed=0, a=A_FCC):
    """n x n x n conventional fcc cells, Pt fraction x placed at random (numpy Generator seeded with `seed`)."""
    at = bulk_fcc("Ni", a, cubic=True).repeat((n, n, n))
    rng = np.random.default_rng(seed)
    Z = np.full(len(at), 28, np.int64)
    Z[rng.choice(len(at), int(round(x * len(at))), replace=False)] = 78
    at.numbers = Z
    return at
