"""``.ce`` cluster-expansion files (icet 3.2 format) without icet.

A ``.ce`` file is a tar with ``items`` (pickled dict: ``parameters`` float64 array, metadata) and ``cluster_space``
(a nested tar with ``items`` = cutoffs, chemical_symbols, symprec, ... and ``atoms`` = the primitive structure as an
ase-db JSON document).  Read by the reference through ``icet.ClusterExpansion.read`` (neuralce/ce/icet_mc.py:131).
The pickle streams are decoded with the data-only reader in ``safe_pickle``.
"""
from __future__ import annotations

import io
import json
import tarfile
import numpy as np

from . import safe_pickle
from .clusterspace import ClusterSpace
from .lattice import Structure


def _ase_json_atoms(txt):
    doc = json.loads(txt)
    key = str(doc["ids"][0]) if "ids" in doc else sorted(k for k in doc if k.isdigit())[0]
    row = doc[key]

    def arr(v):
        if isinstance(v, dict) and "__ndarray__" in v:
            shape, dt, flat = v["__ndarray__"]
            return np.array(flat, dtype=dt).reshape(shape)
        if isinstance(v, dict) and "array" in v:
            return arr(v["array"])
        return np.asarray(v)
    cell = arr(row["cell"]); pos = arr(row["positions"]); num = arr(row["numbers"])
    return Structure(cell, pos, num)


class ClusterExpansion:
    """Parameters + cluster space, the slice of ``icet.ClusterExpansion`` the MC path uses."""

    def __init__(self, cluster_space, parameters, metadata=None):
        self._cluster_space = cluster_space
        self.parameters = np.asarray(parameters, np.float64)
        self.metadata = metadata or {}
        if len(self.parameters) != len(cluster_space):
            raise ValueError(f"{len(self.parameters)} parameters for a cluster space of size {len(cluster_space)}")

    @property
    def cluster_space(self):
        return self._cluster_space

    @classmethod
    def read(cls, path):
        with tarfile.open(path) as t:
            items = safe_pickle.loads(t.extractfile("items").read())
            cs_tar = t.extractfile("cluster_space").read()
        with tarfile.open(fileobj=io.BytesIO(cs_tar)) as t2:
            cs_items = safe_pickle.loads(t2.extractfile("items").read())
            prim = _ase_json_atoms(t2.extractfile("atoms").read().decode())
        cs = ClusterSpace(prim, cs_items["cutoffs"], cs_items["chemical_symbols"])
        return cls(cs, items["parameters"], items.get("metadata"))

    def predict(self, atoms):
        """parameters . cluster_vector (the CE value of the structure)."""
        return float(self.parameters @ self._cluster_space.cluster_vector(atoms))
