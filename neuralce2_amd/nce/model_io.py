"""NeuralCE model files (models/nce/<tag>/model.pkl) without unpickling.

The reference reads them with ``pickle.load`` (neuralce/nce/mc_model.py:20-21); the stream names numpy reconstructors
and ``ase.atoms.Atoms`` / ``ase.cell.Cell``.  ``load_model`` decodes it with the data-only reader and turns the
primitive ``Atoms`` object's state into a ``Structure``; every array stays a numpy array.

``expand_params`` follows neuralce/nce/mc_model.py:28-57: the channel selection (``keep``) is folded into zero rows of
the full-width layer-0 weights, and dropped channels get mu = 0, sd = 1.
"""
from __future__ import annotations

import numpy as np

from .. import safe_pickle
from ..lattice import Structure


def _structure_from_obj(obj):
    st = obj.state
    cell = st["_cellobj"].state["array"] if hasattr(st["_cellobj"], "state") else st["_cellobj"]
    arr = st["arrays"]
    return Structure(np.asarray(cell, float), np.asarray(arr["positions"], float), np.asarray(arr["numbers"]))


def _plain(x):
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_plain(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_plain(v) for v in x)
    if isinstance(x, safe_pickle.Obj):
        q = x.cls.qualname if isinstance(x.cls, safe_pickle.Global) else str(x.cls)
        if q == "ase.atoms.Atoms":
            return _structure_from_obj(x)
        return x
    return x


def load_model(path):
    """model dict: cutoffs, shell_edges, params (one per seed), mu, sd, ym, ys, keep, K_chem, primitive, ..."""
    with open(path, "rb") as f:
        mdl = safe_pickle.loads(f.read())
    mdl = _plain(mdl)
    if not isinstance(mdl.get("primitive"), Structure):
        raise ValueError(f"{path}: no primitive structure in the model")
    return mdl


def n_full_channels(mdl):
    return int(mdl["K_chem"]) + (int(mdl["n_spin"]) if mdl.get("spin", 0) else 0)


def expand_params(mdl, seed_idx=0):
    """params on the full K_chem-channel layer 0 (zero rows for channels dropped by the selection), mu, sd."""
    keep = np.asarray(mdl["keep"])
    Kfull = n_full_channels(mdl)
    P = mdl["params"][seed_idx]
    lin = np.zeros(Kfull); lin[keep] = np.asarray(P["lin"], float)
    layers = []
    for L in P["layers"]:
        L = {k: np.asarray(v, float) for k, v in L.items()}
        if L["self"].shape[0] == len(keep):
            Lx = dict(L)
            Lx["self"] = np.zeros((Kfull, L["self"].shape[1])); Lx["self"][keep] = L["self"]
            for key in ("shell", "gate", "shell_max"):
                if key in L:
                    Lx[key] = np.zeros((L[key].shape[0], Kfull, L[key].shape[2])); Lx[key][:, keep] = L[key]
            layers.append(Lx)
        else:
            layers.append(L)
    mu = np.zeros(Kfull); sd = np.ones(Kfull)
    mu[keep] = np.asarray(mdl["mu"], float); sd[keep] = np.asarray(mdl["sd"], float)
    head = {k: np.asarray(v, float) for k, v in P["head"].items()}
    return dict(lin=lin, layers=layers, head=head), mu, sd
