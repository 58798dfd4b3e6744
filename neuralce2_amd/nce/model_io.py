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


def ensemble_params(mdl):
    """The seed ensemble (reference mc_model.py:175-190, ``seed_idx < 0``: energy = mean over the saved parameter
    sets) as ONE network of width K D, for K seeds of width D.

    Every operation of a layer is either a linear map or elementwise per output channel (activation, gate, shell
    mean), so K networks side by side are one network whose first layer concatenates the K maps along the output
    channels and whose deeper layers are block-diagonal.  The readout ``h0 . lin + h_L . w + b`` is linear in the
    per-seed terms, so the mean over seeds takes lin = mean(lin_k), w = concat(w_k) / K and b = mean(b_k).  The
    result equals the mean of the K per-seed energies and energy changes exactly (up to rounding)."""
    sets = [expand_params(mdl, s) for s in range(len(mdl["params"]))]
    K = len(sets)
    Ps = [s[0] for s in sets]
    mu, sd = sets[0][1], sets[0][2]
    depth = len(Ps[0]["layers"])
    layers = []
    for l in range(depth):
        Ls = [P["layers"][l] for P in Ps]
        if any(set(L) != set(Ls[0]) for L in Ls):
            raise ValueError("the saved parameter sets differ in their layer structure; no ensemble")
        out = {}
        for key in Ls[0]:
            parts = [np.asarray(L[key], float) for L in Ls]
            if key in ("b", "gb"):
                out[key] = np.concatenate(parts)
            elif l == 0:                               # (C0, D) or (P, C0, D): concatenate along D
                out[key] = np.concatenate(parts, axis=-1)
            else:                                      # (D, D) or (P, D, D): block-diagonal over the seeds
                cin, d = parts[0].shape[-2], parts[0].shape[-1]
                lead = parts[0].shape[:-2]
                M = np.zeros(lead + (K * cin, K * d))
                for k, A in enumerate(parts):
                    M[..., k * cin:(k + 1) * cin, k * d:(k + 1) * d] = A
                out[key] = M
        layers.append(out)
    lin = np.mean([P["lin"] for P in Ps], axis=0)
    head = dict(b=np.asarray(np.mean([float(P["head"]["b"]) for P in Ps])))
    if depth:
        head["w"] = np.concatenate([np.asarray(P["head"]["w"], float) for P in Ps]) / K
    return dict(lin=lin, layers=layers, head=head), mu, sd


def check_supported(P):
    """Refuse parameter sets the gfx950 kernels do not evaluate, rather than run them with a term missing: the
    max-over-shell branch (``shell_max``, reference local_update_implicit.py:49-51,69-73) and a gate present in some
    layers but not in all."""
    layers = P["layers"]
    for l, L in enumerate(layers):
        if "shell_max" in L:
            raise NotImplementedError(f"layer {l} has a max-over-shell branch (shell_max); the gfx950 local update "
                                      f"evaluates mean-over-shell layers only")
    gated = ["gate" in L for L in layers]
    if any(gated) and not all(gated):
        raise NotImplementedError(f"gate maps in layers {[l for l, g in enumerate(gated) if g]} only; the kernels "
                                  f"take a gate in every layer or in none")
    for l, L in enumerate(layers):
        if ("gate" in L) != ("gb" in L):
            raise ValueError(f"layer {l}: gate maps and gate bias must come together")
