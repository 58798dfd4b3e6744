"""Per-replica starting structures (--init_npz) for the PT drivers (reference neuralce/mc/init_structures.py:1-61).

The npz file (an interchange format; read with numpy, no pickles):
    numbers   (R, N) uint8   atomic numbers of replica r's starting structure
    positions (N, 3) float   Cartesian positions of the N sites (one geometry for all replicas)
    cell      (3, 3) float   cell of that geometry; must equal the MC cell
    temps     (R,)   float   temperature (K) of each structure; must equal the run's temperatures

Sites are matched onto the MC cell by lattice site, not by atom index: both geometries are put in the canonical
site order of the primitive's supercell (Smith-normal-form cell index and basis, lattice.CanonicalCell), which
identifies a lattice site whatever the atom order or periodic image a position is written in.
"""
from __future__ import annotations

import numpy as np

from ..lattice import CanonicalCell, OffLatticeError, Structure


def map_to_reference(prim, ref_atoms, cell, positions):
    """idx (N,): the npz site idx[i] is the lattice site of ref_atoms[i]."""
    cell = np.asarray(cell, float)
    positions = np.asarray(positions, float)
    if cell.shape != (3, 3) or not np.allclose(cell, ref_atoms.cell, rtol=0, atol=1e-3):
        raise ValueError(f"the init_npz cell {np.round(cell, 4).tolist()} is not the MC cell "
                         f"{np.round(ref_atoms.cell, 4).tolist()}")
    if positions.shape != (len(ref_atoms), 3):
        raise ValueError(f"init_npz lists {len(positions)} sites; the MC cell has {len(ref_atoms)}")
    ref = CanonicalCell(prim, ref_atoms)
    try:
        other = CanonicalCell(prim, Structure(ref_atoms.cell, positions, np.zeros(len(positions), int)))
    except OffLatticeError as e:
        raise ValueError(f"init_npz sites are not the lattice sites of the MC cell: {e}") from None
    if not np.array_equal(ref.T, other.T):
        raise ValueError(f"supercell matrices differ: {ref.T.tolist()} vs {other.T.tolist()}")
    # canonical index -> npz position of that site, then looked up for every reference site
    npz_of_canon = np.empty(len(positions), np.int64)
    npz_of_canon[other.index] = np.arange(len(positions))
    return npz_of_canon[ref.index]


def load_init_npz(path, prim, ref_atoms, temps, tol_T=0.5):
    """-> numbers (R, N) in the atom order of ref_atoms, and the npz temperatures.  Checked: the cell, the lattice
    sites, one structure per run temperature (|dT| <= max(tol_T, 1e-5 T)), and every replica's composition equal
    to ref_atoms'."""
    with np.load(path, allow_pickle=False) as z:
        numbers = np.asarray(z["numbers"])
        T_npz = np.asarray(z["temps"], float).reshape(-1)
        cell, positions = z["cell"], z["positions"]
    temps = np.asarray(temps, float).reshape(-1)
    if numbers.ndim != 2 or numbers.shape[0] != len(T_npz):
        raise ValueError(f"init_npz: numbers {numbers.shape} and temps {T_npz.shape} do not pair up")
    off = np.abs(temps - T_npz) if len(temps) == len(T_npz) else None
    if off is None or np.any(off > np.maximum(tol_T, 1e-5 * T_npz)):
        raise ValueError(f"init_npz holds structures for {np.round(T_npz, 1).tolist()} K; the run's temperatures "
                         f"are {np.round(temps, 1).tolist()} K")
    out = numbers[:, map_to_reference(prim, ref_atoms, cell, positions)]
    want = np.bincount(np.asarray(ref_atoms.numbers, np.int64), minlength=256)
    for r, row in enumerate(out):
        if not np.array_equal(np.bincount(row.astype(np.int64), minlength=256), want):
            raise ValueError(f"init_npz replica {r} has another composition than the MC cell")
    return out, T_npz
