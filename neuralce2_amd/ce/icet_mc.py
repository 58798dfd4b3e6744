"""Cluster-expansion MC entry (reference neuralce/ce/icet_mc.py:76-235): ``.ce`` file + initial cif -> PT run.

``run_pt_mc`` reads the ``.ce`` (data-only), puts the cif on a clean integer supercell of the expansion's primitive
(species moved to the nearest ideal site), optionally tiles it ``supercell_reps`` times with the species reshuffled
within every sublattice from ``seed`` (``CE_SUPERCELL_PERIODIC=1`` keeps the periodic images), scales the expansion
by the number of tiles (the CE value is the total energy of the base cell), encodes the start, reads per-replica
starting structures from ``init_npz``, fingerprints the run (``run_id`` = md5 of the parameters and the start) and
calls ``run_pt_mcmc``.  ``run_mc`` / ``run_multiT_mc`` are the same machinery with one replica per temperature and no
exchange.

Not restated: icet's primitive detection for relaxed cifs with species mapping (``ideal_map`` / ``inert_species``,
map_structure_to_reference, spglib); a cif must sit on the expansion's lattice up to the nearest-site transfer below.
"""
from __future__ import annotations

import hashlib
import os
from pathlib import Path

import numpy as np

from ..ce_io import ClusterExpansion
from ..lattice import OffLatticeError, Structure, read_cif, site_decompose


def _rebuild_as_clean_supercell(atoms, primitive, cell_atol=0.1):
    """Ideal integer supercell P_int = round(S P^-1) of `primitive`, each site taking the species of the nearest
    (periodic) atom of `atoms`, one atom per site (icet_mc.py:20-73)."""
    P = atoms.cell @ np.linalg.inv(primitive.cell)
    P_int = np.round(P).astype(np.int64)
    residual = float(np.max(np.abs(P - P_int)))
    if residual > cell_atol:
        raise ValueError(f"cell deviates from an integer supercell of the expansion's primitive (max residual "
                         f"{residual:.4f} > {cell_atol})")
    cell = P_int @ primitive.cell
    n_cells = int(round(abs(np.linalg.det(P_int))))
    L = int(np.abs(P_int).sum()) + 1
    rng = np.arange(-L, L + 1)
    grid = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    fr = (grid @ primitive.cell) @ np.linalg.inv(cell)
    inside = np.all((fr > -1e-9) & (fr < 1 - 1e-9), axis=1)
    trans = grid[inside]
    if len(trans) != n_cells:
        raise ValueError("could not enumerate the supercell's primitive cells")
    pos = (trans[:, None, :] @ primitive.cell + primitive.positions[None, :, :]).reshape(-1, 3)
    ideal = Structure(cell, pos, np.zeros(len(pos), np.int64))
    if len(ideal) != len(atoms):
        raise ValueError(f"clean supercell has {len(ideal)} sites, the cif {len(atoms)} atoms")
    inv = np.linalg.inv(cell)
    used = np.zeros(len(atoms), bool)
    nums = np.zeros(len(ideal), np.int64)
    for i, p in enumerate(ideal.positions):
        f = (atoms.positions - p) @ inv
        f -= np.round(f)
        d2 = ((f @ cell) ** 2).sum(1)
        d2[used] = np.inf
        j = int(np.argmin(d2))
        used[j] = True
        nums[i] = atoms.numbers[j]
    ideal.numbers = nums
    return ideal


def _prepare_cell(ce, init_cif, supercell_reps, seed):
    prim = ce._cluster_space.primitive_structure
    atoms = read_cif(init_cif)
    energy_scale = 1.0
    if supercell_reps is not None and tuple(supercell_reps) != (1, 1, 1):
        atoms = _rebuild_as_clean_supercell(atoms, prim).repeat(tuple(int(r) for r in supercell_reps))
        energy_scale = float(np.prod([int(r) for r in supercell_reps]))
        if os.environ.get("CE_SUPERCELL_PERIODIC", "0") != "1":
            rng = np.random.default_rng(seed)
            Z = atoms.numbers.copy()
            basis, _, _, _, _ = site_decompose(prim, atoms)
            b2sl = np.asarray(ce._cluster_space.basis_to_sl)
            for sl in np.unique(b2sl):
                m = np.isin(basis, np.where(b2sl == sl)[0])
                Z[m] = rng.permutation(Z[m])
            atoms.numbers = Z
    else:
        try:
            site_decompose(prim, atoms)
        except OffLatticeError:
            atoms = _rebuild_as_clean_supercell(atoms, prim)
    return atoms, energy_scale


def run_pt_mc(ce_path, init_cif, n_steps, T_min=100.0, T_max=2000.0, n_replicas=32, n_ladders=1,
              swap_mode="uniform", seed=0, log_every=0, n_equilib=0, exchange_every=100, chunk_size=1000,
              sample_interval=100, out_dir=".", ideal_map=None, inert_species=None, symprec=1e-2, swap_pairs=None,
              ref_cif=None, fractional_position_tolerance=None, energy_mode="delta", spin_move_prob=0.0,
              spin_orbits_pkl="", spin_states=None, supercell_reps=None, table_mode=None, checkpoint_every=0,
              T_grid=None, init_npz=None, precision="mixed", tile=0):
    """PT run of a ``.ce`` cluster expansion (see the module docstring); returns the MCMCResult."""
    from .classical_ce_gpu import GPUClusterExpansion
    from .pt_mcmc_ce import MCMCConfig, run_pt_mcmc
    if swap_mode not in ("uniform", "weighted"):
        raise ValueError(f"swap_mode must be 'uniform' or 'weighted'; got {swap_mode!r}")
    if ideal_map or inert_species:
        raise NotImplementedError("ideal_map / inert_species need icet's structure mapping, not part of this path")
    if spin_move_prob or spin_orbits_pkl or spin_states:
        raise NotImplementedError("spin moves are not part of the canonical swap path")
    Path(out_dir).mkdir(parents=True, exist_ok=True)
    ce = ClusterExpansion.read(ce_path)
    atoms, energy_scale = _prepare_cell(ce, init_cif, supercell_reps, seed)
    gce = GPUClusterExpansion(ce, atoms, fractional_position_tolerance=fractional_position_tolerance,
                              table_mode=table_mode, energy_scale=energy_scale)
    initial_sigma = gce.encode(atoms)
    print(f"PT-MC: {len(atoms)} sites, energy_scale {energy_scale:g}, init E = {float(gce.energy(initial_sigma)):.4f} eV",
          flush=True)
    if init_npz:
        if n_ladders != 1:
            raise ValueError("--init_npz assigns one structure per temperature of a single ladder; n_ladders must be 1")
        from ..mc.init_structures import load_init_npz
        run_T = np.asarray(T_grid, float) if T_grid is not None else np.geomspace(T_min, T_max, n_replicas)
        nums, _ = load_init_npz(init_npz, ce._cluster_space.primitive_structure, atoms, run_T)
        a_r = atoms.copy()
        rows = []
        for r in range(nums.shape[0]):
            a_r.numbers = nums[r]
            rows.append(gce.encode(a_r))
        initial_sigma = np.stack(rows)
    h = hashlib.md5(np.ascontiguousarray(ce.parameters, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(initial_sigma, dtype=np.int32).tobytes())
    cfg = MCMCConfig(T_min=T_min, T_max=T_max, n_replicas=(len(T_grid) if T_grid is not None else n_replicas),
                     n_ladders=n_ladders, n_steps=n_steps, n_equilib=n_equilib, exchange_interval=exchange_every,
                     chunk_size=chunk_size, sample_interval=sample_interval, output_dir=out_dir, seed=seed,
                     swap_pairs=swap_pairs, initial_sigma=initial_sigma, energy_mode=energy_mode,
                     checkpoint_every=checkpoint_every,
                     T_grid=(np.asarray(T_grid, float) if T_grid is not None else None), run_id=h.hexdigest()[:16],
                     precision=precision, tile=tile)
    return run_pt_mcmc(gce, cfg)


def run_multiT_mc(ce_path, init_cif, temps, n_steps, seed=0, log_every=0, out_dir=".", energy_mode="delta",
                  chunk_size=1000, sample_interval=100, supercell_reps=None, precision="mixed", **kw):
    """independent chains, one per temperature (no exchange), on the PT machinery."""
    temps = np.asarray(temps, float).reshape(-1)
    chunk = min(int(chunk_size), int(n_steps))
    n_chunks = max(1, int(n_steps) // chunk)
    return run_pt_mc(ce_path, init_cif, n_chunks * chunk, n_replicas=len(temps), seed=seed, log_every=log_every,
                     out_dir=out_dir, energy_mode=energy_mode, chunk_size=chunk,
                     sample_interval=min(int(sample_interval), chunk), exchange_every=(n_chunks + 1) * chunk,
                     supercell_reps=supercell_reps, T_grid=temps, precision=precision)


def run_mc(ce_path, init_cif, n_steps, temperature=300.0, **kw):
    """one chain at one temperature."""
    return run_multiT_mc(ce_path, init_cif, [temperature], n_steps, **kw)
