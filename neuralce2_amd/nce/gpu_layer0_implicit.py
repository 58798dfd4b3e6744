"""Cell-size-independent NeuralCE geometry (reference: neuralce/nce/gpu_layer0_implicit.py:17-118).

Per (single) basis site, on the Smith-normal-form lattice of the supercell:
  window    the sites sharing a cluster with the site (slot 0 = the site itself, ordered by distance)
  entries   every (cluster instance, member) with that member at the site, as window slots, with its layer-0 channel
            (1 + orbit) and coefficient nA / (n_instances * n_cells * order): h0[i, col] = sum over the entries of
            site i of coef * prod of the members' point functions (gpu_layer0_implicit.py:151-171)
  shells    partner offsets at distance in (edges[p], edges[p+1]] (np.digitize(..., right=True) bins, as
            gpu_layer0_implicit.py:79-90), periodic images kept as separate entries
  balls     receptive-field sizes: ball_0 = window, ball_{l+1} = ball_l dilated by the partner offsets
            (gpu_layer0_implicit.py:91-107); the kernels build the union of the two flipped sites' balls per step.
"""
from __future__ import annotations

import numpy as np

from ..ce.ce_gpu_implicit import CETemplate, SLOT_ONE
from ..lattice import ATOMIC_NUMBERS, CanonicalCell


def build_geometry(cs, atoms, shell_edges, depth):
    prim = cs.primitive_structure
    if len(prim) != 1:
        raise NotImplementedError("the GPU NeuralCE path supports one-basis lattices")
    if max(cs.M_basis) != 2:
        raise NotImplementedError("the GPU NeuralCE path supports binary species")
    cc = CanonicalCell(prim, atoms)
    tpl = CETemplate(cs, np.zeros(len(cs)))
    N = cc.N; n_cells = cc.n_cells; nA = N
    uo = lambda off: np.mod(np.asarray(off, np.int64) @ cc.U.T, cc.d).astype(np.int32)
    W0 = int(tpl.W_b[0])
    window = np.zeros((W0, 4), np.int32)
    window[:, 1:] = uo(tpl.win_off[0, :W0])
    J0 = int(tpl.J_b[0])
    ent_slot = np.ascontiguousarray(tpl.ent_slot[0, :J0])
    rows = tpl.Rv_row[0, :J0]
    orbit = tpl.T_orbit[rows]
    order = tpl.T_order[rows]
    n_inst = np.array([len(o.clusters) for o in cs.orbits])
    ent_col = (1 + orbit).astype(np.int32)
    ent_coef = nA / (n_inst[orbit] * n_cells * order.astype(float))
    # shells
    P = len(shell_edges) - 1
    cell = prim.cell
    rmax = float(shell_edges[-1]) + 1e-6
    L = int(np.ceil(rmax / min(np.linalg.norm(cell, axis=1)))) + 2
    rng = np.arange(-L, L + 1)
    grid = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    dist = np.linalg.norm(grid @ cell, axis=1)
    keep = (dist > 1e-6) & (dist <= rmax)
    grid, dist = grid[keep], dist[keep]
    shell = np.digitize(dist, np.asarray(shell_edges[1:-1], float), right=True)
    part_off = []; counts = []
    for p in range(P):
        o = grid[shell == p]
        o = o[np.lexsort((o[:, 2], o[:, 1], o[:, 0]))]
        part_off.append(o); counts.append(len(o))
    part_off = np.concatenate(part_off) if part_off else np.zeros((0, 3), np.int64)
    partners = np.zeros((len(part_off), 4), np.int32)
    partners[:, 1:] = uo(part_off)
    # ball sizes (lattice offsets; an upper bound of the distinct sites of one flip)
    ball = {tuple(o) for o in tpl.win_off[0, :W0].tolist()}
    sizes = [len(ball)]
    for _ in range(depth):
        ball = ball | {(a[0] + b[0], a[1] + b[1], a[2] + b[2]) for a in ball for b in part_off.tolist()}
        sizes.append(len(ball))
    sig_orig = np.zeros(N, np.int32)
    spec = sorted(cs.chemical_symbols[0], key=lambda x: ATOMIC_NUMBERS[x])
    lut = {ATOMIC_NUMBERS[s]: k for k, s in enumerate(spec)}
    for i, z in enumerate(atoms.numbers):
        if int(z) not in lut:
            raise ValueError(f"species {z} not in {spec}")
        sig_orig[i] = lut[int(z)]
    return dict(cell=cc, d=cc.d, n_cells=n_cells, N=N, nA=nA, window=window, ent_slot=ent_slot, ent_col=ent_col,
                ent_coef=ent_coef, shell_count=np.array(counts, np.int32), partners=partners, ball_sizes=sizes,
                ball_cap=[2 * s for s in sizes], K=len(cs), sig=sig_orig[cc.perm], perm=cc.perm, inv_perm=cc.inv_perm,
                species=spec, depth=depth)
