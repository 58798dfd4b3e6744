"""Template-only cluster-expansion tables (reference: neuralce/ce/ce_gpu_implicit.py).

Nothing proportional to the number of sites is built: the primitive orbit list is a template of cluster instances
(member basis, integer primitive-cell offset), and the member sites around any site are found by integer arithmetic
k' = (k + U off) mod d on its Smith-normal-form cell coordinate k (ce_gpu_implicit.py:130-139,166-170).

Energy (ce_gpu_implicit.py:147-163):  E = zerolet + sum_cells sum_rows eci_row * prod_m pf[sl_m, bf_m, sigma(member m)]
with eci_row = eci_orbit * energy_scale / (n_instances_orbit * n_cells).  Here every row carries the cell-independent
coefficient c_row = eci_orbit / n_instances_orbit and the job factor f = energy_scale / n_cells multiplies the sum;
instances are the translation-distinct clusters of each orbit.

Tables built here, per basis site b:
  reverse entries  every (row, member) with member basis b, members written relative to that member
                   (ce_gpu_implicit.py:96-107): Rv_b (nB, J, mo), Rv_off (nB, J, mo, 3), Rv_bf, Rv_coef (nB, J)
  window           the distinct (basis, offset) member positions over all reverse entries of b, slot 0 = the site
                   itself; entries refer to window slots (ent_slot, uint16; padding members -> slot ONE whose point
                   function value is 1).  The GPU kernel gathers each window once per proposal instead of once per
                   member.
"""
from __future__ import annotations

import numpy as np

from ..lattice import ATOMIC_NUMBERS, CHEMICAL_SYMBOLS, CanonicalCell

SLOT_ONE = 0xFFFF          # padding member: point-function value 1 (the kernel maps it to its constant slot)


class CETemplate:
    """Cell-independent template of a cluster space (orbit rows, reverse entries, windows)."""

    def __init__(self, cs, parameters):
        prim = cs.primitive_structure
        nB = len(prim)
        self.nB = nB
        self.parameters = np.asarray(parameters, np.float64)
        self.basis_to_sl = np.asarray(cs.basis_to_sl, np.int32)
        self.M_basis = np.asarray(cs.M_basis, np.int32)
        self.active_b = np.where(self.M_basis > 1)[0].astype(np.int32)
        # binary point functions pf[sl, bf, species]; bf 0 == 1 (padding), bf 1 = theta_0 (reference: pf[0,1,:] = -1 Ni, +1 Pt)
        n_sl = int(self.basis_to_sl.max()) + 1
        self.pf = np.ones((n_sl, 2, 2), np.float64)
        for sl in range(n_sl):
            self.pf[sl, 1, :] = (-1.0, 1.0)
        tb, to, tbf, tcoef, torb = [], [], [], [], []
        for oi, o in enumerate(cs.orbits):
            c = float(self.parameters[oi + 1]) / len(o.clusters)
            for inst in o.clusters:
                tb.append([s[0] for s in inst]); to.append([s[1] for s in inst])
                tbf.append([1] * len(inst)); tcoef.append(c); torb.append(oi)
        Rt = len(tb); mo = max(len(r) for r in tb)
        self.Rt, self.mo = Rt, mo
        T_b = np.zeros((Rt, mo), np.int32); T_off = np.zeros((Rt, mo, 3), np.int64)
        T_bf = np.zeros((Rt, mo), np.int32); T_msk = np.zeros((Rt, mo), bool)
        for r in range(Rt):
            n = len(tb[r])
            T_b[r, :n] = tb[r]; T_off[r, :n] = to[r]; T_bf[r, :n] = tbf[r]; T_msk[r, :n] = True
            T_b[r, n:] = tb[r][0]; T_off[r, n:] = to[r][0]                     # padding: anchor member, bf 0 (pf 1)
        self.T_b, self.T_off, self.T_bf, self.T_msk = T_b, T_off, T_bf, T_msk
        self.T_coef = np.asarray(tcoef, np.float64); self.T_orbit = np.asarray(torb, np.int32)
        self.T_order = T_msk.sum(1).astype(np.int32)
        # reverse entries
        ent = [[] for _ in range(nB)]
        for r in range(Rt):
            for m in range(mo):
                if T_msk[r, m]:
                    ent[T_b[r, m]].append((r, m))
        J = max(max(len(e) for e in ent), 1)
        self.J = J
        self.J_b = np.array([len(e) for e in ent], np.int32)
        Rv_b = np.zeros((nB, J, mo), np.int32); Rv_off = np.zeros((nB, J, mo, 3), np.int64)
        Rv_bf = np.zeros((nB, J, mo), np.int32); Rv_coef = np.zeros((nB, J), np.float64)
        Rv_msk = np.zeros((nB, J, mo), bool); Rv_row = np.full((nB, J), -1, np.int32)
        for b in range(nB):
            Rv_b[b] = b
            for j, (r, m) in enumerate(ent[b]):
                Rv_b[b, j] = T_b[r]; Rv_off[b, j] = T_off[r] - T_off[r, m]; Rv_bf[b, j] = T_bf[r]
                Rv_coef[b, j] = self.T_coef[r]; Rv_msk[b, j] = T_msk[r]; Rv_row[b, j] = r
        self.Rv_b, self.Rv_off, self.Rv_bf, self.Rv_coef, self.Rv_msk, self.Rv_row = Rv_b, Rv_off, Rv_bf, Rv_coef, Rv_msk, Rv_row
        # windows: distinct member positions per basis, slot 0 = the site itself
        W = []
        ent_slot = np.full((nB, J, 4), SLOT_ONE, np.uint16)
        if mo > 4:
            raise NotImplementedError("clusters of more than 4 sites are not supported by the swap kernel")
        for b in range(nB):
            slots = {(b, (0, 0, 0)): 0}
            for j in range(len(ent[b])):
                for m in range(mo):
                    if not Rv_msk[b, j, m]:
                        continue
                    key = (int(Rv_b[b, j, m]), tuple(int(x) for x in Rv_off[b, j, m]))
                    if key not in slots:
                        slots[key] = len(slots)
                    ent_slot[b, j, m] = slots[key]
            # order slots by distance from the site (slot 0 = the site): compact clusters then use the first slots,
            # which the swap kernel's bit-packed windows rely on (every triplet / quadruplet in the first 64 slots)
            fb = prim.get_scaled_positions()
            x0 = fb[b] @ prim.cell
            dist = {key: float(np.linalg.norm((fb[key[0]] + np.asarray(key[1])) @ prim.cell - x0)) for key in slots}
            order = sorted(slots, key=lambda key: (round(dist[key], 6), key[0], key[1]))
            remap = {slots[key]: i for i, key in enumerate(order)}
            for j in range(len(ent[b])):
                for m in range(mo):
                    if ent_slot[b, j, m] != SLOT_ONE:
                        ent_slot[b, j, m] = remap[int(ent_slot[b, j, m])]
            W.append([(key, i) for i, key in enumerate(order)])
        self.W_b = np.array([len(w) for w in W], np.int32)
        Wmax = int(self.W_b.max())
        if Wmax >= SLOT_ONE:
            raise NotImplementedError("window too large")
        self.Wmax = Wmax
        self.win_b = np.zeros((nB, Wmax), np.int32); self.win_off = np.zeros((nB, Wmax, 3), np.int64)
        for b in range(nB):
            for (bb, off), s in W[b]:
                self.win_b[b, s] = bb; self.win_off[b, s] = off
        self.ent_slot = ent_slot
        self.ent_coef = Rv_coef.copy()
        # proposal groups: active sublattices, their basis sites, selection probability proportional to size
        act_sl = sorted(set(int(self.basis_to_sl[b]) for b in self.active_b))
        maxb = max((int((self.basis_to_sl[self.active_b] == sl).sum()) for sl in act_sl), default=1)
        self.act_sl = act_sl
        self.act_b = np.zeros((len(act_sl), maxb), np.int32); self.act_nb = np.zeros(len(act_sl), np.int32)
        for a, sl in enumerate(act_sl):
            bs = [int(b) for b in self.active_b if self.basis_to_sl[b] == sl]
            self.act_b[a, :len(bs)] = bs; self.act_nb[a] = len(bs)
        self.act_p = (self.act_nb / max(self.act_nb.sum(), 1)).astype(np.float64)


def build_ce_gpu_implicit(ce, supercell, energy_scale=1.0, template=None):
    """Tables for ``supercell`` (an integer supercell of the cluster space's primitive), reference
    ce_gpu_implicit.py:50-126.  Returns a dict with the template, SNF (U, d), perm / inv_perm, reduced window
    offsets, forward / reverse offsets reduced mod d (for the oracle), proposal groups and the energy factor."""
    cs = ce._cluster_space
    tpl = template or CETemplate(cs, ce.parameters)
    cc = CanonicalCell(cs.primitive_structure, supercell)
    U, d = cc.U, cc.d
    nB = tpl.nB; N = cc.N; n_cells = cc.n_cells
    uo = lambda off: np.mod(np.asarray(off, np.int64) @ U.T, d).astype(np.int32)
    win = np.zeros((nB, tpl.Wmax, 4), np.int32)
    win[:, :, 0] = tpl.win_b
    win[:, :, 1:] = uo(tpl.win_off)
    site_to_sl = tpl.basis_to_sl[np.arange(N) % nB]
    sl_species = {}
    for b in range(nB):
        sl_species.setdefault(int(tpl.basis_to_sl[b]), sorted(cs.chemical_symbols[b], key=lambda s: ATOMIC_NUMBERS[s]))
    return dict(template=tpl, n_atoms=N, n_cells=n_cells, n_basis=nB, U=U, d=np.asarray(d, np.int64), T=cc.T,
                perm=cc.perm, inv_perm=cc.inv_perm, basis_orig=cc.basis_orig, site_to_sl=site_to_sl,
                basis_to_sl=tpl.basis_to_sl, sl_species=sl_species, act_sl=tpl.act_sl,
                zerolet=float(tpl.parameters[0]) * float(energy_scale), energy_scale=float(energy_scale),
                factor=float(energy_scale) / n_cells, win=win,
                T_uo=uo(tpl.T_off), Rv_uo=uo(tpl.Rv_off), n_rows_template=tpl.Rt, max_size=tpl.mo, J=tpl.J,
                n_instances=tpl.Rt * n_cells, cell=cc)


def encode_canonical(info, cs, atoms):
    """species index per site in canonical order (sigma_canon = sigma_orig[perm]); reference
    ce_gpu_implicit.py:214-228 (species index = rank of the atomic number among the sublattice's species)."""
    Z = np.asarray(atoms.numbers); N = len(Z)
    sigma = np.zeros(N, np.int32)
    site_sl_orig = info['site_to_sl'][info['inv_perm']]
    for sl, species in info['sl_species'].items():
        m = site_sl_orig == sl
        if not m.any():
            continue
        lut = {ATOMIC_NUMBERS[s]: i for i, s in enumerate(species)}
        u, inv = np.unique(Z[m], return_inverse=True)
        bad = [CHEMICAL_SYMBOLS[z] for z in u if z not in lut]
        if bad:
            raise ValueError(f"sublattice {sl}: species {bad} not in {species}")
        sigma[m] = np.array([lut[z] for z in u], np.int32)[inv]
    return sigma[info['perm']]
