"""FP64 NumPy oracle of the NeuralCE (neural cluster expansion) energy.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.

Restates, independently of the product's canonical-order / template code (clusters and shell partners are placed by
CARTESIAN position in the supercell):
  h0_raw        layer 0: h0[i, 1 + k] = nA / (n_instances_k * n_cells) * sum over clusters of orbit k that contain
                site i (once per member equal to i) of prod(point functions) / order_k;  h0[i, 0] = 1
                (neuralce/nce/gpu_layer0_implicit.py:151-171, neuralce/nce/general.py:26-43)
  shells        partners of every site per distance shell (edges (lo, hi]), periodic images kept as separate
                entries (neuralce/nce/gpu_layer0_implicit.py:79-90, neuralce/nce/general.py:77-89)
  forward       reference layer: z_p = mean over shell p of h, pre = h W_self + sum_p z_p W_p + b,
                out = tanh(pre) * sigmoid(sum_p z_p G_p + gb) (neuralce/nce/local_update_implicit.py:43-54,
                neuralce/nce/general.py:213-235); the SiLU variant out = silu(pre) for the 12-neighbour network
  energy        E_std = sum_i e_i / nA + b with e_i = h0s_i . lin + h_L,i . w (local_update_implicit.py:110-112),
                eV/atom = E_std * ys + ym (mc_model.py:217-219); the summed readout variant E = sum_i e_i + b
"""
from __future__ import annotations

import numpy as np


_GRID = 1 << 20          # fractional coordinates of the supercell quantised to 2^-20 for the site key


def _keys(f):
    g = np.mod(np.rint(f * _GRID).astype(np.int64), _GRID)
    return (g[:, 0] * _GRID + g[:, 1]) * _GRID + g[:, 2], np.abs(f * _GRID - np.rint(f * _GRID)).max(-1)


def _site_lookup(atoms):
    """xyz (M, 3) -> index of the supercell site at that position, modulo the supercell: sites are keyed by their
    wrapped fractional coordinates on a 2^-20 grid and found by binary search (O(M log N))."""
    Sinv = np.linalg.inv(atoms.cell)
    k_site, _ = _keys(np.mod(atoms.positions @ Sinv, 1.0))
    order = np.argsort(k_site, kind="stable")
    k_sorted = k_site[order]
    assert len(np.unique(k_sorted)) == len(k_sorted), "two sites share a position"

    def site_at(xyz):
        k, off = _keys(np.mod(np.asarray(xyz, float).reshape(-1, 3) @ Sinv, 1.0))
        pos = np.minimum(np.searchsorted(k_sorted, k), len(k_sorted) - 1)
        assert np.all(k_sorted[pos] == k) and np.all(off < 20.0), "position not on the supercell lattice"
        return order[pos]
    return site_at


def _translations(prim, atoms):
    P = prim.cell
    T = np.round(atoms.cell @ np.linalg.inv(P)).astype(np.int64)
    n_cells = int(round(abs(np.linalg.det(T))))
    Sinv = np.linalg.inv(atoms.cell)
    L = int(np.abs(T).sum()) + 1
    rng = np.arange(-L, L + 1)
    grid = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    fr = (grid @ P) @ Sinv
    inside = np.all((fr > -1e-9) & (fr < 1 - 1e-9), axis=1)
    trans = grid[inside]
    assert len(trans) == n_cells
    return trans, n_cells


def h0_raw(cs, atoms, spins):
    """(N, 1 + n_orbits) raw layer 0 of a binary cluster space; spins = -1 (species 0) / +1 (species 1) per atom."""
    prim = cs.primitive_structure
    fb = prim.get_scaled_positions()
    trans, n_cells = _translations(prim, atoms)
    site_at = _site_lookup(atoms)
    N = len(atoms)
    nA = N                                              # every site active (binary, one sublattice)
    h0 = np.zeros((N, 1 + len(cs.orbits)))
    h0[:, 0] = 1.0
    for k, o in enumerate(cs.orbits):
        acc = np.zeros(N)
        for inst in o.clusters:
            sites = [site_at((fb[b] + np.asarray(n) + trans) @ prim.cell) for b, n in inst]
            prod = np.ones(len(trans))
            for s in sites:
                prod = prod * spins[s]
            for s in sites:
                np.add.at(acc, s, prod / o.order)
        h0[:, 1 + k] = acc * nA / (len(o.clusters) * n_cells)
    return h0


def shells(prim, atoms, edges):
    """partners[p] (N, W_p) site indices of shell p: lattice points at distance in (edges[p], edges[p+1]]."""
    site_at = _site_lookup(atoms)
    P = prim.cell
    fb = prim.get_scaled_positions()
    rmax = float(edges[-1]) + 1e-6
    L = int(np.ceil(rmax / min(np.linalg.norm(P, axis=1)))) + 2
    rng = np.arange(-L, L + 1)
    grid = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    assert len(prim) == 1, "one-basis lattices"
    vec = grid @ P
    d = np.linalg.norm(vec, axis=1)
    keep = (d > 1e-6) & (d <= rmax)
    vec, d = vec[keep], d[keep]
    shell = np.digitize(d, np.asarray(edges[1:-1]), right=True)
    out = []
    for p in range(len(edges) - 1):
        offs = vec[shell == p]
        part = np.stack([site_at(atoms.positions + o) for o in offs], axis=1)   # (N, W_p)
        out.append(part)
    return out


def layer(L, h, partners, kind="gate"):
    zs = [h[part].mean(axis=1) for part in partners]                    # (N, C) per shell
    pre = h @ L["self"] + sum(z @ L["shell"][p] for p, z in enumerate(zs)) + L["b"]
    if kind == "silu":
        return pre / (1.0 + np.exp(-pre))
    out = np.tanh(pre)
    if "gate" in L:
        g = sum(z @ L["gate"][p] for p, z in enumerate(zs)) + L["gb"]
        out = out / (1.0 + np.exp(-g))
    return out


def site_energies(P, h0s, partners, kind="gate"):
    h = h0s
    for Lr in P["layers"]:
        h = layer(Lr, h, partners, kind)
    e = h0s @ P["lin"]
    if P["layers"]:
        e = e + h @ P["head"]["w"]
    return e


def energy_std(P, h0s, partners, kind="gate", readout="mean"):
    e = site_energies(P, h0s, partners, kind)
    tot = e.sum() / (len(e) if readout == "mean" else 1.0)
    return float(tot + float(P["head"]["b"]))


def standardize(h0, mu, sd):
    return (h0 - mu) / sd


class Oracle:
    """Precomputed cluster site lists and shell partners of one cell (placed by Cartesian position, as h0_raw and
    shells above), so that E(sigma) is a vectorised FP64 full recompute."""

    def __init__(self, cs, atoms, shell_edges, P, mu, sd, kind="gate", readout="mean"):
        prim = cs.primitive_structure
        fb = prim.get_scaled_positions()
        trans, n_cells = _translations(prim, atoms)
        site_at = _site_lookup(atoms)
        self.N = len(atoms)
        self.orbits = []
        for o in cs.orbits:
            lists = []
            for inst in o.clusters:
                lists.append(np.stack([site_at((fb[b] + np.asarray(n) + trans) @ prim.cell) for b, n in inst], 1))
            sites = np.concatenate(lists, 0)                   # (n_inst * n_cells, order)
            self.orbits.append((sites, o.order, self.N / (len(o.clusters) * n_cells)))
        self.partners = shells(prim, atoms, shell_edges)
        self.P, self.mu, self.sd, self.kind, self.readout = P, mu, sd, kind, readout

    def h0(self, spins):
        h = np.zeros((self.N, 1 + len(self.orbits)))
        h[:, 0] = 1.0
        for k, (sites, order, scale) in enumerate(self.orbits):
            prod = np.prod(spins[sites], axis=1) / order
            h[:, 1 + k] = np.bincount(sites.ravel(), weights=np.repeat(prod, sites.shape[1]), minlength=self.N) * scale
        return h

    def h0s(self, sig):
        spins = np.where(np.asarray(sig) == 0, -1.0, 1.0)
        return standardize(self.h0(spins), self.mu, self.sd)

    def energy(self, sig):
        """model units (E_std for the reference readout)."""
        return energy_std(self.P, self.h0s(sig), self.partners, self.kind, self.readout)

    def layers(self, sig):
        h = self.h0s(sig)
        out = [h]
        for Lr in self.P["layers"]:
            h = layer(Lr, h, self.partners, self.kind)
            out.append(h)
        return out

    def run_chunk(self, occ, E, betas, inj, de_scale):
        """sequential swap Metropolis with injected proposals (mcmc.py:167-185): a same-species or same-site
        proposal is accepted with dE = 0; else dE from two full recomputes; accept dE_eV <= 0 or u < exp(-b dE_eV)."""
        ia, ib, u = inj
        R, n = ia.shape
        nacc = np.zeros(R, np.int64)
        for r in range(R):
            for t in range(n):
                a, b = int(ia[r, t]), int(ib[r, t])
                if a == b or occ[r, a] == occ[r, b]:
                    nacc[r] += 1
                    continue
                s1 = occ[r].copy(); s1[a], s1[b] = occ[r, b], occ[r, a]
                E1 = self.energy(s1)
                dE = E1 - E[r]
                dE_eV = dE * de_scale
                if dE_eV <= 0 or u[r, t] < np.exp(-betas[r] * dE_eV):
                    occ[r] = s1; E[r] = E1; nacc[r] += 1
        return nacc
