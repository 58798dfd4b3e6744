"""Cluster space of a primitive structure: symmetry, orbits, point functions, cluster vectors.

Replaces the parts of icet 3.2 (an external C++ dependency of the reference, not vendored) that the Monte Carlo
path reads: the primitive orbit list with every translation-distinct cluster of each orbit
(``o.clusters`` / ``lattice_sites`` at neuralce/ce/ce_gpu_implicit.py:76-86), the cluster-vector element order that
the ``.ce`` parameters follow, and the orthonormal point functions
theta_p(s) = -cos(2 pi k s / M) (p even), -sin(2 pi k s / M) (p odd), k = (p + 2) // 2
(neuralce/nce/gpu_layer0.py:32-35).

Clusters are included when every pair distance is <= the cutoff of their order (cutoffs[0] pairs, [1] triplets,
...).  Orbits are classes under the space group; they are ordered by (order, radius, sorted pair distances,
multiplicity), radius = mean distance of the sites to their centroid.  For the binary PtNi fcc spaces this order
reproduces the reference's shipped model energies (models/ce/ptni/fits.json), which the tests check.

Only binary sublattices (M = 2, one basis function) are supported for the cluster-vector element order:
multicomponent decorations need icet's site-permutation bookkeeping and are outside the configurations built here.
"""
from __future__ import annotations

import itertools
import numpy as np

from .lattice import ATOMIC_NUMBERS, Structure

_TOL = 1e-5


def theta(M, p, s):
    """orthonormal point function value (reference: neuralce/nce/gpu_layer0.py:32-35)."""
    k = (p + 2) // 2
    x = 2 * np.pi * k * s / M
    return -np.cos(x) if p % 2 == 0 else -np.sin(x)


def point_function_table(M):
    """pf[bf, species]: bf 0 is the constant 1, bf 1..M-1 the orthonormal functions theta_{bf-1}."""
    pf = np.ones((M, M))
    for p in range(M - 1):
        for s in range(M):
            pf[p + 1, s] = theta(M, p, s)
    return pf


def lattice_point_ops(prim, basis_species):
    """Space-group operations (R, t, site map) of the decorated primitive: f' = R f + t in fractional coordinates,
    R integer, metric preserving; a basis site may only map onto a site with the same allowed species."""
    P = prim.cell
    G = P @ P.T
    fb = prim.get_scaled_positions()
    nB = len(prim)
    ops = []
    for ent in itertools.product((-1, 0, 1), repeat=9):
        R = np.array(ent, np.int64).reshape(3, 3)
        if abs(round(np.linalg.det(R))) != 1:
            continue
        if not np.allclose(R.T @ G @ R, G, atol=1e-6):
            continue
        for b_img in range(nB):
            t = fb[b_img] - R @ fb[0]
            smap = []
            shifts = []
            ok = True
            for b in range(nB):
                f = R @ fb[b] + t
                hit = None
                for b2 in range(nB):
                    dd = f - fb[b2]
                    if np.allclose(dd, np.round(dd), atol=_TOL) and basis_species[b] == basis_species[b2]:
                        hit = (b2, np.round(dd).astype(np.int64)); break
                if hit is None:
                    ok = False; break
                smap.append(hit[0]); shifts.append(hit[1])
            if ok:
                ops.append((R, np.array(shifts), np.array(smap)))
    return ops


class Orbit:
    def __init__(self, order, clusters, radius, distances):
        self.order = order
        self.clusters = clusters          # translation-distinct instances: list of tuples ((b, (n0,n1,n2)), ...)
        self.radius = radius
        self.distances = distances
        self.multiplicity = len(clusters)

    def __repr__(self):
        return f"Orbit(order={self.order}, radius={self.radius:.4f}, multiplicity={self.multiplicity})"


class ClusterSpace:
    """Orbit list of ``prim`` for ``cutoffs`` (per order starting at pairs) and allowed species per basis site."""

    def __init__(self, prim, cutoffs, chemical_symbols):
        self.primitive_structure = prim
        self.cutoffs = [float(c) for c in cutoffs]
        if len(chemical_symbols) == 1 and len(prim) > 1:
            chemical_symbols = list(chemical_symbols) * len(prim)
        self.chemical_symbols = [list(s) for s in chemical_symbols]
        nB = len(prim)
        self.n_basis = nB
        spec = [tuple(sorted(s, key=lambda x: ATOMIC_NUMBERS[x])) for s in self.chemical_symbols]
        # sublattices: groups of basis sites with the same allowed species, ordered by first appearance
        self.sublattice_species = []
        self.basis_to_sl = np.zeros(nB, np.int32)
        for b in range(nB):
            if spec[b] not in self.sublattice_species:
                self.sublattice_species.append(spec[b])
            self.basis_to_sl[b] = self.sublattice_species.index(spec[b])
        self.M_basis = np.array([len(s) for s in spec], np.int32)
        for m in self.M_basis:
            if m > 2:
                raise NotImplementedError("multicomponent (M > 2) sublattices are not supported")
        self.active = self.M_basis > 1
        self._ops = lattice_point_ops(prim, spec)
        self._build_orbits()

    # ── geometry helpers ──
    def _cart(self, b, n):
        return (self.primitive_structure.get_scaled_positions()[b] + np.asarray(n)) @ self.primitive_structure.cell

    def _neighbors(self, b, rmax):
        """lattice sites (b2, n) of active basis sites within rmax of basis site b (offset 0), excluding itself."""
        P = self.primitive_structure.cell
        fb = self.primitive_structure.get_scaled_positions()
        hmin = min(abs(np.linalg.det(P)) / np.linalg.norm(np.cross(P[(i + 1) % 3], P[(i + 2) % 3])) for i in range(3))
        L = int(np.ceil(rmax / hmin)) + 1
        rng = range(-L, L + 1)
        out = []
        x0 = fb[b] @ P
        for n in itertools.product(rng, rng, rng):
            for b2 in range(self.n_basis):
                if not self.active[b2]:
                    continue
                dd = np.linalg.norm((fb[b2] + np.array(n)) @ P - x0)
                if 1e-8 < dd <= rmax + _TOL:
                    out.append((b2, tuple(int(v) for v in n)))
        return out

    def _translate_key(self, sites):
        """translation-canonical form: sorted tuple, anchored (min over anchors) so the anchor offset is 0."""
        best = None
        for (ba, na) in sites:
            s = tuple(sorted((b, (n[0] - na[0], n[1] - na[1], n[2] - na[2])) for b, n in sites))
            if s[0][1] != (0, 0, 0):
                # anchor must be the first element after sorting for a unique key; re-anchor on the first
                f = s[0][1]
                s = tuple(sorted((b, (n[0] - f[0], n[1] - f[1], n[2] - f[2])) for b, n in s))
            if best is None or s < best:
                best = s
        return best

    def _apply(self, op, sites):
        R, shifts, smap = op
        out = []
        for b, n in sites:
            n2 = R @ np.asarray(n) + shifts[b]
            out.append((int(smap[b]), tuple(int(v) for v in n2)))
        return out

    def _sym_key(self, sites):
        return min(self._translate_key(self._apply(op, sites)) for op in self._ops)

    def _build_orbits(self):
        cut = self.cutoffs
        nB = self.n_basis
        orbits = {}

        def dist(s1, s2):
            return np.linalg.norm(self._cart(*s1) - self._cart(*s2))

        # singlets
        for b in range(nB):
            if not self.active[b]:
                continue
            key = self._sym_key([(b, (0, 0, 0))])
            orbits.setdefault(key, set()).add(self._translate_key([(b, (0, 0, 0))]))
        for order, rc in enumerate(cut, start=2):
            for b in range(nB):
                if not self.active[b]:
                    continue
                home = (b, (0, 0, 0))
                nb = self._neighbors(b, rc)
                for combo in itertools.combinations(nb, order - 1):
                    sites = [home] + list(combo)
                    if any(dist(x, y) > rc + _TOL for x, y in itertools.combinations(sites, 2)):
                        continue
                    tk = self._translate_key(sites)
                    key = self._sym_key(sites)
                    orbits.setdefault(key, set()).add(tk)
        ol = []
        for key, inst in orbits.items():
            rep = list(key)
            xyz = np.array([self._cart(*s) for s in rep])
            c = xyz.mean(0)
            radius = float(np.mean(np.linalg.norm(xyz - c, axis=1)))
            dists = tuple(sorted(round(float(np.linalg.norm(xyz[i] - xyz[j])), 5)
                                 for i, j in itertools.combinations(range(len(rep)), 2)))
            ol.append(Orbit(len(rep), sorted(inst), radius, dists))
        ol.sort(key=lambda o: (o.order, round(o.radius, 5), o.distances, o.multiplicity))
        self.orbits = ol

    # ── icet-like API ──
    def __len__(self):
        """number of cluster-vector elements including the zerolet."""
        return 1 + len(self.orbits)

    def get_sublattice_index(self, b):
        return int(self.basis_to_sl[b])

    def species_index(self, b, z):
        """icet species index: rank of the atomic number among the basis site's allowed species."""
        spec = sorted(self.chemical_symbols[b], key=lambda x: ATOMIC_NUMBERS[x])
        zs = [ATOMIC_NUMBERS[s] for s in spec]
        return zs.index(int(z))

    def cluster_vector(self, atoms):
        """[1, Phi_1, ...]: orbit averages of the point-function products over ``atoms`` (brute force through the
        canonical template; used by tests and by ``ClusterExpansion.predict``)."""
        from .lattice import CanonicalCell
        cc = CanonicalCell(self.primitive_structure, atoms)
        sig = np.zeros(len(atoms), np.int64)
        for i in range(len(atoms)):
            b = int(cc.basis_orig[i])
            sig[i] = self.species_index(b, atoms.numbers[i]) if self.active[b] else 0
        sig_c = sig[cc.perm]
        d = cc.d; nB = self.n_basis
        kall = np.stack(np.meshgrid(np.arange(d[0]), np.arange(d[1]), np.arange(d[2]), indexing="ij"), -1).reshape(-1, 3)
        spins = np.where(sig_c == 0, -1.0, 1.0)        # binary: theta_0(0) = -1, theta_0(1) = +1
        cv = [1.0]
        for o in self.orbits:
            tot = 0.0
            for inst in o.clusters:
                prod = np.ones(len(kall))
                for b, n in inst:
                    kk = np.mod(kall + cc.uo(n), d)
                    site = ((kk[:, 0] * d[1] + kk[:, 1]) * d[2] + kk[:, 2]) * nB + b
                    prod = prod * spins[site]
                tot += prod.sum()
            cv.append(tot / (len(o.clusters) * len(kall)))
        return np.array(cv)
