"""Crystal structures, integer supercell algebra and the canonical site order.

A minimal stand-in for the parts of ``ase`` the Monte Carlo path needs (cell + positions + atomic numbers,
``repeat``, fcc builder, cif read/write) plus the lattice arithmetic of the reference's template-only tables:

* ``site_decompose``  every supercell site -> (basis index, integer primitive-cell offset), integer supercell matrix T
  (contract of neuralce/nce/gpu_layer0_tiled.py:16-33; bases chosen by Cartesian distance, offsets reduced exactly
  in integers through the adjugate of T).
* ``smith_normal_form``  U T V = diag(d) with U, V unimodular (reference: neuralce/ce/ce_gpu_implicit.py:28-39 uses
  sympy; this is an exact integer implementation with the same contract).
* ``canonical_cells``  k = U n mod d and the cell index (k0 d1 + k1) d2 + k2 (ce_gpu_implicit.py:42-46).

Canonical site order: i = cell * n_basis + basis.  ``perm`` maps canonical -> input order (sigma_canon =
sigma_orig[perm]), ``inv_perm`` the reverse, as in ce_gpu_implicit.py:71-74.
"""
from __future__ import annotations

import numpy as np

CHEMICAL_SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y "
    "Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os "
    "Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh "
    "Fl Mc Lv Ts Og").split()
ATOMIC_NUMBERS = {s: z for z, s in enumerate(CHEMICAL_SYMBOLS)}


class Structure:
    """Periodic structure: ``cell`` rows are the lattice vectors (Angstrom), ``positions`` Cartesian (N, 3),
    ``numbers`` atomic numbers (N,).  Same conventions as ``ase.Atoms`` for the attributes used here."""

    def __init__(self, cell, positions, numbers, pbc=True):
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.numbers = np.array(numbers, dtype=np.int64).reshape(-1)
        if len(self.numbers) != len(self.positions):
            raise ValueError(f"{len(self.numbers)} numbers for {len(self.positions)} positions")
        self.pbc = np.array([bool(pbc)] * 3) if np.ndim(pbc) == 0 else np.asarray(pbc, bool)

    def __len__(self):
        return len(self.numbers)

    def copy(self):
        return Structure(self.cell, self.positions, self.numbers, self.pbc)

    def get_scaled_positions(self, wrap=True):
        f = np.linalg.solve(self.cell.T, self.positions.T).T
        if wrap:
            f = np.mod(f, 1.0)
            f[np.isclose(f, 1.0, atol=1e-12)] = 0.0
        return f

    def get_chemical_symbols(self):
        return [CHEMICAL_SYMBOLS[int(z)] for z in self.numbers]

    def set_chemical_symbols(self, symbols):
        self.numbers = np.array([ATOMIC_NUMBERS[s] for s in symbols], np.int64)

    def repeat(self, reps):
        """Tile the cell; atom order ((m0*n1 + m1)*n2 + m2)*n_atoms + a, as ase.Atoms.repeat."""
        reps = tuple(int(r) for r in (reps if np.ndim(reps) else (reps,) * 3))
        n = len(self)
        m = np.indices(reps, dtype=np.int64).reshape(3, -1).T          # (m0, m1, m2), m2 fastest
        pos = (self.positions[None, :, :] + (m @ self.cell)[:, None, :]).reshape(-1, 3)
        num = np.tile(self.numbers, len(m))
        cell = self.cell * np.array(reps, np.float64)[:, None]
        out = Structure(cell, pos, num, self.pbc)
        assert len(out) == n * len(m)
        return out

    def __repr__(self):
        from collections import Counter
        c = Counter(self.get_chemical_symbols())
        return f"Structure({''.join(f'{k}{v}' for k, v in sorted(c.items()))}, cell={np.round(self.cell, 4).tolist()})"


def bulk_fcc(symbol, a, cubic=True):
    """fcc crystal; cubic=True gives the 4-atom conventional cell in ase's order, else the 1-atom primitive."""
    z = ATOMIC_NUMBERS[symbol]
    if cubic:
        f = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
        return Structure(np.eye(3) * a, f * a, [z] * 4)
    h = a / 2.0
    return Structure([[0, h, h], [h, 0, h], [h, h, 0]], [[0, 0, 0]], [z])


# ── integer lattice algebra ───────────────────────────────────────────────────────────────────────
def smith_normal_form(T):
    """U, d, V with U @ T @ V = diag(d), U and V unimodular integer matrices, d > 0 and d[i] | d[i+1].

    Exact elimination over Python integers (same contract as ce_gpu_implicit.py:28-39, which calls sympy)."""
    A = [[int(x) for x in row] for row in np.asarray(T, dtype=np.int64).tolist()]
    n = 3
    U = [[int(i == j) for j in range(n)] for i in range(n)]
    V = [[int(i == j) for j in range(n)] for i in range(n)]

    def swap_rows(M, i, j): M[i], M[j] = M[j], M[i]

    def swap_cols(M, i, j):
        for r in M: r[i], r[j] = r[j], r[i]

    def add_row(M, src, dst, f):          # row dst += f * row src
        M[dst] = [a + f * b for a, b in zip(M[dst], M[src])]

    def add_col(M, src, dst, f):          # col dst += f * col src
        for r in M: r[dst] += f * r[src]

    for t in range(n):
        while True:
            piv = [(abs(A[i][j]), i, j) for i in range(t, n) for j in range(t, n) if A[i][j] != 0]
            if not piv:
                raise ValueError("singular supercell matrix")
            _, pi, pj = min(piv)
            swap_rows(A, t, pi); swap_rows(U, t, pi)
            swap_cols(A, t, pj); swap_cols(V, t, pj)
            done = True
            for i in range(t + 1, n):
                q = A[i][t] // A[t][t]
                if q: add_row(A, t, i, -q); add_row(U, t, i, -q)
                if A[i][t] != 0: done = False
            for j in range(t + 1, n):
                q = A[t][j] // A[t][t]
                if q: add_col(A, t, j, -q); add_col(V, t, j, -q)
                if A[t][j] != 0: done = False
            if not done:
                continue
            bad = [(i, j) for i in range(t + 1, n) for j in range(t + 1, n) if A[i][j] % A[t][t] != 0]
            if bad:                                      # divisibility: pull an offending row into row t
                add_row(A, bad[0][0], t, 1); add_row(U, bad[0][0], t, 1)
                continue
            break
        if A[t][t] < 0:
            A[t] = [-x for x in A[t]]; U[t] = [-x for x in U[t]]
    U = np.array(U, np.int64); V = np.array(V, np.int64); D = np.array(A, np.int64)
    d = np.diag(D).copy()
    assert np.array_equal(U @ np.asarray(T, np.int64) @ V, np.diag(d)), "SNF does not reproduce T"
    assert abs(round(np.linalg.det(U))) == 1 and abs(round(np.linalg.det(V))) == 1
    return U, d, V


def canonical_cells(off, U, d):
    """integer offsets (..., 3) -> (k = U n mod d, cell index (k0 d1 + k1) d2 + k2)."""
    k = np.mod(np.asarray(off, np.int64) @ U.T, d)
    return k, (k[..., 0] * d[1] + k[..., 1]) * d[2] + k[..., 2]


class OffLatticeError(ValueError):
    """A site of the structure is not a lattice site of the primitive (or the cell is not an integer supercell)."""


def integer_supercell(prim, atoms):
    """T with S = T P (rows: lattice vectors) and n_cells = |det T|; refuses a cell that is no integer multiple."""
    Tf = np.linalg.solve(prim.cell.T, atoms.cell.T).T
    T = np.rint(Tf).astype(np.int64)
    if not np.allclose(Tf, T, atol=1e-4):
        raise OffLatticeError(f"the cell is not an integer supercell of the primitive: T = {np.round(Tf, 4).tolist()}")
    n_cells = abs(int(round(np.linalg.det(T))))
    if n_cells == 0:
        raise OffLatticeError("singular supercell matrix")
    return T, n_cells


def lattice_coordinates(prim, atoms, tol=5e-3):
    """Every site x -> (basis b, integer cell vector n) with x = (n + f_b) P, f_b the basis' fractional position.

    The basis is the one whose nearest lattice point is closest in Cartesian distance; a site farther than `tol`
    Angstrom from every lattice point raises OffLatticeError.  n is not reduced into the supercell."""
    P = prim.cell
    frac = atoms.positions @ np.linalg.inv(P)
    fb = prim.get_scaled_positions()
    best = basis = n = None
    for b in range(len(prim)):
        r = frac - fb[b]
        nb = np.rint(r)
        r -= nb
        miss = np.sqrt(np.einsum("ij,ij->i", r @ P, r @ P))
        if b == 0:
            best, basis, n = miss, np.zeros(len(atoms), np.int64), nb.astype(np.int64)
            continue
        take = miss < best
        best = np.where(take, miss, best)
        basis[take] = b
        n[take] = nb[take].astype(np.int64)
    if len(atoms) and best.max() > tol:
        i = int(best.argmax())
        raise OffLatticeError(f"site {i} lies {best[i]:.3g} A from the nearest lattice site of the primitive")
    return basis, n, best


def _adjugate3(T):
    """integer adjugate of a 3 x 3 integer matrix: T @ adj(T) = det(T) I."""
    t = [[int(x) for x in row] for row in np.asarray(T)]
    adj = [[0] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            r = [k for k in range(3) if k != j]
            c = [k for k in range(3) if k != i]
            minor = t[r[0]][c[0]] * t[r[1]][c[1]] - t[r[0]][c[1]] * t[r[1]][c[0]]
            adj[i][j] = (-1) ** (i + j) * minor
    return np.array(adj, np.int64)


def reduce_offsets(n, T):
    """Integer cell vectors (..., 3) -> the representative n - q T inside the supercell, q = floor(n T^-1),
    evaluated exactly in integers as floor(n adj(T) / det T)."""
    n = np.asarray(n, np.int64)
    adj = _adjugate3(T)
    det = int(round(np.linalg.det(np.asarray(T, float))))
    num = n @ adj
    if det < 0:
        num, det = -num, -det
    return n - np.floor_divide(num, det) @ np.asarray(T, np.int64)


def site_decompose(prim, atoms, tol=5e-3):
    """Every supercell site -> (basis index, integer primitive-cell offset reduced into the supercell), the integer
    supercell matrix T (S = T P), n_cells and the reduction function; the contract of the reference's
    site_decompose (nce/gpu_layer0_tiled.py:16-33).  Off-lattice sites raise OffLatticeError."""
    T, n_cells = integer_supercell(prim, atoms)
    basis, n, _ = lattice_coordinates(prim, atoms, tol)
    return basis, reduce_offsets(n, T), T, n_cells, lambda t: reduce_offsets(t, T)


class CanonicalCell:
    """Canonical site order of a supercell of ``prim``: SNF (U, d), perm / inv_perm, basis of every site.

    k = U n mod d is unchanged by a supercell translation n -> n + m T, so the cell index needs no reduction of the
    site offsets first; ``index`` holds each input site's canonical index."""

    def __init__(self, prim, atoms):
        self.prim = prim
        self.nB = len(prim)
        T, n_cells = integer_supercell(prim, atoms)
        basis, n, _ = lattice_coordinates(prim, atoms)
        N = len(atoms)
        if N != self.nB * n_cells:
            raise OffLatticeError(f"{N} sites for {n_cells} primitive cells of {self.nB} sites")
        U, d, _ = smith_normal_form(T.T)            # rows of T = supercell vectors -> lattice T^T Z^3
        if int(np.prod(d)) != n_cells:
            raise OffLatticeError("Smith normal form volume differs from |det T|")
        _, cell_idx = canonical_cells(n, U, d)
        canon = cell_idx * self.nB + basis
        if np.bincount(canon, minlength=N).max() > 1:
            raise OffLatticeError("two sites fall on the same lattice site: not a clean tiling of the primitive")
        self.index = canon
        self.perm = np.argsort(canon).astype(np.int64)
        self.inv_perm = np.empty(N, np.int64); self.inv_perm[self.perm] = np.arange(N)
        self.U, self.d, self.T = U, d, T
        self.n_cells = n_cells
        self.N = N
        self.basis_orig = basis

    def uo(self, off):
        """U off mod d for integer offsets (..., 3)."""
        return np.mod(np.asarray(off, np.int64) @ self.U.T, self.d)


# ── cif I/O (the subset the reference's structure files use) ──────────────────────────────────────
def _cif_float(s):
    s = s.strip()
    if "(" in s:
        s = s[:s.index("(")]
    return float(s)


def read_cif(path):
    """Read a P1 (or symmetry-expanded) cif with a single atom_site loop."""
    with open(path) as f:
        lines = [ln.strip() for ln in f if ln.strip() and not ln.strip().startswith("#")]
    vals = {}
    for ln in lines:
        if ln.startswith("_cell_") and len(ln.split()) >= 2:
            k, v = ln.split()[:2]; vals[k] = _cif_float(v)
    a, b, c = vals["_cell_length_a"], vals["_cell_length_b"], vals["_cell_length_c"]
    al, be, ga = (np.radians(vals[k]) for k in ("_cell_angle_alpha", "_cell_angle_beta", "_cell_angle_gamma"))
    cell = _cell_from_par(a, b, c, al, be, ga)
    # atom_site loop
    i = 0; frac = []; syms = []
    while i < len(lines):
        if lines[i] == "loop_" and i + 1 < len(lines) and lines[i + 1].startswith("_atom_site_"):
            keys = []; i += 1
            while i < len(lines) and lines[i].startswith("_"):
                keys.append(lines[i].split()[0]); i += 1
            while i < len(lines) and not lines[i].startswith("_") and lines[i] != "loop_":
                tok = lines[i].split()
                if len(tok) < len(keys):
                    break
                row = dict(zip(keys, tok))
                sym = row.get("_atom_site_type_symbol", row.get("_atom_site_label"))
                sym = "".join(ch for ch in sym if ch.isalpha())
                sym = sym[0].upper() + sym[1:].lower()
                if sym not in ATOMIC_NUMBERS and sym[:2] in ATOMIC_NUMBERS:
                    sym = sym[:2]
                syms.append(sym)
                frac.append([_cif_float(row[f"_atom_site_fract_{x}"]) for x in "xyz"])
                i += 1
            break
        i += 1
    if not frac:
        raise ValueError(f"{path}: no _atom_site loop found")
    frac = np.array(frac)
    return Structure(cell, frac @ cell, [ATOMIC_NUMBERS[s] for s in syms])


def _cell_from_par(a, b, c, al, be, ga):
    """Lattice vectors from cell parameters (a along x, b in the xy plane), the convention of ase.geometry."""
    ca, cb, cg, sg = np.cos(al), np.cos(be), np.cos(ga), np.sin(ga)
    for v in (ca, cb, cg):
        pass
    ca = 0.0 if abs(ca) < 1e-14 else ca
    cb = 0.0 if abs(cb) < 1e-14 else cb
    cg = 0.0 if abs(cg) < 1e-14 else cg
    sg = np.sqrt(1.0 - cg * cg)
    va = [a, 0.0, 0.0]
    vb = [b * cg, b * sg, 0.0]
    cx = cb; cy = (ca - cb * cg) / sg; cz = np.sqrt(max(1.0 - cx * cx - cy * cy, 0.0))
    vc = [c * cx, c * cy, c * cz]
    return np.array([va, vb, vc])


def write_cif(path, atoms):
    cell = atoms.cell
    a, b, c = np.linalg.norm(cell, axis=1)
    al = np.degrees(np.arccos(np.dot(cell[1], cell[2]) / (b * c)))
    be = np.degrees(np.arccos(np.dot(cell[0], cell[2]) / (a * c)))
    ga = np.degrees(np.arccos(np.dot(cell[0], cell[1]) / (a * b)))
    f = atoms.get_scaled_positions()
    syms = atoms.get_chemical_symbols()
    with open(path, "w") as fh:
        fh.write("data_image0\n_symmetry_space_group_name_H-M    'P 1'\n_symmetry_int_tables_number    1\n\n")
        fh.write(f"_cell_length_a       {a:.8f}\n_cell_length_b       {b:.8f}\n_cell_length_c       {c:.8f}\n")
        fh.write(f"_cell_angle_alpha    {al:.8f}\n_cell_angle_beta     {be:.8f}\n_cell_angle_gamma    {ga:.8f}\n\n")
        fh.write("loop_\n  _atom_site_label\n  _atom_site_occupancy\n  _atom_site_fract_x\n  _atom_site_fract_y\n"
                 "  _atom_site_fract_z\n  _atom_site_type_symbol\n")
        for k, (s, p) in enumerate(zip(syms, f)):
            fh.write(f"  {s}{k + 1} 1.0000 {p[0]:.8f} {p[1]:.8f} {p[2]:.8f} {s}\n")
