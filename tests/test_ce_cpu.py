"""CPU tests: the oracle pinned against the reference's fixtures, closed forms and brute force; host-side geometry;
the C-ABI library loads and exports every declared symbol (no GPU compute here)."""
import json
import os
import re

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT, random_fcc

# ── Philox known answers (Random123 kat_vectors, philox4x32_10) ──
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    from oracle.ce_oracle import philox4x32_10 as oracle_philox
    from neuralce2_amd._lib import philox4x32_10 as lib_philox
    assert list(oracle_philox(ctr, key)) == out
    assert list(lib_philox(ctr, key)) == out


def test_library_exports_every_declared_symbol():
    from neuralce2_amd import _lib
    decl = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        txt = open(os.path.join(ROOT, "include", fn)).read()
        decl |= set(re.findall(r"\b(nce_[a-z0-9_]+)\s*\(", txt))
    assert decl, "no declarations found"
    for name in sorted(decl):
        assert hasattr(_lib.lib, name), name
    assert decl <= set(_lib.EXPORTED) | {"nce_last_error"}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """every field offset and struct size of the ctypes mirrors equals the C header's (compiled with g++)."""
    import shutil
    import subprocess
    from neuralce2_amd import _lib
    if not shutil.which("g++"):
        pytest.skip("no g++")
    pairs = [("nce_ce_model_desc", _lib.ModelDesc), ("nce_pt_job", _lib.PTJob),
             ("nce_pt_exchange_job", _lib.ExchangeJob), ("nce_nn_model_desc", _lib.NNModelDesc),
             ("nce_nn_job", _lib.NNJob)]
    lines = ['#include "%s"' % os.path.join(ROOT, "include", "neuralce_hip.h"), "#include <cstdio>",
             "#include <cstddef>", "int main() {"]
    want = []
    for cname, py in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(ctypes_sizeof(py))
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
            want.append(getattr(py, f).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    subprocess.run(["g++", "-o", str(tmp_path / "layout"), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(tmp_path / "layout")], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == want


def ctypes_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)


def test_orbit_order_reproduces_reference_model_energies(ptni_ce):
    """icet orbit order pinned by fits.json: ce.predict of the per-atom model (params / 256) on six orderings."""
    from neuralce2_amd.lattice import bulk_fcc
    fits = json.load(open(os.path.join(GOLDEN, "ptni_fits_c8_6_5.json")))["icet_c8_6_5"]
    assert len(ptni_ce.parameters) == fits["n_params"] == 38
    assert int((np.abs(ptni_ce.parameters) > 1e-8).sum()) == fits["n_nonzero"]
    conv = bulk_fcc("Ni", 3.924, cubic=True).repeat((4, 4, 4))
    frac = conv.get_scaled_positions() * 4
    code = ((np.round(frac * 2).astype(int) % 2) * np.array([1, 2, 4])).sum(1)
    sub = np.array([{0: 0, 3: 1, 5: 2, 6: 3}[int(c)] for c in code])
    h = np.round(frac.sum(1) * 2).astype(int); zc = np.floor(frac[:, 2] + 1e-6).astype(int)

    def mk(mask):
        a = conv.copy(); a.set_chemical_symbols(np.where(mask, "Pt", "Ni").tolist()); return a
    do22 = np.where((zc // 2) % 2 == 0, sub == 0, sub == 1)
    refs = {"L12_Ni3Pt": mk(sub == 0), "DO22_Ni3Pt": mk(do22), "L12_Pt3Ni": mk(sub != 0), "DO22_Pt3Ni": mk(~do22),
            "L10_NiPt": mk(np.isin(sub, [0, 1])), "L11_NiPt": mk((h // 2) % 2 == 0)}
    cs = ptni_ce.cluster_space
    for k, a in refs.items():
        e = float((ptni_ce.parameters / fits["eci_scale"]) @ cs.cluster_vector(a))
        assert abs(e - fits["refs"][k]) < 1e-12, (k, e, fits["refs"][k])


def test_orbit_multiplicities(ptni_ce):
    """8 pair shells (140 neighbours), 630 triplets and 1332 quadruplets contain a site (lattice count)."""
    cs = ptni_ce.cluster_space
    by = {}
    for o in cs.orbits:
        by[o.order] = by.get(o.order, 0) + o.multiplicity * o.order
    assert by == {1: 1, 2: 140, 3: 630, 4: 1332}
    assert len([o for o in cs.orbits if o.order == 2]) == 8


def test_snf_matches_sympy_invariants():
    import sympy
    from sympy.matrices.normalforms import smith_normal_form as sympy_snf
    from neuralce2_amd.lattice import smith_normal_form
    rng = np.random.default_rng(0)
    mats = [np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]]) * n for n in (1, 2, 4, 8, 64)]
    mats += [rng.integers(-4, 5, (3, 3)) for _ in range(30)]
    for T in mats:
        if round(abs(np.linalg.det(T))) == 0:
            continue
        U, d, V = smith_normal_form(T)
        assert np.array_equal(U @ T @ V, np.diag(d))
        assert abs(round(np.linalg.det(U))) == 1 and abs(round(np.linalg.det(V))) == 1
        assert all(d[i + 1] % d[i] == 0 for i in range(2))
        ref = np.abs(np.diag(np.array(sympy_snf(sympy.Matrix(T.tolist())).tolist(), dtype=np.int64)))
        assert sorted(ref.tolist()) == sorted(d.tolist())


def test_cubic_fcc_snf_diagonal():
    from neuralce2_amd.lattice import CanonicalCell, bulk_fcc
    for n in (1, 2, 8):
        cc = CanonicalCell(bulk_fcc("Ni", 3.924, cubic=False), bulk_fcc("Ni", 3.924, cubic=True).repeat((n, n, n)))
        assert cc.d.tolist() == [n, 2 * n, 2 * n]
        assert sorted(cc.perm.tolist()) == list(range(4 * n ** 3))


def _oracle(ce, tpl, at, scale=1.0):
    from neuralce2_amd.ce.ce_gpu_implicit import build_ce_gpu_implicit, encode_canonical
    from oracle.ce_oracle import OracleCE
    info = build_ce_gpu_implicit(ce, at, energy_scale=scale, template=tpl)
    return OracleCE(info), encode_canonical(info, ce.cluster_space, at).astype(np.int8), info


@pytest.mark.parametrize("n", [1, 2, 4])
def test_oracle_energy_equals_brute_force(ptni_ce, ptni_template, n):
    from oracle.ce_oracle import brute_energy
    at = random_fcc(n, seed=n)
    o, sig, _ = _oracle(ptni_ce, ptni_template, at)
    assert abs(o.energy(sig) - brute_energy(ptni_ce.cluster_space, ptni_ce.parameters, at)) < 1e-9


def test_oracle_energy_closed_forms(ptni_ce, ptni_template):
    """pure Ni: every spin -1 -> E = J0 + sum_o J_o (-1)^order; pure Pt: E = sum_o J_o (all spins +1)."""
    at = random_fcc(2, seed=0)
    o, sig, _ = _oracle(ptni_ce, ptni_template, at)
    p = ptni_ce.parameters; orders = [ob.order for ob in ptni_ce.cluster_space.orbits]
    e_ni = p[0] + sum(p[i + 1] * (-1) ** k for i, k in enumerate(orders))
    e_pt = p.sum()
    assert abs(o.energy(np.zeros_like(sig)) - e_ni) < 1e-9
    assert abs(o.energy(np.ones_like(sig)) - e_pt) < 1e-9


def test_oracle_delta_is_energy_difference_and_antisymmetric(ptni_ce, ptni_template, ptni_atoms_256):
    o, sig, _ = _oracle(ptni_ce, ptni_template, ptni_atoms_256)
    E = o.energy(sig); rng = np.random.default_rng(1); N = len(sig)
    for _ in range(40):
        i, j = rng.integers(0, N, 2)
        if rng.random() < 0.4:
            j = (i + rng.integers(1, 5)) % N               # near pairs share clusters
        d = o.delta(sig, i, j)
        s2 = sig.copy(); s2[i], s2[j] = sig[j], sig[i]
        assert abs(d - (o.energy(s2) - E)) < 1e-9
        assert abs(o.delta(sig, j, i) - d) < 1e-9       # same swap, other order
        assert abs(o.delta(s2, i, j) + d) < 1e-9        # reverse move


def test_oracle_extensive_under_tiling(ptni_ce, ptni_template, ptni_atoms_256):
    """the periodic 2x2x2 image has 8x the energy with energy_scale 8, and one swap keeps its energy change
    (reference tests/test_ce_implicit.py:56-72)."""
    a4 = ptni_atoms_256; a8 = a4.repeat((2, 2, 2))
    o4, s4, i4 = _oracle(ptni_ce, ptni_template, a4)
    o8, s8, i8 = _oracle(ptni_ce, ptni_template, a8, scale=8.0)
    assert abs(o8.energy(s8) / o4.energy(s4) - 8.0) < 1e-12
    ia, ib = 3, 7
    while s4[ia] == s4[ib]:
        ib += 1
    ja, jb = int(i8['inv_perm'][int(i4['perm'][ia])]), int(i8['inv_perm'][int(i4['perm'][ib])])
    assert abs(o8.delta(s8, ja, jb) / o4.delta(s4, ia, ib) - 1.0) < 1e-9


def test_template_invariants(ptni_template):
    t = ptni_template
    assert (t.Rt, t.J, t.mo) == (614, 2103, 4)
    assert t.W_b.tolist() == [141]                    # the 8 A covering window
    slots = t.ent_slot[0, :t.J_b[0]]
    assert np.all((slots == 0).sum(1) == 1)           # every reverse entry holds the site itself exactly once
    # reverse coefficients add up to order x forward coefficients
    assert abs(t.ent_coef[0].sum() - (t.T_coef * t.T_order).sum()) < 1e-12


def test_oracle_chunk_running_total_and_injected_decisions(ptni_ce, ptni_template):
    at = random_fcc(2, seed=3)
    o, sig, _ = _oracle(ptni_ce, ptni_template, at)
    R, steps, N = 4, 300, len(sig)
    rng = np.random.default_rng(5)
    occ = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    E = np.array([o.energy(x) for x in occ])
    betas = np.full(R, 1.0 / (8.617333e-5 * 900.0))
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    occ_py = occ.copy(); E_py = E.copy()
    o.run_chunk(occ, E, betas, steps, inj=(ia, ib, u))
    for r in range(R):                                 # pure-Python restatement of the step (pt_mcmc_ce.py:558-566)
        for t in range(steps):
            a, b = ia[r, t], ib[r, t]
            if occ_py[r, a] != occ_py[r, b]:
                d = o.delta(occ_py[r], a, b)
                if np.log(u[r, t]) < -betas[r] * d:
                    occ_py[r, a], occ_py[r, b] = occ_py[r, b], occ_py[r, a]; E_py[r] += d
    assert np.array_equal(occ, occ_py)
    assert np.max(np.abs(E - E_py)) < 1e-9
    assert np.max(np.abs(E - np.array([o.energy(x) for x in occ]))) < 1e-9


def test_oracle_exchange_semantics():
    """neighbour pairs of the chosen parity; accept iff log u < (b_j - b_i)(E_j - E_i) (pt_mcmc_ce.py:748-760)."""
    from oracle.ce_oracle import exchange
    R = 6
    beta = np.linspace(40, 10, R)
    E = np.array([0.0, -1.0, 0.5, 0.4, -3.0, 1.0])
    ros = np.arange(R, dtype=np.int32); brep = np.zeros(R)
    u = np.array([0.5, 0.5, 0.5])
    n = exchange(E, ros, beta, brep, inj_parity=0, inj_u=u)
    exp = np.arange(R)
    for p, i in enumerate(range(0, R - 1, 2)):
        la = (beta[i + 1] - beta[i]) * (E[i + 1] - E[i])
        if np.log(u[p]) < la:
            exp[i], exp[i + 1] = exp[i + 1], exp[i]
    assert ros.tolist() == exp.tolist() and n == int((exp != np.arange(R)).sum() // 2)
    assert np.allclose(brep[ros], beta)


def test_safe_pickle_reads_ce_without_unpickling():
    from neuralce2_amd import safe_pickle
    import tarfile
    with tarfile.open(os.path.join(GOLDEN, "icet_c8_6_5.ce")) as t:
        items = safe_pickle.loads(t.extractfile("items").read())
    assert items["parameters"].dtype == np.float64 and items["parameters"].shape == (38,)
    assert items["metadata"]["icet_version"] == "3.2"


def test_cif_roundtrip(tmp_path):
    from neuralce2_amd.lattice import read_cif, write_cif
    at = random_fcc(2, seed=9)
    p = tmp_path / "a.cif"
    write_cif(str(p), at)
    b = read_cif(str(p))
    assert np.allclose(b.cell, at.cell, atol=1e-6)
    assert np.array_equal(b.numbers, at.numbers)
    assert np.allclose(b.positions, at.positions, atol=1e-6)


def test_every_module_compiles():
    """every Python file of the package, the oracle, the tests and the entry points is valid source."""
    files = [os.path.join(ROOT, f) for f in ("bench.py", "__graft_entry__.py")]
    for top in ("neuralce2_amd", "oracle", "tests"):
        for dp, _, fs in os.walk(os.path.join(ROOT, top)):
            files += [os.path.join(dp, f) for f in fs if f.endswith(".py")]
    bad = []
    for path in files:
        try:
            compile(open(path).read(), path, "exec")
        except SyntaxError as e:
            bad.append(f"{path}:{e.lineno}: {e.msg}")
    assert not bad, bad


def _exact_cell():
    """2 x 3 x 4 primitive fcc supercell (24 sites; shortest image 5.55 A), pairs <= 5 A and triplets <= 4 A, so no
    cluster holds a site twice; per-site ECIs N(0, 20 meV) from a seed."""
    from neuralce2_amd.clusterspace import ClusterSpace
    from neuralce2_amd.lattice import bulk_fcc
    cs = ClusterSpace(bulk_fcc("Ni", 3.924, cubic=False), [5.0, 4.0], [["Ni", "Pt"]])
    at = bulk_fcc("Ni", 3.924, cubic=False).repeat((2, 3, 4))
    params = np.zeros(len(cs)); params[1:] = np.random.default_rng(0).normal(0, 0.02, len(cs) - 1)
    return cs, at, params


def test_exact_enumeration_energies_match_brute_force():
    """the enumerator's energies of C(24, 12) configurations (lexicographic k-subsets) equal the brute-force energy
    of the same configurations, and every energy is invariant under the lattice's translations."""
    from itertools import combinations, islice
    from math import comb
    from oracle.ce_oracle import brute_terms, brute_energy, enumerate_energies
    cs, at, params = _exact_cell()
    N = len(at)
    sites, coef, e0 = brute_terms(cs, params, at, float(N))
    E = enumerate_energies(sites, coef, e0, N, 12)
    assert len(E) == comb(24, 12)
    picks = [0, 1, 17, 5000, 123456, len(E) // 2, len(E) - 2, len(E) - 1]
    subsets = {}
    for r, c in enumerate(combinations(range(N), 12)):
        if r in picks:
            subsets[r] = c
        if r > max(picks):
            break
    for r, c in subsets.items():
        a = at.copy(); Z = np.full(N, 28); Z[list(c)] = 78; a.numbers = Z
        assert abs(E[r] - brute_energy(cs, params, a, float(N))) < 1e-12, r
    # the energies take few distinct values per symmetry class: the minimum is reached at least N / |orbit| times
    assert np.sum(np.isclose(E, E.min(), rtol=0, atol=1e-12)) >= 2


def test_mchammer_style_swaps_keep_the_energy_and_composition(ptni_ce, ptni_template, ptni_atoms_256):
    """the single-thread CPU sampler (timed beside the GPU as the mchammer-style baseline): every proposal changes
    the configuration, the composition is kept, and the carried energy equals a full recompute."""
    o, sig, _ = _oracle(ptni_ce, ptni_template, ptni_atoms_256)
    occ = sig.copy()
    E0 = o.energy(occ)
    acc, E = o.mchammer_swaps(occ, E0, 1.0 / (8.617333e-5 * 900.0), 5000, seed=3)
    assert 0 < acc < 5000
    assert np.bincount(occ, minlength=2).tolist() == np.bincount(sig, minlength=2).tolist()
    assert abs(E - o.energy(occ)) < 1e-9
