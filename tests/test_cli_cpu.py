"""CPU tests of the command line, the per-replica starting-structure reader, the clean-supercell rebuild and the
heat-capacity estimator (no GPU calls)."""
import json
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN


def test_cli_accepts_the_reference_command_lines():
    from neuralce2_amd.mc.run import parse_args
    a = parse_args("--framework nce --ckpt model.pkl --init_cif x.cif --supercell 2 2 2 --pt --T_min 300 --T_max 1400 "
                   "--n_replicas 32 --n_steps 200000 --energy_mode local --out_dir runs/demo_nce".split())
    assert a.framework == "nce" and a.pt and a.supercell == [2, 2, 2] and a.energy_mode == "local"
    b = parse_args("--framework icet --ce m.ce --init_cif x.cif --supercell 2 2 2 --pt --T_min 300 --T_max 1400 "
                   "--n_replicas 32 --n_steps 200000 --energy_mode delta --out_dir runs/demo_icet".split())
    assert b.framework == "icet" and b.ce == "m.ce" and b.n_replicas == 32 and b.exchange_every == 100
    c = parse_args("--framework ce_cnn --ckpt m.pkl --init_cif x.cif --temps 300 600 --dtype float64".split())
    assert c.temps == [300.0, 600.0] and c.dtype == "float64"


def test_summary_keeps_scalars_and_lists():
    from dataclasses import dataclass
    from neuralce2_amd.mc.run import summarize

    @dataclass
    class R:
        a: float
        b: np.ndarray
        c: str
        d: dict
    s = summarize(R(np.float64(1.5), np.arange(3.0), "x", {"k": 1}))
    assert s == {"a": 1.5, "b": [0.0, 1.0, 2.0], "c": "x"}
    json.dumps(s)


def _cell():
    from neuralce2_amd.lattice import bulk_fcc
    at = bulk_fcc("Ni", 3.924, cubic=True).repeat((2, 2, 2))
    rng = np.random.default_rng(0)
    Z = np.full(len(at), 28); Z[rng.choice(len(at), 12, replace=False)] = 78
    at.numbers = Z
    return at


def test_init_npz_matches_sites_by_lattice_position(tmp_path):
    from neuralce2_amd.lattice import bulk_fcc
    from neuralce2_amd.mc.init_structures import load_init_npz
    prim = bulk_fcc("Ni", 3.924, cubic=False)
    at = _cell()
    rng = np.random.default_rng(1)
    R = 3
    nums = np.stack([rng.permutation(at.numbers) for _ in range(R)]).astype(np.uint8)
    order = rng.permutation(len(at))                       # the npz lists the sites in another order
    shift = np.array([1, 0, -1]) @ at.cell                 # and some through a periodic image
    pos = at.positions[order] + shift * (rng.random((len(at), 1)) < 0.5)
    p = tmp_path / "init.npz"
    temps = np.array([300.0, 600.0, 900.0])
    np.savez(p, numbers=nums[:, order], positions=pos, cell=at.cell, temps=temps)
    out, T = load_init_npz(str(p), prim, at, temps)
    assert np.array_equal(out, nums) and np.allclose(T, temps)
    with pytest.raises(ValueError, match="temperatures"):
        load_init_npz(str(p), prim, at, temps + 5.0)
    bad = nums.copy(); bad[1, 0] = 28 if bad[1, 0] == 78 else 78
    np.savez(p, numbers=bad[:, order], positions=pos, cell=at.cell, temps=temps)
    with pytest.raises(ValueError, match="composition"):
        load_init_npz(str(p), prim, at, temps)
    np.savez(p, numbers=nums[:, order], positions=pos, cell=at.cell * 1.01, temps=temps)
    with pytest.raises(ValueError, match="cell"):
        load_init_npz(str(p), prim, at, temps)


def test_clean_supercell_rebuild_transfers_species_by_position():
    from neuralce2_amd.lattice import Structure, bulk_fcc, site_decompose
    from neuralce2_amd.ce.icet_mc import _rebuild_as_clean_supercell
    prim = bulk_fcc("Ni", 3.924, cubic=False)
    at = _cell()
    rng = np.random.default_rng(2)
    order = rng.permutation(len(at))
    noisy = Structure(at.cell, at.positions[order] + rng.normal(0, 0.05, (len(at), 3)), at.numbers[order])
    clean = _rebuild_as_clean_supercell(noisy, prim)
    site_decompose(prim, clean)                             # on the lattice
    # every clean site carries the species of the original atom at that position
    for p, z in zip(clean.positions, clean.numbers):
        f = (at.positions - p) @ np.linalg.inv(at.cell); f -= np.round(f)
        j = int(np.argmin(((f @ at.cell) ** 2).sum(1)))
        assert at.numbers[j] == z


def test_heat_capacity_and_peak():
    from neuralce2_amd.analysis import heat_capacity, peak_interp, KB
    rng = np.random.default_rng(3)
    T = np.array([500.0, 800.0])
    e = rng.normal(0, 1, (1000, 2)) * np.array([0.01, 0.02])
    cv = heat_capacity(e, T, 48)
    assert np.allclose(cv, 48 * e.var(0) / (KB * T) ** 2)
    # a parabola in log T: the vertex is recovered exactly (equal and unequal spacing)
    for Ts in (np.geomspace(400, 1600, 9), np.array([400, 500, 620, 800, 1000, 1400.0])):
        x0 = np.log(777.0)
        cvs = 2.0 - 3.0 * (np.log(Ts) - x0) ** 2
        assert abs(peak_interp(Ts, cvs) - 777.0) < 1e-6
    assert peak_interp([1, 2, 3], [3, 2, 1]) == 1          # maximum at the edge


def test_exact_reference_thermodynamics_fixture_is_consistent():
    """the exact N = 48 table: C_v = d<E>/dT * N / k_B (finite differences) and the stated peak."""
    from neuralce2_amd.analysis import KB
    d = json.load(open(os.path.join(GOLDEN, "exact_N48.json")))
    T = np.array(d["T"]); E = np.array(d["E"]); Cv = np.array(d["Cv"])
    sel = (T > 400) & (T < 2000)
    dEdT = np.gradient(E, T)
    assert np.max(np.abs(dEdT[sel] / KB - Cv[sel])) < 0.02 * Cv[sel].max()
    k = int(np.argmax(Cv))
    assert abs(Cv[k] - 1.3305) < 2e-3 and abs(T[k] - 799) < 15
