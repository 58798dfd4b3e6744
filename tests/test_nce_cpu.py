"""CPU tests of the NeuralCE oracle and host geometry, pinned by the reference's shipped data:
layer-0 features of the 121 training cells, the model's own predictions on them, and the exact-enumeration
ground-state energy on 48 sites."""
import json
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN


@pytest.fixture(scope="module")
def model():
    from neuralce2_amd.nce.model_io import load_model
    from neuralce2_amd.clusterspace import ClusterSpace
    mdl = load_model(os.path.join(GOLDEN, "nipt_v3_nomax_sel.pkl"))
    cs = ClusterSpace(mdl["primitive"], mdl["cutoffs"], mdl["chemical_symbols"])
    return mdl, cs


@pytest.fixture(scope="module")
def feats():
    return np.load(os.path.join(GOLDEN, "nce_features_subset.npz"), allow_pickle=False)


def test_model_file_reads_without_unpickling(model):
    mdl, cs = model
    assert mdl["cutoffs"] == [7.0, 5.0, 4.5] and mdl["shell_edges"] == [0.0, 3.3, 4.3, 5.2]
    assert mdl["K_chem"] == len(cs) == 18
    assert list(mdl["keep"]) == [0, 1, 2, 3, 6, 8, 10, 11, 12, 13, 14, 16]
    assert abs(mdl["ym"] + 5.929046420849914) < 1e-15 and abs(mdl["ys"] - 0.09400637081046036) < 1e-15
    assert len(mdl["params"]) == 3


def test_layer0_matches_reference_features(model, feats):
    """raw h0 of all 121 training cells (3x3x3 conventional fcc, ase atom order; the singlet column is the spin)."""
    from neuralce2_amd.lattice import bulk_fcc
    from oracle.nce_oracle import h0_raw
    mdl, cs = model
    keep = np.asarray(mdl["keep"])
    conv = bulk_fcc("Ni", 3.924, cubic=True).repeat((3, 3, 3))
    err = 0.0
    for n in range(0, 121, 5):
        h0 = h0_raw(cs, conv, feats["spins"][n].astype(float))
        err = max(err, float(np.abs(h0[:, keep] - feats["H"][n]).max()))
    assert err < 1e-12, err


def test_shells_match_reference_tables(model, feats):
    from neuralce2_amd.lattice import bulk_fcc
    from oracle.nce_oracle import shells
    mdl, _ = model
    part = shells(mdl["primitive"], bulk_fcc("Ni", 3.924, cubic=True).repeat((3, 3, 3)), mdl["shell_edges"])
    for p in range(3):
        for s in range(108):
            assert sorted(feats["SI"][p, s][feats["SM"][p, s] > 0].tolist()) == sorted(part[p][s].tolist())


def test_forward_matches_reference_predictions(model, feats):
    """the network on the reference's own features reproduces its saved predictions (seed 0) exactly."""
    from neuralce2_amd.lattice import bulk_fcc
    from neuralce2_amd.nce.model_io import expand_params
    from oracle.nce_oracle import shells, energy_std, standardize
    mdl, _ = model
    keep = np.asarray(mdl["keep"])
    part = shells(mdl["primitive"], bulk_fcc("Ni", 3.924, cubic=True).repeat((3, 3, 3)), mdl["shell_edges"])
    P, mu, sd = expand_params(mdl, 0)
    for n in range(121):
        H = np.zeros((108, 18)); H[:, keep] = feats["H"][n]
        e = energy_std(P, standardize(H, mu, sd), part) * mdl["ys"] + mdl["ym"]
        assert abs(e - feats["pred"][n]) < 1e-12, (n, e, feats["pred"][n])


def test_l10_ground_state_energy(model):
    from neuralce2_amd.lattice import read_cif
    from neuralce2_amd.nce.model_io import expand_params
    from oracle.nce_oracle import Oracle
    mdl, cs = model
    at = read_cif(os.path.join(GOLDEN, "init_L10_N48.cif"))
    P, mu, sd = expand_params(mdl, 0)
    o = Oracle(cs, at, mdl["shell_edges"], P, mu, sd)
    e = o.energy((at.numbers == 78).astype(int)) * mdl["ys"] + mdl["ym"]
    assert abs(e - json.load(open(os.path.join(GOLDEN, "exact_N48.json")))["E_L10"]) < 1e-12


def test_geometry_template_consistency(model):
    """the product's template h0 (window + reverse entries) equals the oracle's positional h0."""
    from neuralce2_amd.lattice import bulk_fcc
    from neuralce2_amd.nce.gpu_layer0_implicit import build_geometry
    from oracle.nce_oracle import h0_raw
    mdl, cs = model
    at = bulk_fcc("Ni", 3.924, cubic=True).repeat((3, 3, 3))
    rng = np.random.default_rng(0)
    spins = rng.choice([-1.0, 1.0], len(at))
    geo = build_geometry(cs, at, mdl["shell_edges"], 1)
    d = geo["d"]; N = geo["N"]
    sp_c = spins[geo["perm"]]
    h0 = np.zeros((N, len(cs))); h0[:, 0] = 1.0
    k = np.stack(np.meshgrid(np.arange(d[0]), np.arange(d[1]), np.arange(d[2]), indexing="ij"), -1).reshape(-1, 3)
    for e in range(len(geo["ent_col"])):
        prod = np.ones(N)
        for q in range(4):
            s = int(geo["ent_slot"][e, q])
            if s == 0xFFFF:
                continue
            kk = np.mod(k + geo["window"][s, 1:], d)
            prod = prod * sp_c[(kk[:, 0] * d[1] + kk[:, 1]) * d[2] + kk[:, 2]]
        h0[:, geo["ent_col"][e]] += prod * geo["ent_coef"][e]
    ref = h0_raw(cs, at, spins)[geo["perm"]]
    assert np.abs(h0 - ref).max() < 1e-12


# ── the 12-neighbour SiLU network with a summed readout (the 16^3 NeuralCE workload's network) ──────────────────
def _silu_setup(n=3, seed=3):
    from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
    from neuralce2_amd.nce.mc_model import silu_network
    cs = fcc_cluster_space((9.0, 6.0, 5.0))
    P = silu_network(len(cs), 64, 3, 1, seed=seed)
    return cs, fcc_cell(n, 0.5, seed), P


def test_silu_network_matches_dense_torch_restatement():
    """oracle energy_std(kind='silu', readout='sum') against an independent torch evaluation: the shell mean as a
    dense (N, N) neighbour-average matrix built from the partner tables (pinned by the reference's SI/SM tables),
    torch.nn.functional.silu, per-site readout summed."""
    import torch
    from oracle.nce_oracle import Oracle
    cs, at, P = _silu_setup()
    C0 = len(cs)
    o = Oracle(cs, at, [0.0, 3.3], P, np.zeros(C0), np.ones(C0), kind="silu", readout="sum")
    part = o.partners[0]
    N = len(at)
    assert part.shape == (N, 12)
    A = torch.zeros((N, N), dtype=torch.float64)
    for j in range(12):
        A[torch.arange(N), torch.as_tensor(part[:, j])] += 1.0 / 12.0
    rng = np.random.default_rng(5)
    for _ in range(3):
        sig = rng.permutation((at.numbers == 78).astype(int))
        h0s = torch.as_tensor(o.h0s(sig))
        h = h0s
        for L in P["layers"]:
            pre = h @ torch.as_tensor(L["self"]) + (A @ h) @ torch.as_tensor(L["shell"][0]) + torch.as_tensor(L["b"])
            h = torch.nn.functional.silu(pre)
        e = h0s @ torch.as_tensor(P["lin"]) + h @ torch.as_tensor(P["head"]["w"])
        E_t = float(e.sum()) + float(P["head"]["b"])
        assert abs(o.energy(sig) - E_t) < 1e-12 * max(1.0, abs(E_t)), (o.energy(sig), E_t)


def test_silu_network_closed_forms():
    """(1) with every hidden weight zero the energy is the linear layer-0 term sum_i h0s_i . lin + b;
    (2) on a pure cell every site sees the same h0 and every shell mean equals the site's own row, so the network
    reduces to one site's scalar recursion h <- silu(h (W_self + W_shell) + b), E = N (h0 . lin + h_L . w) + b."""
    import math
    from oracle.nce_oracle import Oracle
    cs, at, P = _silu_setup()
    C0 = len(cs); N = len(at)
    mu, sd = np.zeros(C0), np.ones(C0)
    P0 = dict(lin=P["lin"], head=dict(w=P["head"]["w"], b=np.asarray(0.375)),
              layers=[dict(self=np.zeros_like(L["self"]), shell=np.zeros_like(L["shell"]), b=np.zeros_like(L["b"]))
                      for L in P["layers"]])
    o0 = Oracle(cs, at, [0.0, 3.3], P0, mu, sd, kind="silu", readout="sum")
    sig = np.random.default_rng(2).permutation((at.numbers == 78).astype(int))
    want = float((o0.h0s(sig) @ P["lin"]).sum()) + 0.375
    assert abs(o0.energy(sig) - want) < 1e-11
    Pb = dict(P, layers=[dict(L, b=np.linspace(-0.3, 0.4, L["b"].size)) for L in P["layers"]])
    o = Oracle(cs, at, [0.0, 3.3], Pb, mu, sd, kind="silu", readout="sum")
    for s in (0, 1):
        h0 = o.h0s(np.full(N, s))[0]
        h = h0
        for L in Pb["layers"]:
            pre = h @ (L["self"] + L["shell"][0]) + L["b"]
            h = np.array([x / (1.0 + math.exp(-x)) for x in pre])
        want = N * float(h0 @ Pb["lin"] + h @ Pb["head"]["w"]) + float(Pb["head"]["b"])
        assert abs(o.energy(np.full(N, s)) - want) < 1e-10 * max(1.0, abs(want)), (s, o.energy(np.full(N, s)), want)


# ── seed ensemble and refused parameter sets ──────────────────────────────────────────────────────────────────
def test_seed_ensemble_is_the_mean_of_the_seeds(model):
    """the fused ensemble network (width 3 x 16) gives the mean of the three per-seed energies, on the reference's
    own 3x3x3 feature sets."""
    from neuralce2_amd.lattice import bulk_fcc
    from neuralce2_amd.nce.model_io import ensemble_params, expand_params
    from oracle.nce_oracle import shells, energy_std, standardize
    mdl, _ = model
    feats = np.load(os.path.join(GOLDEN, "nce_features_subset.npz"), allow_pickle=False)
    keep = np.asarray(mdl["keep"])
    part = shells(mdl["primitive"], bulk_fcc("Ni", 3.924, cubic=True).repeat((3, 3, 3)), mdl["shell_edges"])
    Pe, mu, sd = ensemble_params(mdl)
    assert Pe["layers"][0]["self"].shape == (18, 48) and Pe["head"]["w"].shape == (48,)
    sets = [expand_params(mdl, s) for s in range(3)]
    for n in range(0, 121, 10):
        H = np.zeros((108, 18)); H[:, keep] = feats["H"][n]
        X = standardize(H, mu, sd)
        per_seed = [energy_std(P, X, part) for P, _, _ in sets]
        assert abs(energy_std(Pe, X, part) - np.mean(per_seed)) < 1e-12
        assert abs(energy_std(sets[0][0], X, part) * mdl["ys"] + mdl["ym"] - feats["pred"][n]) < 1e-12


def test_seed_ensemble_block_diagonal_for_deeper_layers():
    """two-layer networks: the fused second layer is block-diagonal and the energy is still the seed mean."""
    from neuralce2_amd.nce.model_io import ensemble_params
    from oracle.nce_oracle import layer
    rng = np.random.default_rng(0)
    C0, D, Pn, N = 5, 4, 2, 30

    def net():
        L1 = dict(self=rng.normal(size=(C0, D)), shell=rng.normal(size=(Pn, C0, D)), b=rng.normal(size=D),
                  gate=rng.normal(size=(Pn, C0, D)), gb=rng.normal(size=D))
        L2 = dict(self=rng.normal(size=(D, D)), shell=rng.normal(size=(Pn, D, D)), b=rng.normal(size=D),
                  gate=rng.normal(size=(Pn, D, D)), gb=rng.normal(size=D))
        return dict(lin=rng.normal(size=C0), layers=[L1, L2], head=dict(w=rng.normal(size=D), b=rng.normal()))
    nets = [net() for _ in range(3)]
    mdl_like = dict(params=nets)
    import neuralce2_amd.nce.model_io as mio
    orig = mio.expand_params
    mio.expand_params = lambda m, s: (m["params"][s], np.zeros(C0), np.ones(C0))
    try:
        Pe, _, _ = ensemble_params(mdl_like)
    finally:
        mio.expand_params = orig
    h0 = rng.normal(size=(N, C0))
    partners = [rng.integers(0, N, size=(N, 6)) for _ in range(Pn)]

    def E(P):
        h = h0
        for L in P["layers"]:
            h = layer(L, h, partners)
        return float((h0 @ P["lin"] + h @ P["head"]["w"]).sum() / N + float(P["head"]["b"]))
    assert abs(E(Pe) - np.mean([E(P) for P in nets])) < 1e-12
    assert np.all(Pe["layers"][1]["self"][:D, D:] == 0)


def test_unsupported_parameter_sets_are_refused(model):
    """a max-over-shell branch or a gate in some layers only must raise, not run with the term dropped."""
    from neuralce2_amd.nce.model_io import check_supported, expand_params
    mdl, _ = model
    P, _, _ = expand_params(mdl, 0)
    check_supported(P)
    L = dict(P["layers"][0]); L["shell_max"] = np.zeros_like(L["shell"])
    with pytest.raises(NotImplementedError, match="shell_max"):
        check_supported(dict(P, layers=[L]))
    L2 = {k: v for k, v in P["layers"][0].items() if k not in ("gate", "gb")}
    with pytest.raises(NotImplementedError, match="gate"):
        check_supported(dict(P, layers=[P["layers"][0], L2]))
    # a spin model (the shipped PtNi model marked as one) is refused by setup before anything is built
    from neuralce2_amd.lattice import read_cif
    from neuralce2_amd.nce.mc_model import setup
    at = read_cif(os.path.join(GOLDEN, "init_L10_N48.cif"))
    with pytest.raises(NotImplementedError, match="spin"):
        setup(dict(mdl, spin=1, n_spin=12), at)


def test_default_cache_mode_follows_the_reference_rule(model):
    """the cache layout is chosen as the reference's setup_implicit does (nce/mc_model.py:202-206): 'h0' while the
    layer-0 width C0 is at most the first layer's projection width (P shells, x2 with a gate, + 1) x D, else
    'proj'; no hidden layer: 'h0'.  Shipped PtNi model: C0 = 18 <= (3 x 2 + 1) x 16 = 112; the 3 x 64 SiLU
    network: 41 <= 2 x 64; a Fe-Ni-Cr-like width of 131 channels with the PtNi layer shapes: 131 > 112."""
    from neuralce2_amd.nce.mc_model import default_cache_mode
    from neuralce2_amd.nce.model_io import expand_params
    mdl, _ = model
    P, _, _ = expand_params(mdl, 0)
    assert len(np.ravel(P["lin"])) == 18 and default_cache_mode(P) == "h0"
    cs, _, Ps = _silu_setup()
    assert default_cache_mode(Ps) == "h0"
    rng = np.random.default_rng(0)
    L1 = dict(P["layers"][0])
    for k in ("self", "shell", "gate"):
        shp = list(np.shape(L1[k]))
        shp[-2] = 131
        L1[k] = rng.normal(size=shp)
    wide = dict(P, lin=rng.normal(size=131), layers=[L1])
    assert default_cache_mode(wide) == "proj"
    narrow = dict(wide, lin=rng.normal(size=112))
    assert default_cache_mode(narrow) == "h0"
    assert default_cache_mode(dict(P, layers=[])) == "h0"
