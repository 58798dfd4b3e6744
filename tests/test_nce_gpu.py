"""GPU parity of the NeuralCE kernels (full cache, receptive-field local-update chunk) against the FP64 oracle.

Bars: FP64 energies within 1e-10 eV/atom of the oracle and the exact-enumeration fixture; occupations bit-exact under
injected proposals; caches after a chain of committed swaps equal a full rebuild (reference
tests/test_nce_local_update.py:27-53, 1e-9); FP32 caches within 2e-5 eV/atom."""
import json
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def ptni_nce():
    from neuralce2_amd.nce.model_io import load_model, expand_params
    from neuralce2_amd.clusterspace import ClusterSpace
    mdl = load_model(os.path.join(GOLDEN, "nipt_v3_nomax_sel.pkl"))
    cs = ClusterSpace(mdl["primitive"], mdl["cutoffs"], mdl["chemical_symbols"])
    P, mu, sd = expand_params(mdl, 0)
    return mdl, cs, P, mu, sd


def _cell(kind, seed=0):
    from neuralce2_amd.lattice import bulk_fcc, read_cif
    if kind == "L10_48":
        return read_cif(os.path.join(GOLDEN, "init_L10_N48.cif"))
    n = int(kind)
    at = bulk_fcc("Ni", 3.924, cubic=True).repeat((n, n, n))
    rng = np.random.default_rng(seed)
    Z = np.full(len(at), 28); Z[rng.choice(len(at), len(at) // 2, replace=False)] = 78
    at.numbers = Z
    return at


def _models(ptni_nce, at, kind="gate", readout="mean"):
    from neuralce2_amd.nce.mc_model import NNModel
    from oracle.nce_oracle import Oracle
    mdl, cs, P, mu, sd = ptni_nce
    nn = NNModel(cs, at, P, mu, sd, mdl["shell_edges"], ys=mdl["ys"], ym=mdl["ym"])
    o = Oracle(cs, at, mdl["shell_edges"], P, mu, sd)
    return nn, o


@pytest.mark.parametrize("cell", ["L10_48", "3", "4"])
def test_full_cache_energy(torch_dev, ptni_nce, cell):
    mdl = ptni_nce[0]
    at = _cell(cell)
    nn, o = _models(ptni_nce, at)
    sig_o = (at.numbers == 78).astype(np.int8)
    rng = np.random.default_rng(1)
    batch_o = np.stack([sig_o] + [rng.permutation(sig_o) for _ in range(3)])
    E_ref = np.array([o.energy(s) for s in batch_o]) * mdl["ys"] + mdl["ym"]
    if cell == "L10_48":
        assert abs(E_ref[0] - json.load(open(os.path.join(GOLDEN, "exact_N48.json")))["E_L10"]) < 1e-12
    perm = nn.geo["perm"]
    for prec, tol in ((0, 1e-10), (1, 2e-5)):
        _, E = nn.full_cache(nn.occ(batch_o[:, perm]), precision=prec)
        got = E.cpu().numpy() * mdl["ys"] + mdl["ym"]
        assert np.max(np.abs(got - E_ref)) < tol, (prec, got, E_ref)


def test_h0_cache_rows(torch_dev, ptni_nce):
    at = _cell("3", 4)
    nn, o = _models(ptni_nce, at)
    sig_o = (at.numbers == 78).astype(np.int8)
    cache, _ = nn.full_cache(nn.occ(sig_o[nn.geo["perm"]]), precision=0)
    C0 = len(ptni_nce[2]["lin"])
    h0 = cache[0, :nn.N * C0].reshape(nn.N, C0).cpu().numpy()
    ref = o.h0s(sig_o)[nn.geo["perm"]]
    assert np.max(np.abs(h0 - ref)) < 1e-12


@pytest.mark.parametrize("cell,prec,steps", [("3", 0, 120), ("4", 0, 80), ("3", 1, 120), ("L10_48", 0, 100)])
def test_chunk_injected_streams(torch_dev, ptni_nce, cell, prec, steps):
    import torch
    from neuralce2_amd.nce.mcmc import NNJobState
    mdl = ptni_nce[0]
    at = _cell(cell, 2)
    nn, o = _models(ptni_nce, at)
    perm, N = nn.geo["perm"], nn.N
    rng = np.random.default_rng(3)
    R = 4
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = (1.0 / (8.617333262e-5 * np.geomspace(400, 1600, R))).astype(np.float32).astype(np.float64)
    st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    ib[:, ::17] = ia[:, ::17]                                  # same-site proposals
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
    st.run(steps, inj)
    torch.cuda.synchronize()
    E = np.array([o.energy(s) for s in occ_o])
    occ_ref = occ_o.copy()
    nacc = o.run_chunk(occ_ref, E, betas, (perm[ia], perm[ib], u), nn.de_scale)
    got = st.occ.cpu().numpy()
    assert np.array_equal(got, occ_ref[:, perm])
    assert np.array_equal(st.n_accept.cpu().numpy(), nacc)
    tol = 1e-10 if prec == 0 else 2e-5 / mdl["ys"]
    assert np.max(np.abs(st.energy.cpu().numpy() - E)) < tol
    # caches after the chain == a full rebuild
    cache_chain = st.cache.clone()
    st.anchor()
    assert torch.max(torch.abs(cache_chain - st.cache)).item() < (1e-9 if prec == 0 else 1e-4)


def test_silu_three_layer_network(torch_dev):
    """the 12-neighbour SiLU variant (3 layers, width 64, summed readout) on 40 cluster channels + the zerolet."""
    import torch
    from neuralce2_amd.synthetic import fcc_cluster_space
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    from neuralce2_amd.lattice import bulk_fcc
    from oracle.nce_oracle import Oracle
    cs = fcc_cluster_space((9.0, 6.0, 5.0))
    assert len(cs) == 41
    P = silu_network(41, 64, 3, 1, seed=0)
    mu, sd = np.zeros(41), np.ones(41)
    at = bulk_fcc("Ni", 3.924, cubic=True).repeat((5, 5, 5))
    rng = np.random.default_rng(0)
    Z = np.full(len(at), 28); Z[rng.choice(len(at), len(at) // 2, replace=False)] = 78; at.numbers = Z
    edges = [0.0, 3.3]
    nn = NNModel(cs, at, P, mu, sd, edges, act="silu", readout="sum")
    o = Oracle(cs, at, edges, P, mu, sd, kind="silu", readout="sum")
    perm, N = nn.geo["perm"], nn.N
    R, steps = 2, 40
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = np.array([1.0 / (8.617333262e-5 * 900), 1.0 / (8.617333262e-5 * 1400)])
    for prec in (0, 1):
        st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
        E0 = np.array([o.energy(s) for s in occ_o])
        assert np.max(np.abs(st.energy.cpu().numpy() - E0)) < (1e-9 if prec == 0 else 1e-3)
        ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
        inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
        st.run(steps, inj)
        torch.cuda.synchronize()
        E = E0.copy(); occ_ref = occ_o.copy()
        o.run_chunk(occ_ref, E, betas, (perm[ia], perm[ib], u), 1.0)
        assert np.array_equal(st.occ.cpu().numpy(), occ_ref[:, perm])
        assert np.max(np.abs(st.energy.cpu().numpy() - E)) < (1e-9 if prec == 0 else 1e-3)


@pytest.mark.parametrize("kind", ["silu", "ptni"])
def test_bf16_mfma_energy_change(torch_dev, ptni_nce, kind, monkeypatch):
    """BF16 layers on MFMA (FP32 accumulation, FP32 caches): the energy change of one accepted swap per replica
    against the FP64 oracle's difference of full energies, within 3% of |dE| + 1e-5 (model units per site for the
    mean readout, totals for the summed one); after a chain of such steps the running total matches a full recompute
    within the same bound per step.  The PtNi model (gated, three shells, 12 used channels) takes the fused BF16
    layers by NCE_NN_FUSED=1 (by default its BF16 requests run the FP32 kernel, whose step fits in LDS)."""
    monkeypatch.setenv("NCE_NN_FUSED", "1")
    import torch
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    from neuralce2_amd.synthetic import fcc_cluster_space
    from oracle.nce_oracle import Oracle
    rng = np.random.default_rng(7)
    if kind == "silu":
        cs = fcc_cluster_space((9.0, 6.0, 5.0))
        P = silu_network(41, 64, 3, 1, seed=3)
        mu, sd = np.zeros(41), np.ones(41)
        at = _cell("5", 1)
        edges = [0.0, 3.3]
        nn = NNModel(cs, at, P, mu, sd, edges, act="silu", readout="sum")
        o = Oracle(cs, at, edges, P, mu, sd, kind="silu", readout="sum")
    else:
        at = _cell("4", 1)
        nn, o = _models(ptni_nce, at)
    perm, N = nn.geo["perm"], nn.N
    R = 16
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = np.full(R, 1.0 / (8.617333262e-5 * 1000.0))
    st = NNJobState(nn, occ_o[:, perm], betas, precision=2, sample_interval=1, samples_per_chunk=1)
    E0 = st.energy.cpu().numpy().copy()
    ia = np.zeros((R, 1), np.int64); ib = np.zeros((R, 1), np.int64)
    for r in range(R):                                   # a real swap (two different species) per replica
        a = int(rng.integers(N)); b = int(rng.choice(np.flatnonzero(occ_o[r, perm] != occ_o[r, perm][a])))
        ia[r, 0], ib[r, 0] = a, b
    u = np.zeros((R, 1))                                 # u = 0: every swap is accepted
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
    st.run(1, inj)
    torch.cuda.synchronize()
    dE = st.energy.cpu().numpy() - E0
    ref = []
    for r in range(R):
        s1 = occ_o[r].copy(); a, b = perm[ia[r, 0]], perm[ib[r, 0]]
        s1[a], s1[b] = s1[b], s1[a]
        ref.append(o.energy(s1) - o.energy(occ_o[r]))
    ref = np.array(ref)
    err = np.abs(dE - ref)
    print(f"BF16 single-step dE ({kind}): max relative error {np.max(err / np.maximum(np.abs(ref), 1e-12)):.2e}")
    assert np.all(err <= 3e-2 * np.abs(ref) + 1e-5), (err, ref)
    # chain: running total against a full recompute of the caches
    steps = 30
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    st2 = NNJobState(nn, occ_o[:, perm], betas, precision=2, sample_interval=1, samples_per_chunk=steps)
    st2.run(steps, tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u)))
    torch.cuda.synchronize()
    E_run = st2.energy.cpu().numpy().copy()
    st2.anchor()
    E_full = st2.energy.cpu().numpy()
    scale = np.max(np.abs(ref)) + 1e-5
    assert np.max(np.abs(E_run - E_full)) <= steps * 3e-2 * scale, (np.abs(E_run - E_full).max(), scale)
    print(kind, "max |dE err| / |dE|", float(np.max(err / (np.abs(ref) + 1e-12))), "chain drift", float(np.abs(E_run - E_full).max()))


def test_setup_local_delta_and_commit(torch_dev, ptni_nce):
    """setup(mdl, atoms): local_delta gives the oracle's energy difference without changing the state; commit applies
    the swap so that the carried caches equal a rebuild (the reference's setup()['local_delta'] / ['commit'])."""
    from neuralce2_amd.nce.mc_model import setup
    from oracle.nce_oracle import Oracle
    mdl, cs, P, mu, sd = ptni_nce
    at = _cell("3", 5)
    mc = setup(mdl, at, dtype="float64")
    o = Oracle(cs, at, mdl["shell_edges"], P, mu, sd)
    perm = mc["perm"]
    sig = mc["sig0"].astype(np.int8).copy()
    cache = mc["full_cache"](sig)
    rng = np.random.default_rng(9)
    for _ in range(6):
        i, j = mc["propose"](int(rng.integers(1 << 40)))
        if sig[i] == sig[j]:
            j = int(np.flatnonzero(sig != sig[i])[0])
        before = cache.cache.clone()
        dE, patch, sig1 = mc["local_delta"](cache, sig, [[i, sig[j]], [j, sig[i]]])
        assert cache.cache.equal(before)
        s_o = np.empty_like(sig); s_o[perm] = sig
        s1_o = np.empty_like(sig1); s1_o[perm] = sig1
        assert abs(dE - (o.energy(s1_o) - o.energy(s_o))) < 1e-10
        cache = mc["commit"](cache, patch, True)
        sig = sig1
        assert np.array_equal(cache.occ.cpu().numpy()[0], sig)
    carried = cache.cache.clone()
    cache.anchor()
    assert (carried - cache.cache).abs().max().item() < 1e-9


def _apart(nn, ia, ib):
    """True where the two flips' last-layer balls share no site on the cell: the cell vector ib - ia lies outside
    ball_L - ball_L (one flip's ball: the window dilated depth times by the partner offsets, sites deduplicated)."""
    d = np.array([int(x) for x in nn.geo["d"]], np.int64)
    ball = np.unique(nn.geo["window"][:, 1:].astype(np.int64) % d, axis=0)
    part = nn.geo["partners"][:, 1:].astype(np.int64)
    for _ in range(nn.depth):
        ball = np.unique(np.concatenate([ball, ((ball[:, None, :] + part[None]) % d).reshape(-1, 3)]), axis=0)
    diff = ((ball[:, None, :] - ball[None, :, :]) % d).reshape(-1, 3)
    near = np.zeros(nn.N, bool)
    near[(diff[:, 0] * d[1] + diff[:, 1]) * d[2] + diff[:, 2]] = True
    co = lambda x: np.stack([x // (d[1] * d[2]), (x // d[2]) % d[1], x % d[2]], -1)
    v = (co(ib) - co(ia)) % d
    return ~near[(v[..., 0] * d[1] + v[..., 1]) * d[2] + v[..., 2]]


@pytest.mark.parametrize("kind,n,prec,steps", [("ptni", 12, 0, 150), ("ptni", 12, 1, 150), ("ptni", 12, 2, 150),
                                               ("silu", 20, 0, 40), ("silu", 20, 1, 40)])
def test_fixed_balls_match_hash_path(torch_dev, ptni_nce, monkeypatch, kind, n, prec, steps):
    """Swaps whose balls share no site run on tables built once per model (slots, hits sorted by target); the rest
    build their balls in the LDS hash table.  The same injected stream with the tables switched off
    (NCE_NN_FIXED_BALLS=0, every step on the hash path, which the oracle tests above pin on small cells) gives the
    same occupations and acceptances and the same energies to the precision's bar, and the running totals and caches
    of the table run equal a full recompute (sizes where the oracle's brute-force placement is too slow)."""
    import torch
    from neuralce2_amd.lattice import bulk_fcc
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    at = bulk_fcc("Ni", 3.924, cubic=True).repeat((n, n, n))
    rng = np.random.default_rng(5)
    Z = np.full(len(at), 28); Z[rng.choice(len(at), len(at) // 2, replace=False)] = 78; at.numbers = Z
    if kind == "ptni":
        mdl, cs, P, mu, sd = ptni_nce
        make = lambda: NNModel(cs, at, P, mu, sd, mdl["shell_edges"], ys=mdl["ys"], ym=mdl["ym"])
        tol_e = 1e-10 if prec == 0 else 2e-5 / mdl["ys"]
    else:
        from neuralce2_amd.synthetic import fcc_cluster_space
        cs = fcc_cluster_space((9.0, 6.0, 5.0))
        P = silu_network(len(cs), 64, 3, 1, seed=0)
        make = lambda: NNModel(cs, at, P, np.zeros(len(cs)), np.ones(len(cs)), [0.0, 3.3], act="silu", readout="sum")
        tol_e = 1e-8 if prec == 0 else 1e-3
    monkeypatch.setenv("NCE_NN_FIXED_BALLS", "0")
    nn_hash = make()
    monkeypatch.delenv("NCE_NN_FIXED_BALLS")
    nn_fix = make()
    perm, N = nn_fix.geo["perm"], nn_fix.N
    R = 4 if kind == "ptni" else 2
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = (1.0 / (8.617333262e-5 * np.geomspace(500, 1500, R))).astype(np.float32).astype(np.float64)
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    apart = _apart(nn_fix, ia, ib) & (ia != ib)
    assert 10 < apart.sum() < apart.size - 10, apart.sum()       # both kinds of step occur
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
    runs = []
    for nn in (nn_hash, nn_fix):
        st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
        st.run(steps, inj)
        torch.cuda.synchronize()
        runs.append(st)
    a, b = runs
    assert np.array_equal(a.occ.cpu().numpy(), b.occ.cpu().numpy())
    assert np.array_equal(a.n_accept.cpu().numpy(), b.n_accept.cpu().numpy())
    assert np.max(np.abs(a.energy.cpu().numpy() - b.energy.cpu().numpy())) < tol_e
    if prec < 2:                 # BF16 projections drift from a full FP32 rebuild by design (re-anchored by drivers)
        cache_chain = b.cache.clone()
        E_run, E_full = b.anchor()
        assert torch.max(torch.abs(E_run - E_full)).item() < tol_e
        assert torch.max(torch.abs(cache_chain - b.cache)).item() < (1e-9 if prec == 0 else 1e-4)
    # the evaluation entry point (nce_nn_delta: the energy change of one swap per replica, nothing committed) on
    # both paths, each replica's swap one whose balls share no site
    occ_now = b.occ.cpu().numpy()
    cand = rng.integers(0, N, (R, 64, 2))
    ok = _apart(nn_fix, cand[..., 0], cand[..., 1]) & \
        (np.take_along_axis(occ_now, cand[..., 0], 1) != np.take_along_axis(occ_now, cand[..., 1], 1))
    assert ok.any(axis=1).all()
    pick = cand[np.arange(R), ok.argmax(axis=1)]
    inj1 = (torch.tensor(pick[:, :1], device=torch_dev), torch.tensor(pick[:, 1:], device=torch_dev),
            torch.zeros((R, 1), dtype=torch.float64, device=torch_dev))
    dEs = []
    for st in runs:
        occ_before = st.occ.clone()
        dE = torch.empty(R, dtype=torch.float64, device=torch_dev)
        st.nn.ctx.nn_delta([st.job(inj1)], dE, prec)
        torch.cuda.synchronize()
        assert torch.equal(st.occ, occ_before)
        dEs.append(dE.cpu().numpy())
    assert np.all(dEs[1] != 0.0)
    assert np.max(np.abs(dEs[0] - dEs[1])) < tol_e


@pytest.mark.parametrize("cell", ["4", "12"])
def test_lds_placements_agree(torch_dev, ptni_nce, monkeypatch, cell):
    """The per-step arrays (layer-0 change, sorted hits, projections) live in LDS where they fit and in the replica's
    scratch otherwise; NCE_NN_LDS_MODE_MAX caps how many go to LDS.  Every placement, with the gather form each one
    selects, gives the same occupations and acceptances as the default on the same injected stream, and energies to
    the precision's bar: FP64, FP32 and BF16, on a cell where every swap builds its balls in the hash table (256
    sites) and one where most use the fixed tables (6,912 sites)."""
    import torch
    from neuralce2_amd.nce.mc_model import NNModel
    from neuralce2_amd.nce.mcmc import NNJobState
    mdl, cs, P, mu, sd = ptni_nce
    at = _cell(cell, 7)
    models = {}
    for cap in (None, 0, 1, 2):
        if cap is None:
            monkeypatch.delenv("NCE_NN_LDS_MODE_MAX", raising=False)
        else:
            monkeypatch.setenv("NCE_NN_LDS_MODE_MAX", str(cap))
        models[cap] = NNModel(cs, at, P, mu, sd, mdl["shell_edges"], ys=mdl["ys"], ym=mdl["ym"])
    monkeypatch.delenv("NCE_NN_LDS_MODE_MAX", raising=False)
    N, perm = models[None].N, models[None].geo["perm"]
    rng = np.random.default_rng(11)
    R, steps = 4, 100
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = (1.0 / (8.617333262e-5 * np.geomspace(500, 1500, R))).astype(np.float32).astype(np.float64)
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
    for prec in (0, 1, 2):
        tol = 1e-10 if prec == 0 else 2e-5 / mdl["ys"]
        res = {}
        for cap, nn in models.items():
            st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
            st.run(steps, inj)
            torch.cuda.synchronize()
            res[cap] = (st.occ.cpu().numpy(), st.n_accept.cpu().numpy(), st.energy.cpu().numpy())
        ref = res[None]
        for cap in (0, 1, 2):
            assert np.array_equal(res[cap][0], ref[0]), (prec, cap)
            assert np.array_equal(res[cap][1], ref[1]), (prec, cap)
            assert np.max(np.abs(res[cap][2] - ref[2])) < tol, (prec, cap)


@pytest.mark.parametrize("net", ["ptni", "silu"])
def test_proj_cache_mode_equals_h0_mode(torch_dev, ptni_nce, monkeypatch, net):
    """The 'proj' cache layout (no layer-0 rows in the cache; a full recompute builds them in a temporary buffer,
    here at most two replicas at a time) runs the same chain as the 'h0' layout: three injected chunks of 50 steps,
    each followed by a full recompute, give identical occupations and acceptances and the same energies (1e-12
    relative) in FP64, FP32 and BF16, on the shipped PtNi model (gated, three shells, 256 sites) and the 3 x 64
    SiLU network (500 sites); the cache is N x C0 elements smaller."""
    import torch
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    monkeypatch.setenv("NCE_NN_H0_BATCH", "2")
    if net == "ptni":
        mdl, cs, P, mu, sd = ptni_nce
        at = _cell("4", 3)

        def mk(mode):
            return NNModel(cs, at, P, mu, sd, mdl["shell_edges"], ys=mdl["ys"], ym=mdl["ym"], cache_mode=mode)
    else:
        from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
        cs = fcc_cluster_space((9.0, 6.0, 5.0))
        P = silu_network(len(cs), 64, 3, 1, seed=2)
        at = fcc_cell(5, 0.5, 1)
        mu, sd = np.zeros(len(cs)), np.ones(len(cs))

        def mk(mode):
            return NNModel(cs, at, P, mu, sd, [0.0, 3.3], act="silu", readout="sum", cache_mode=mode)
    m_h0, m_pr = mk("h0"), mk("proj")
    N = m_h0.N
    C0 = len(np.ravel(P["lin"]))
    assert m_h0.cache_mode == "h0" and m_pr.cache_mode == "proj"
    assert m_h0.cache_elems - m_pr.cache_elems == (N * C0 + 3) // 4 * 4
    rng = np.random.default_rng(21)
    R, steps = 5, 50
    perm, inv = m_h0.geo["perm"], m_h0.geo["inv_perm"]
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    betas = (1.0 / (8.617333262e-5 * np.geomspace(500, 1500, R))).astype(np.float32).astype(np.float64)
    streams = [tuple(torch.tensor(x, device=torch_dev) for x in
                     (inv[rng.integers(0, N, (R, steps))], inv[rng.integers(0, N, (R, steps))], rng.random((R, steps))))
               for _ in range(3)]
    for prec in (0, 1, 2):
        res = {}
        for mode, nn in (("h0", m_h0), ("proj", m_pr)):
            st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
            fulls = []
            for inj in streams:
                st.run(steps, inj)
                _, E_full = st.anchor()
                fulls.append(E_full.cpu().numpy().copy())
            torch.cuda.synchronize()
            res[mode] = (st.occ.cpu().numpy(), st.n_accept.cpu().numpy(), np.stack(fulls))
        assert np.array_equal(res["h0"][0], res["proj"][0]), prec
        assert np.array_equal(res["h0"][1], res["proj"][1]), prec
        scale = np.abs(res["h0"][2]).max()
        assert np.max(np.abs(res["h0"][2] - res["proj"][2])) <= 1e-12 * scale, prec


def test_silu_network_16cubed_chain_matches_oracle(torch_dev):
    """The 16,384-site NeuralCE workload (fcc 16^3 cubic, x = 0.5, 3 SiLU layers of width 64 over the 12 nearest
    neighbours, 41 layer-0 channels, summed readout), 8 replicas, an injected chain of 8 steps: half of the
    proposals pair sites half a cell apart (their balls share no site: the fixed-table path), half are uniform.
    FP64: occupations bit-exact with the FP64 oracle's replay, running totals within 1e-9 eV.  FP32: the same
    decisions, totals within 1e-3 eV of the oracle energy of the final configuration (16,384 sites, |E| ~ 1.7e3 eV).
    BF16 projections: totals within 0.5 eV of the oracle energy of their own final configuration.  First-step energy
    changes without commit (nce_nn_delta): FP64 1e-9 eV, FP32 1e-5 relative, BF16 3 % + 1e-3 eV."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    from oracle.nce_oracle import Oracle, _site_lookup
    cs = fcc_cluster_space((9.0, 6.0, 5.0))
    P = silu_network(len(cs), 64, 3, 1, seed=0)
    mu, sd = np.zeros(len(cs)), np.ones(len(cs))
    at = fcc_cell(16, 0.5, 0)
    nn = NNModel(cs, at, P, mu, sd, [0.0, 3.3], act="silu", readout="sum")
    o = Oracle(cs, at, [0.0, 3.3], P, mu, sd, kind="silu", readout="sum")
    perm, inv, N = nn.geo["perm"], nn.geo["inv_perm"], nn.N
    R, steps = 8, 8
    rng = np.random.default_rng(12)
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    ia = rng.integers(0, N, (R, steps))
    far = _site_lookup(at)(at.positions[ia.ravel()] + at.cell.sum(0) / 2).reshape(R, steps)
    ib = np.where(np.arange(steps)[None, :] % 2 == 0, far, rng.integers(0, N, (R, steps)))
    u = rng.random((R, steps))
    betas = (1.0 / (8.617333262e-5 * np.geomspace(600, 2400, R))).astype(np.float32).astype(np.float64)
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (inv[ia], inv[ib], u))
    runs = {}
    for prec in (0, 1, 2):
        st = NNJobState(nn, occ_o[:, perm], betas, precision=prec, sample_interval=1, samples_per_chunk=steps)
        first = tuple(torch.tensor(x[:, :1].copy(), device=torch_dev) for x in (inv[ia], inv[ib], u))
        dE = torch.empty(R, dtype=torch.float64, device=torch_dev)
        nn.ctx.nn_delta([st.job(first)], dE, prec)
        st.run(steps, inj)
        torch.cuda.synchronize()
        runs[prec] = (st.occ.cpu().numpy()[:, inv], st.energy.cpu().numpy(), dE.cpu().numpy())

    def replay(r):
        occ = occ_o[r:r + 1].copy()
        E0 = np.array([o.energy(occ[0])])
        s1 = occ[0].copy(); s1[ia[r, 0]], s1[ib[r, 0]] = occ[0, ib[r, 0]], occ[0, ia[r, 0]]
        d1 = o.energy(s1) - E0[0]
        E = E0.copy()
        o.run_chunk(occ, E, betas[r:r + 1], (ia[r:r + 1], ib[r:r + 1], u[r:r + 1]), 1.0)
        return occ[0], E[0], d1
    with ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(replay, range(R)))
    occ_ref = np.stack([x[0] for x in ref]); E_ref = np.array([x[1] for x in ref]); d_ref = np.array([x[2] for x in ref])
    occ64, E64, d64 = runs[0]
    assert np.array_equal(occ64, occ_ref)
    assert np.max(np.abs(E64 - E_ref)) < 1e-9, np.abs(E64 - E_ref)
    assert np.max(np.abs(d64 - d_ref)) < 1e-9
    occ32, E32, d32 = runs[1]
    assert np.array_equal(occ32, occ_ref)
    assert np.max(np.abs(E32 - E_ref)) < 1e-3, np.abs(E32 - E_ref)
    assert np.all(np.abs(d32 - d_ref) <= 1e-5 * np.abs(d_ref) + 1e-6)
    occ16, E16, d16 = runs[2]
    with ThreadPoolExecutor(8) as ex:
        E16_ref = np.array(list(ex.map(o.energy, occ16)))
    moved = np.abs(d_ref) > 0
    print(f"16^3 SiLU: FP32 total error {np.abs(E32 - E_ref).max():.2e} eV, BF16 {np.abs(E16 - E16_ref).max():.2e} eV; "
          f"BF16 first-step dE relative error {np.max(np.abs(d16 - d_ref)[moved] / np.abs(d_ref)[moved]):.2e}; "
          f"{int(np.sum(np.all(occ16 == occ_ref, axis=1)))}/{R} BF16 chains on the FP64 path")
    assert np.max(np.abs(E16 - E16_ref)) < 0.5
    assert np.all(np.abs(d16 - d_ref) <= 0.03 * np.abs(d_ref) + 1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["silu16", "silu5"])
def test_split_replicas_equal_one_workgroup_per_replica(torch_dev, monkeypatch, kind):
    """BF16 layers with each replica over S = 2, 4 and 8 workgroups of one XCD (a cooperative launch; the parts
    exchange their input rows, neighbour tables and energy partials through the XCD's L2) against one workgroup per
    replica (NCE_NN_SPLIT=1), 8 replicas, the same injected chain of 12 steps: identical occupations, running totals
    within 1e-6 relative (the energy partials are summed in another grouping), and no NaN (a lost barrier returns
    NaN).  silu16: the 16,384-site cell, half of the proposals half a cell apart (fixed tables), half uniform (hash
    tables, built by part 0 and published); silu5: a 500-site cell, where every step builds its balls in the hash
    table."""
    import torch
    from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    from oracle.nce_oracle import _site_lookup
    cs = fcc_cluster_space((9.0, 6.0, 5.0))
    P = silu_network(len(cs), 64, 3, 1, seed=2)
    n = 16 if kind == "silu16" else 5
    at = fcc_cell(n, 0.5, 1)
    nn = NNModel(cs, at, P, np.zeros(len(cs)), np.ones(len(cs)), [0.0, 3.3], act="silu", readout="sum")
    inv, N = nn.geo["inv_perm"], nn.N
    R, steps = 8, 12
    rng = np.random.default_rng(5)
    sig_o = (at.numbers == 78).astype(np.int8)
    occ_o = np.stack([rng.permutation(sig_o) for _ in range(R)])
    ia = rng.integers(0, N, (R, steps))
    ib = rng.integers(0, N, (R, steps))
    if n % 2 == 0:                                       # (half of an odd cubic cell is not a lattice vector)
        far = _site_lookup(at)(at.positions[ia.ravel()] + at.cell.sum(0) / 2).reshape(R, steps)
        ib = np.where(np.arange(steps)[None, :] % 2 == 0, far, ib)
    u = rng.random((R, steps))
    betas = (1.0 / (8.617333262e-5 * np.geomspace(600, 2400, R))).astype(np.float32).astype(np.float64)
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (inv[ia], inv[ib], u))
    perm = nn.geo["perm"]
    runs = {}
    monkeypatch.setenv("NCE_NN_FUSED", "1")
    for S in (1, 2, 4, 8):
        monkeypatch.setenv("NCE_NN_SPLIT", str(S))
        st = NNJobState(nn, occ_o[:, perm], betas, precision=2, sample_interval=1, samples_per_chunk=steps)
        st.run(steps, inj)
        torch.cuda.synchronize()
        runs[S] = (st.occ.cpu().numpy(), st.energy.cpu().numpy(), st.traj.cpu().numpy())
    occ1, E1, T1 = runs[1]
    assert np.all(np.isfinite(E1))
    for S in (2, 4, 8):
        occ, E, T = runs[S]
        assert np.all(np.isfinite(E)), (S, E)
        assert np.array_equal(occ, occ1), S
        assert np.allclose(E, E1, rtol=1e-6, atol=1e-6), (S, np.abs(E - E1).max())
        assert np.allclose(T, T1, rtol=1e-6, atol=1e-6), S
