"""GPU parity of the cluster-expansion kernels against the FP64 oracle (calls go through libneuralce_hip.so).

Bars: occupations bit-exact; FP64 energies and swap energy changes within 1e-9 eV; mixed precision (FP32 swap energy
change) within 1e-5 eV per swap; Metropolis decisions identical except where |beta dE + ln u| < 1e-6."""
import numpy as np
import pytest

from tests.conftest import random_fcc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _setup(ce, tpl, n, x=0.5, seed=1, scale=1.0, prim_reps=None):
    from neuralce2_amd.ce.ce_gpu_implicit import build_ce_gpu_implicit, encode_canonical
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    from oracle.ce_oracle import OracleCE
    if prim_reps is not None:
        from neuralce2_amd.lattice import bulk_fcc
        at = bulk_fcc("Ni", 3.924, cubic=False).repeat(prim_reps)
        r = np.random.RandomState(seed); Z = np.full(len(at), 28); Z[r.choice(len(at), len(at) // 2, replace=False)] = 78
        at.numbers = Z
    else:
        at = random_fcc(n, x=x, seed=seed)
    g = GPUClusterExpansion(ce, at, energy_scale=scale)
    o = OracleCE(g.info)
    sig = g.encode(at).astype(np.int8)
    return g, o, sig


CELLS = [dict(n=4), dict(n=8), dict(n=1), dict(n=2), dict(prim_reps=(2, 2, 2)), dict(prim_reps=(3, 1, 2))]


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: str(c))
def test_energy_matches_oracle(torch_dev, ptni_ce, ptni_template, cell):
    import torch
    g, o, sig = _setup(ptni_ce, ptni_template, cell.get("n", 0), prim_reps=cell.get("prim_reps"))
    rng = np.random.default_rng(3)
    R = 5
    batch = np.stack([rng.permutation(sig) for _ in range(R)])
    E_gpu = g.batched_energy(batch)
    E_ref = np.array([o.energy(b) for b in batch])
    assert np.max(np.abs(E_gpu - E_ref)) < 1e-9, (E_gpu, E_ref)


@pytest.mark.parametrize("cell", CELLS, ids=lambda c: str(c))
def test_swap_delta_matches_oracle(torch_dev, ptni_ce, ptni_template, cell):
    import torch
    from neuralce2_amd._lib import PREC_FP64, PREC_MIXED
    g, o, sig = _setup(ptni_ce, ptni_template, cell.get("n", 0), prim_reps=cell.get("prim_reps"))
    N = len(sig)
    rng = np.random.default_rng(7)
    R = 64
    occ = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    ia = rng.integers(0, N, R); ib = rng.integers(0, N, R)
    # near pairs: b inside a's window (clusters containing both sites)
    nb = g.info['win'][0, :, 1:]
    for k in range(0, R, 4):
        ib[k] = (ia[k] + rng.integers(1, 4)) % N
    ib[1] = ia[1]                                   # same site
    ref = np.array([o.delta(occ[r], ia[r], ib[r]) for r in range(R)])
    t_occ = torch.tensor(occ, device=torch_dev)
    t_ia = torch.tensor(ia, dtype=torch.int64, device=torch_dev); t_ib = torch.tensor(ib, dtype=torch.int64, device=torch_dev)
    dE = torch.empty(R, dtype=torch.float64, device=torch_dev)
    g.ctx.ce_delta(g.model, t_occ, t_ia, t_ib, dE, PREC_FP64)
    got = dE.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-9, np.max(np.abs(got - ref))
    g.ctx.ce_delta(g.model, t_occ, t_ia, t_ib, dE, PREC_MIXED)
    got32 = dE.cpu().numpy()
    assert np.max(np.abs(got32 - ref)) < 1e-5 * max(1.0, np.abs(ref).max()), np.max(np.abs(got32 - ref))


def _near_tie(o, occ_before_path, *a):
    return False


CHUNK_CASES = [(4, 8, 400, 0, 0), (8, 32, 300, 0, 0), (2, 4, 500, 0, 0), (4, 8, 400, 1, 0),
               (8, 8, 301, 0, 4), (8, 8, 302, 0, 8), (8, 8, 303, 1, 4), (8, 8, 305, 1, 8),
               (16, 6, 400, 0, 4), (16, 6, 400, 1, 8), (1, 4, 203, 0, 8), (2, 4, 203, 1, 4),
               (8, 8, 403, 1, 16), (4, 6, 301, 0, 16), (16, 4, 500, 1, 16)]


@pytest.mark.parametrize("n,R,steps,precision,tile", CHUNK_CASES)
def test_chunk_injected_streams_bitexact(torch_dev, ptni_ce, ptni_template, n, R, steps, precision, tile):
    """same injected proposals and uniforms -> identical occupations; energies within tolerance."""
    import torch
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    g, o, sig = _setup(ptni_ce, ptni_template, n)
    N = len(sig)
    rng = np.random.default_rng(11)
    occ0 = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    betas = 1.0 / (8.617333e-5 * np.geomspace(300, 1400, R).astype(np.float32))
    betas = betas.astype(np.float32).astype(np.float64)
    si = 10 if steps % 10 == 0 else 1
    st = PTJobState(g, occ0, betas, seed=5, sample_interval=si, samples_per_chunk=steps // si)
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    inj = (torch.tensor(ia, dtype=torch.int64, device=torch_dev), torch.tensor(ib, dtype=torch.int64, device=torch_dev),
           torch.tensor(u, dtype=torch.float64, device=torch_dev))
    E0 = st.energy.cpu().numpy().copy()
    g.ctx.pt_run_chunk([st.job(accumulate=True, inj=inj)], steps, precision, tile)
    torch.cuda.synchronize()
    occ_ref = occ0.copy(); E_ref = np.array([o.energy(x) for x in occ0])
    assert np.max(np.abs(E_ref - E0)) < 1e-9
    traj_ref = np.zeros((steps // si, R)); s1 = np.zeros(R); s2 = np.zeros(R)
    nacc_ref, emin_ref = o.run_chunk(occ_ref, E_ref, betas, steps, traj=traj_ref, sample_interval=si, accumulate=True,
                                     e_ref=E0.copy(), s1=s1, s2=s2, e_min=E0.copy(), inj=(ia, ib, u))
    occ_gpu = st.occ.cpu().numpy()
    tol = 1e-9 if precision == 0 else 1e-4
    mism = [r for r in range(R) if not np.array_equal(occ_gpu[r], occ_ref[r])]
    assert not mism, f"occupations differ for replicas {mism}"
    assert np.array_equal(st.n_accept.cpu().numpy(), nacc_ref)
    assert np.max(np.abs(st.energy.cpu().numpy() - E_ref)) < tol
    assert np.max(np.abs(st.traj.cpu().numpy() - traj_ref)) < tol
    # running total == full recompute (cells of one conventional cube hold clusters that contain a site twice
    # through a periodic image; the reference's swap delta counts those once per member position, so the identity
    # only holds from n = 2 on, for the reference and for this kernel alike)
    if n >= 2:
        E_full = np.array([o.energy(x) for x in occ_gpu])
        assert np.max(np.abs(st.energy.cpu().numpy() - E_full)) / N < 1e-9 * (1 if precision == 0 else 1e3)


@pytest.mark.parametrize("n,tile", [(4, 0), (8, 8), (8, 4), (16, 4), (16, 8), (8, 16)])
def test_chunk_philox_stream_matches_oracle_replay(torch_dev, ptni_ce, ptni_template, n, tile):
    """the kernel's own Philox proposals replayed by the oracle give the same trajectory."""
    import torch
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    g, o, sig = _setup(ptni_ce, ptni_template, n)
    R, steps = 16, 500
    rng = np.random.default_rng(2)
    occ0 = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    betas = (1.0 / (8.617333e-5 * np.geomspace(300, 1400, R))).astype(np.float32).astype(np.float64)
    st = PTJobState(g, occ0, betas, seed=123456789, sample_interval=50, samples_per_chunk=steps // 50)
    st.step.fill_(1000)
    g.ctx.pt_run_chunk([st.job()], steps, 0, tile)
    torch.cuda.synchronize()
    occ_ref = occ0.copy(); E_ref = np.array([o.energy(x) for x in occ0])
    nacc, _ = o.run_chunk(occ_ref, E_ref, betas, steps, seed=123456789, step0=1000)
    assert np.array_equal(st.occ.cpu().numpy(), occ_ref)
    assert np.array_equal(st.n_accept.cpu().numpy(), nacc)
    assert np.max(np.abs(st.energy.cpu().numpy() - E_ref)) < 1e-9
    assert np.all(st.step.cpu().numpy() == 1000 + steps)


@pytest.mark.parametrize("R,parity", [(8, 0), (8, 1), (7, 1), (32, -1), (2, 0)])
def test_exchange_matches_oracle(torch_dev, ptni_ce, ptni_template, R, parity):
    import torch
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    from oracle.ce_oracle import exchange
    g, o, sig = _setup(ptni_ce, ptni_template, 2)
    rng = np.random.default_rng(R)
    occ0 = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    betas = (1.0 / (8.617333e-5 * np.geomspace(300, 1400, R))).astype(np.float32).astype(np.float64)
    st = PTJobState(g, occ0, betas, seed=99, sample_interval=1, samples_per_chunk=1)
    E = st.energy.cpu().numpy()
    # make exchanges likely: energies spread over a few kT
    Eset = E.min() + rng.normal(0, 0.05, R)
    st.energy.copy_(torch.tensor(Eset, device=torch_dev))
    inj_u = None if parity < 0 else torch.tensor(rng.random(R), device=torch_dev)
    ej = st.exchange_job(17, inj_parity=parity, inj_u=inj_u)
    g.ctx.pt_exchange([ej])
    torch.cuda.synchronize()
    ros = np.arange(R, dtype=np.int32); brep = np.zeros(R)
    n = exchange(Eset, ros, betas, brep, seed=st.seed ^ 0x9E3779B97F4A7C15, ex_index=17, inj_parity=parity,
                 inj_u=None if inj_u is None else inj_u.cpu().numpy())
    assert np.array_equal(st.rep_of_slot.cpu().numpy(), ros)
    assert np.array_equal(st.beta_rep.cpu().numpy(), brep)
    assert int(st.n_swaps.item()) == n
    sor = np.empty(R, np.int32); sor[ros] = np.arange(R)
    assert np.array_equal(st.slot_of_rep.cpu().numpy(), sor)
    assert np.array_equal(st.traj[-1].cpu().numpy(), Eset[ros])


def test_chunk_large_cell_fast_path(torch_dev):
    """1,048,576-site cell at x = 0.25 (the benchmark workload, synthetic ECIs): a short Philox chunk of every
    replica replayed by the oracle; running totals equal full recomputes."""
    import torch
    from neuralce2_amd.synthetic import synthetic_ce, fcc_cell
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    from oracle.ce_oracle import OracleCE
    ce = synthetic_ce(0)
    at = fcc_cell(64, 0.25, 0)
    g = GPUClusterExpansion(ce, at, energy_scale=float(len(at)))
    o = OracleCE(g.info)
    sig = g.encode(at).astype(np.int8)
    R, steps = 4, 2000
    betas = (1.0 / (8.617333e-5 * np.geomspace(300, 1400, R))).astype(np.float32).astype(np.float64)
    for precision in (0, 1):
        st = PTJobState(g, sig, betas, seed=3 + precision, sample_interval=100, samples_per_chunk=steps // 100)
        g.ctx.pt_run_chunk([st.job()], steps, precision, 8)
        torch.cuda.synchronize()
        occ = np.ascontiguousarray(np.broadcast_to(sig, (R, len(sig))).copy())
        E0 = o.energy(sig); E = np.full(R, E0)
        o.run_chunk(occ, E, betas, steps, seed=3 + precision)
        got = st.occ.cpu().numpy()
        assert np.array_equal(got, occ)
        tol = 1e-9 if precision == 0 else 1e-6 * len(sig) * 1e-3
        assert np.max(np.abs(st.energy.cpu().numpy() - E)) < max(tol, 1e-9)
        Efull = g.batched_energy(got)
        assert np.max(np.abs(Efull - st.energy.cpu().numpy())) / len(sig) < 1e-9


@pytest.mark.parametrize("n,tile,precision", [(4, 1, 0), (8, 2, 0), (8, 4, 1), (16, 8, 0), (2, 1, 1)])
def test_general_kernel_injected_streams(torch_dev, ptni_ce, ptni_template, n, tile, precision, monkeypatch):
    """the general (any-basis, value-table) chunk kernel on the same parity bar."""
    import torch
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    monkeypatch.setenv("NCE_FORCE_GENERAL", "1")
    g, o, sig = _setup(ptni_ce, ptni_template, n)
    N = len(sig); R, steps = 6, 300
    rng = np.random.default_rng(21)
    occ0 = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    betas = (1.0 / (8.617333e-5 * np.geomspace(300, 1400, R))).astype(np.float32).astype(np.float64)
    st = PTJobState(g, occ0, betas, seed=5, sample_interval=10, samples_per_chunk=steps // 10)
    ia = rng.integers(0, N, (R, steps)); ib = rng.integers(0, N, (R, steps)); u = rng.random((R, steps))
    inj = tuple(torch.tensor(x, device=torch_dev) for x in (ia, ib, u))
    g.ctx.pt_run_chunk([st.job(inj=inj)], steps, precision, tile)
    torch.cuda.synchronize()
    occ_ref = occ0.copy(); E_ref = np.array([o.energy(x) for x in occ0])
    nacc, _ = o.run_chunk(occ_ref, E_ref, betas, steps, inj=(ia, ib, u))
    assert np.array_equal(st.occ.cpu().numpy(), occ_ref)
    assert np.array_equal(st.n_accept.cpu().numpy(), nacc)
    assert np.max(np.abs(st.energy.cpu().numpy() - E_ref)) < (1e-9 if precision == 0 else 1e-4)


def _pt_pair(g, sig, R, seed, sample_interval, samples_per_chunk, ladder_id=None):
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    rng = np.random.default_rng(seed)
    occ0 = np.stack([rng.permutation(sig) for _ in range(R)]).astype(np.int8)
    betas = (1.0 / (8.617333e-5 * np.geomspace(300, 1400, R))).astype(np.float32).astype(np.float64)
    mk = lambda: PTJobState(g, occ0, betas, ladder_id=ladder_id, seed=seed, sample_interval=sample_interval,  # noqa
                            samples_per_chunk=samples_per_chunk)
    return mk(), mk(), occ0, betas


def test_device_moment_merge_matches_the_trajectory(torch_dev, ptni_ce, ptni_template):
    """per-slot FP64 moments folded on the device after every chunk equal mean / variance of the every-step
    trajectory by slot, across exchanges that move replicas between slots."""
    import torch
    g, o, sig = _setup(ptni_ce, ptni_template, 4)
    R, chunk, n_chunks = 8, 200, 6
    st, _, _, _ = _pt_pair(g, sig, R, 31, 1, chunk)
    traj = []
    for c in range(n_chunks):
        g.ctx.pt_run_chunk([st.job(accumulate=True)], chunk, 0, 0)
        traj.append(st.traj.cpu().numpy().copy())
        g.ctx.pt_exchange([st.exchange_job(None, reorder_last=False)])
    torch.cuda.synchronize()
    e = np.concatenate(traj, 0)                              # (n_chunks * chunk, R) by slot
    mean, var = st.moments()
    assert np.all(st.acc_n.cpu().numpy() == n_chunks * chunk)
    assert int(st.ex_count.item()) == n_chunks and int(st.n_swaps.item()) > 0
    assert np.allclose(mean, e.mean(0), rtol=0, atol=1e-9 * np.abs(e).max())
    assert np.allclose(var, e.var(0), rtol=1e-9, atol=1e-12 * np.abs(e).max() ** 2)


def test_exchange_counter_on_the_device_equals_the_host_index(torch_dev, ptni_ce, ptni_template):
    import torch
    g, o, sig = _setup(ptni_ce, ptni_template, 2)
    a, b, _, _ = _pt_pair(g, sig, 12, 5, 1, 1)
    for k in range(1, 6):
        g.ctx.pt_exchange([a.exchange_job(k)])
        g.ctx.pt_exchange([b.exchange_job(None)])
        torch.cuda.synchronize()
        assert torch.equal(a.rep_of_slot, b.rep_of_slot) and torch.equal(a.beta_rep, b.beta_rep)
    assert int(b.ex_count.item()) == 5


@pytest.mark.parametrize("precision,n,ladders", [(0, 4, 1), (1, 8, 1), (1, 4, 2)])
def test_captured_cycle_replays_equal_eager_launches(torch_dev, ptni_ce, ptni_template, precision, n, ladders):
    """a HIP-graph cycle (2 chunks + merges + exchange) replayed three times, with eager work and synchronisation
    in between, leaves exactly the state of the same launches issued one by one, after every replay."""
    import torch
    g, o, sig = _setup(ptni_ce, ptni_template, n)
    R, chunk = 16, 300
    lid = None if ladders == 1 else np.repeat(np.arange(ladders), R // ladders).astype(np.int32)
    a, b, _, _ = _pt_pair(g, sig, R, 77, 100, chunk // 100, ladder_id=lid)
    cy = g.ctx.pt_cycle([a.job(accumulate=True)], chunk, 2, precision, 0, [a.exchange_job(None)])
    names = ("occ", "energy", "step", "n_accept", "e_min", "rep_of_slot", "slot_of_rep", "beta_rep", "acc_n",
             "acc_mean", "acc_m2", "s1", "s2", "e_ref", "ex_count", "n_swaps", "traj")
    for rep in range(3):
        cy.launch(1)
        for _ in range(2):
            g.ctx.pt_run_chunk([b.job(accumulate=True)], chunk, precision, 0)
        g.ctx.pt_exchange([b.exchange_job(None)])
        scratch = torch.randn(1 << 16, device=torch_dev).sum()       # eager work between the replays
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        for nm in names:
            assert torch.equal(getattr(a, nm), getattr(b, nm)), (rep, nm)
    assert int(a.ex_count.item()) == 3
    cy.close()


def test_cycle_refuses_what_it_cannot_replay(torch_dev, ptni_ce, ptni_template):
    import torch
    from neuralce2_amd._lib import NCEError
    g, o, sig = _setup(ptni_ce, ptni_template, 2)
    a, _, _, _ = _pt_pair(g, sig, 4, 1, 10, 1)
    with pytest.raises(NCEError, match="ex_count"):
        g.ctx.pt_cycle([a.job()], 10, 1, 0, 0, [a.exchange_job(3)])
    inj = tuple(torch.zeros((4, 10), dtype=t, device=torch_dev) for t in (torch.int64, torch.int64, torch.float64))
    with pytest.raises(NCEError, match="injected"):
        g.ctx.pt_cycle([a.job(inj=inj)], 10, 1, 0, 0, [])


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_pt_statistics_match_exact_enumeration(torch_dev, precision):
    """The CE parallel-tempering sampler against the exact thermodynamics of a 24-site cell (every one of the
    C(24, 12) configurations enumerated by the oracle): <E>(T) within 3e-4 eV/site and C_v(T) within 6 % at 12
    temperatures 250-2000 K, from 3 x 10^5 steps per replica (the first tenth dropped; a few standard errors)."""
    from neuralce2_amd.ce_io import ClusterExpansion
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState, anchor_interval_chunks
    from oracle.ce_oracle import brute_terms, enumerate_energies
    from tests.test_ce_cpu import _exact_cell
    cs, at, params = _exact_cell()
    N = len(at)
    sites, coef, e0 = brute_terms(cs, params, at, float(N))
    E_all = enumerate_energies(sites, coef, e0, N, 12)
    temps = np.geomspace(250.0, 2000.0, 12).astype(np.float32)
    T64 = temps.astype(np.float64)
    ex_E, ex_C = [], []
    for T in T64:
        b = 1.0 / (8.617333e-5 * T)
        w = np.exp(-b * (E_all - E_all.min()))
        m1 = (w * E_all).sum() / w.sum(); m2 = (w * E_all ** 2).sum() / w.sum()
        ex_E.append(m1 / N); ex_C.append((m2 - m1 * m1) * b * b / N)
    ex_E, ex_C = np.array(ex_E), np.array(ex_C)
    Z = np.full(N, 28); Z[np.random.default_rng(4).choice(N, 12, replace=False)] = 78; at.numbers = Z
    g = GPUClusterExpansion(ClusterExpansion(cs, params), at, energy_scale=float(N))
    sig = g.encode(at).astype(np.int8)
    betas = (1.0 / (8.617333e-5 * T64)).astype(np.float32).astype(np.float64)
    st = PTJobState(g, sig, betas, seed=31, sample_interval=1000, samples_per_chunk=1)
    chunk, n_chunks = 1000, 300
    every = anchor_interval_chunks(N, chunk, precision, g.fp32_step_error)
    for c in range(n_chunks):
        g.ctx.pt_run_chunk([st.job(accumulate=c >= n_chunks // 10)], chunk, precision, 0)
        g.ctx.pt_exchange([st.exchange_job(None)])
        if every and (c + 1) % every == 0:
            st.anchor()
        if c + 1 == n_chunks // 10:
            st.reset_moments()
    mean, var = st.moments()
    E_mc = mean / N; C_mc = var / (8.617333e-5 * T64) ** 2 / N
    dE = np.abs(E_mc - ex_E); dC = np.abs(C_mc / ex_C - 1)
    print(f"precision {precision}: max |<E> - exact| {dE.max():.2e} eV/site, C_v ratio "
          f"{np.round(C_mc / ex_C, 3).tolist()}")
    assert dE.max() < 3e-4 and dC.max() < 0.06, (dE, C_mc / ex_C)


@pytest.mark.parametrize("n,x", [(16, 0.3), (64, 0.25), (3, 0.5)])
def test_parity_energy_kernel_equals_forward_rows(torch_dev, n, x, monkeypatch):
    """full energies of the binary fast path (window-parity sums with coefficients / order) against the
    forward-row kernel (every cluster of every cell), on cells of 108 to 1,048,576 sites."""
    import time
    import torch
    from neuralce2_amd.synthetic import synthetic_ce, fcc_cell
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    at = fcc_cell(n, x, 2)
    g = GPUClusterExpansion(synthetic_ce(0), at, energy_scale=float(len(at)))
    sig = g.encode(at).astype(np.int8)
    rng = np.random.default_rng(n)
    R = 8
    occ = torch.tensor(np.stack([rng.permutation(sig) for _ in range(R)]), device=torch_dev)
    E_fast = torch.empty(R, dtype=torch.float64, device=torch_dev)
    E_rows = torch.empty_like(E_fast)
    for _ in range(2):                                   # second round timed (code objects loaded)
        torch.cuda.synchronize(); t0 = time.perf_counter()
        g.ctx.ce_energy(g.model, occ, E_fast)
        torch.cuda.synchronize(); t_fast = time.perf_counter() - t0
        monkeypatch.setenv("NCE_CE_ENERGY_ROWS", "1")
        torch.cuda.synchronize(); t0 = time.perf_counter()
        g.ctx.ce_energy(g.model, occ, E_rows)
        torch.cuda.synchronize(); t_rows = time.perf_counter() - t0
        monkeypatch.delenv("NCE_CE_ENERGY_ROWS")
    print(f"{len(at)} sites x {R}: parity kernel {t_fast * 1e3:.2f} ms, forward rows {t_rows * 1e3:.2f} ms")
    assert torch.max(torch.abs(E_fast - E_rows)).item() / len(at) < 1e-12
