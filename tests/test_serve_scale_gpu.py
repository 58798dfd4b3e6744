"""The served forms at the sizes of the batched and streamed workloads.

* Batched uneven jobs: 512 PT jobs present at once (fcc n^3 cubic cells, n in 4..32 log-uniform in sites,
  x in 0.1..0.9, 16-64 replicas, cluster-expansion : NeuralCE (3-layer SiLU, BF16) = 3 : 1), two chunks each: every
  job's running totals equal a full recompute at retirement, and a sample of CE jobs is bit-identical alone.
* Served stream: Poisson arrivals including a 10,061,824-site job and a 256-temperature job, 10^6 steps each, with a
  memory budget that makes the large job wait for retirements: admission follows the budget and each CE job's
  result is the same alone and in the mix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def test_512_uneven_jobs_at_once(torch_dev):
    from neuralce2_amd.serve import Server, job_stream
    from tests.test_serve_gpu import _silu_factory
    specs = job_stream(512, 0.0, seed=3, n_min=4, n_max=32, t_min=16, t_max=64, n_steps=400, chunk=200, nce_every=4)
    assert sum(s.kind == "nce" for s in specs) == 128
    srv = Server(chunk=200, keep_occ=True, nn_precision="bf16", nce_model=_silu_factory())
    res, stats = srv.run(specs)
    print(f"512 jobs: {stats['wall_s']:.1f} s, {stats['attempts_per_s']:.3e} attempts/s, p50 "
          f"{stats['latency_p50_s']:.2f} s, p99 {stats['latency_p99_s']:.2f} s", flush=True)
    assert stats["n_jobs"] == 512 and stats["attempts"] == sum(s.n_steps * s.n_temps for s in specs)
    for s, r in zip(specs, res):
        assert r.job_id == s.job_id and np.all(np.isfinite(r.E_mean)) and np.all(r.C_v >= 0)
        # CE: FP32 swap energy changes summed in FP64; NeuralCE: BF16 projections of the per-step changes of the
        # synthetic SiLU network (rebuilt every ceil(N / 200) chunks), whose swaps move the energy by ~0.1-1 eV
        assert r.drift < (1e-8 if s.kind == "ce" else 1e-3), (s.job_id, s.kind, s.n, r.drift)
    ce = [s for s in specs if s.kind == "ce"]
    for s in (ce[0], ce[57], ce[200], max(ce, key=lambda q: q.n)):
        a, _ = srv.run([s])
        r = res[s.job_id]
        assert np.array_equal(a[0].occ, r.occ), s.job_id
        assert np.allclose(a[0].E_mean, r.E_mean, rtol=1e-12, atol=0), s.job_id


def test_served_stream_with_a_ten_million_site_job(torch_dev):
    from neuralce2_amd.serve import Server, JobSpec, _CEJob, job_stream
    specs = job_stream(10, 40.0, seed=5, n_min=4, n_max=24, t_min=16, t_max=64, n_steps=10 ** 6, chunk=20000)
    big = JobSpec(10, n=136, x=0.3, n_temps=16, n_steps=10 ** 6, seed=77, arrival=specs[2].arrival + 1e-3)
    wide = JobSpec(11, n=8, x=0.5, n_temps=256, n_steps=10 ** 6, seed=78, arrival=specs[3].arrival + 1e-3)
    specs = specs + [big, wide]
    budget = _CEJob.bytes_needed(big) + _CEJob.bytes_needed(wide)      # the big job cannot join a full house
    srv = Server(chunk=20000, mem_budget=budget, keep_occ=True)
    res, stats = srv.run(specs)
    print(f"stream: {stats['wall_s']:.1f} s, {stats['attempts_per_s']:.3e} attempts/s, p50 "
          f"{stats['latency_p50_s']:.2f} s, p99 {stats['latency_p99_s']:.2f} s", flush=True)
    by = {r.job_id: r for r in res}
    assert len(by) == 12 and by[10].n_sites == 10_061_824 and len(by[11].temps) == 256
    for r in res:
        # mixed form: FP32 swap energy changes in FP64 totals, re-anchored about every N^2 / 8 steps at chunk ends
        # (every 20,000-step chunk for the smallest cells)
        assert r.drift < 5e-9, (r.job_id, r.n_sites, r.drift)
    # admission by memory: no two jobs overlap while their bytes exceed the budget
    for r in res:
        live = [q for q in res if q.admitted <= r.admitted < q.finished]
        assert sum(_CEJob.bytes_needed(next(s for s in specs if s.job_id == q.job_id)) for q in live) <= budget
    for jid in (10, 11, 0):
        s = next(s for s in specs if s.job_id == jid)
        a, _ = srv.run([JobSpec(**{**s.__dict__, "arrival": 0.0})])
        assert np.array_equal(a[0].occ, by[jid].occ) and np.allclose(a[0].E_mean, by[jid].E_mean, rtol=1e-12, atol=0)
