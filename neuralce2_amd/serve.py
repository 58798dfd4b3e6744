"""Served parallel-tempering jobs on one GPU: many PT jobs of uneven cell size, composition and temperature ladder.

A request is a C_v(T) curve: a cell (fcc n x n x n conventional cells, Pt fraction x), a temperature ladder and a number
of MC steps per replica.  The server keeps every admitted job's replica state resident in device memory and runs
continuous batching at chunk granularity:

  * admission: a request that has arrived is admitted when its device memory (occupations, caches, counters) fits in
    the budget (default: the smaller of 288 GB and 90 % of the free device memory); requests are taken in arrival
    order.  Its initial occupations are drawn on the device, and the FP64 energy of that one configuration is
    computed once for all its replicas;
  * one cluster-expansion tick = one chunk of `chunk` steps of every admitted CE job in ONE chunk launch (all
    (job, replica) pairs; the speculative kernel handles jobs of different cells together), the FP64 moment merge
    and the temperature-swap exchange of every CE job, captured once as a HIP graph (nce_pt_cycle) for the current
    set of jobs and replayed each tick with no host work in between; the host runs at most two ticks ahead of the
    device and rebuilds the graph only when a job joins, leaves or ends its equilibration;
  * the NeuralCE jobs run one nce_nn_run_chunk launch on a second stream, left in flight while the CE ticks
    continue; their caches are rebuilt every ceil(N / chunk) chunks (FP32 / BF16 drift) before the exchange;
  * a job that has done its steps retires: its per-slot <E> and <E^2> (FP64, every production step after its
    n_equilib equilibration steps) give E(T) and C_v(T), its memory is released, and waiting requests are admitted
    before the next tick.

Random numbers are Philox keyed by (job seed, replica, step) and the exchange by (job seed, exchange count), so a
cluster-expansion job's trajectory does not depend on which other jobs share its launches: a job gives the same result
alone and in the mix (the speculative kernel's decisions are the sequential chain's, whatever the tile).  NeuralCE
jobs match within the FP32 rounding of their sums (their per-step slot numbering comes from LDS atomics).

Job latency = retirement time - arrival time; the server reports p50 / p99 and the aggregate swap attempts/s.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

KB_CE = 8.617333e-5


@dataclass
class JobSpec:
    job_id: int
    n: int                      # fcc n x n x n conventional cells: 4 n^3 sites
    x: float                    # Pt fraction
    n_temps: int
    n_steps: int
    T_min: float = 300.0
    T_max: float = 1400.0
    seed: int = 0
    arrival: float = 0.0        # seconds after the start of the stream
    kind: str = "ce"            # "ce" (synthetic 8/6/5 cluster expansion) | "nce" (NeuralCE model of the server)
    n_equilib: int = 0          # leading steps left out of the moments (a whole number of chunks)


@dataclass
class JobResult:
    job_id: int
    kind: str
    n_sites: int
    temps: np.ndarray
    E_mean: np.ndarray          # eV / site, per temperature
    C_v: np.ndarray             # k_B / site, per temperature
    attempts: int
    acceptance: float
    arrival: float
    admitted: float
    finished: float
    occ: Optional[np.ndarray] = None   # final occupations by temperature slot (canonical order), on request
    drift: float = 0.0          # largest |running total - full recompute| at retirement (eV per site)

    @property
    def latency(self):
        return self.finished - self.arrival


def _ladder(spec):
    temps = np.array([spec.T_min]) if spec.n_temps == 1 else np.geomspace(spec.T_min, spec.T_max, spec.n_temps)
    return temps.astype(np.float32), (1.0 / (KB_CE * temps)).astype(np.float32).astype(np.float64)


def _initial_sigma(n_sites, x, seed, device=None):
    """random occupation with round(x N) Pt sites (species index 1), canonical order, from the job seed: a numpy
    array, or with `device` an int8 tensor drawn on that device (a seeded permutation on the GPU, on the current
    stream, so that even a 10^7-site cell costs the host nothing)."""
    k = int(round(x * n_sites))
    if device is None:
        rng = np.random.default_rng(seed)
        sig = np.zeros(n_sites, np.int8)
        sig[rng.choice(n_sites, k, replace=False)] = 1
        return sig
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    perm = torch.randperm(n_sites, generator=g, device=device)
    sig = torch.zeros(n_sites, dtype=torch.int8, device=device)
    sig[perm[:k]] = 1
    return sig


class _CEJob:
    def __init__(self, spec, gce, chunk):
        from .ce.pt_mcmc_ce import PTJobState
        self.spec = spec
        self.temps, betas = _ladder(spec)
        self.N = gce.n_atoms
        self.state = PTJobState(gce, _initial_sigma(self.N, spec.x, spec.seed, gce.device), betas, seed=spec.seed,
                                sample_interval=chunk, samples_per_chunk=1, job_id=spec.job_id)
        self.steps_done = 0
        self.equilibrating = spec.n_equilib > 0
        self.drift = 0.0
        self.anchor_every = 0                         # set by the server from the precision
        self.n_ticks = 0

    @staticmethod
    def bytes_needed(spec):
        N = 4 * spec.n ** 3
        return spec.n_temps * (N + 128) + 4096

    def result(self, keep_occ, verify=True):
        import torch
        st = self.state
        mean, var = st.moments()                      # per slot, total energies
        T = self.temps.astype(np.float64)
        E = mean / self.N
        cv = var / (KB_CE * T) ** 2 / self.N
        acc = float(st.n_accept.sum().item()) / max(1, self.steps_done * st.R)
        self.drift = float(st.drift.item()) / self.N
        if verify:
            full = torch.empty_like(st.energy)
            st.gce.ctx.ce_energy(st.gce.model, st.occ, full)
            self.drift = max(self.drift, float((full - st.energy).abs().max().item()) / self.N)
        occ = None
        if keep_occ:
            rep = st.rep_of_slot.cpu().numpy()
            occ = st.occ.cpu().numpy()[rep]
        return E, cv, acc, occ


class _NCEJob:
    """NeuralCE job: FP32 (or BF16-MFMA) local updates; caches and energies rebuilt every ceil(N / chunk) chunks
    (reference mcmc.py:360-366); the exchange permutes configurations on the host every chunk (mcmc.py:370-384)
    with draws from Philox keyed by (job seed, exchange count)."""

    def __init__(self, spec, nn, chunk, precision):
        from .nce.mcmc import NNJobState
        self.spec = spec
        self.nn = nn
        self.chunk = chunk
        self.temps, betas = _ladder(spec)
        self.betas = (1.0 / (8.617333262e-5 * self.temps.astype(np.float64))).astype(np.float32).astype(np.float64)
        self.N = nn.N
        self.state = NNJobState(nn, _initial_sigma(self.N, spec.x, spec.seed, nn.device), self.betas, seed=spec.seed,
                                precision=precision, sample_interval=1, samples_per_chunk=chunk)
        self.anchor_every = max(1, int(math.ceil(self.N / chunk)))
        self.n_chunks = 0
        self.n_ex = 0
        self.steps_done = 0
        self.acc_n = 0
        self.drift = 0.0
        self.s1 = np.zeros(spec.n_temps)
        self.s2 = np.zeros(spec.n_temps)

    @staticmethod
    def bytes_needed(spec, nn, precision):
        el = 8 if precision == 0 else 4
        return spec.n_temps * (nn.N + (nn.cache_elems + nn.scratch_elems) * el + 128) + 4096

    def after_chunk(self):
        """moments of the chunk's trajectory (after equilibration), the periodic rebuild, then the exchange."""
        from .nce.mcmc import exchange_draws, ladder_swaps
        self.n_chunks += 1
        self.steps_done += self.chunk
        if self.steps_done > self.spec.n_equilib:
            tr = self.state.traj.cpu().numpy()        # (chunk, R) model units, replica = temperature
            e = tr * self.nn.ys + self.nn.ym if self.nn.readout == "mean" else tr / self.N
            self.s1 += e.sum(0)
            self.s2 += (e * e).sum(0)
            self.acc_n += e.shape[0]
        if self.n_chunks % self.anchor_every == 0 or self.steps_done >= self.spec.n_steps:
            E_run, E_full = self.state.anchor()
            self.drift = max(self.drift, float((E_full - E_run).abs().max().item()) * self.nn.de_scale / self.N)
        R = self.state.R
        parity, u = exchange_draws(self.spec.seed, self.n_ex, R)
        self.n_ex += 1
        order, _, acc = ladder_swaps(self.state.energy.cpu().numpy() * self.nn.de_scale, self.betas, parity, u)
        if acc:
            self.state.permute(order)

    def result(self, keep_occ):
        n = max(self.acc_n, 1)
        mean = self.s1 / n
        var = np.maximum(self.s2 / n - mean * mean, 0.0)
        T = self.temps.astype(np.float64)
        cv = var * self.N / (8.617333262e-5 * T) ** 2
        acc = float(self.state.n_accept.sum().item()) / max(1, self.steps_done * self.state.R)
        occ = self.state.occ.cpu().numpy() if keep_occ else None
        return mean, cv, acc, occ


def autotune_tile(ctx, states, precision, candidates=(4, 8, 16), n_steps=768, repeats=2, stream=None):
    """Pick the speculative kernel's tile (steps of a replica evaluated in parallel) for this set of CE jobs by timing
    a short chunk per candidate on copies of the jobs' state (the jobs themselves do not move): the best of
    `repeats` timed launches after one untimed one.  Probes of a few hundred steps were too short: their timings
    differed by less than their noise and picked a tile 20 % slower over full chunks.  Everything runs and is timed
    on `stream` (the CE jobs' stream), so launches on other streams are not waited for."""
    import torch
    from .ce.pt_mcmc_ce import PTJobState
    stream = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(stream):
        copies = []
        for st in states:
            c = PTJobState.__new__(PTJobState)
            c.__dict__.update(st.__dict__)
            for name in ("occ", "energy", "beta_rep", "slot_of_rep", "step", "n_accept", "e_min", "e_ref", "s1",
                         "s2"):
                setattr(c, name, getattr(st, name).clone())
            c.traj = None
            copies.append(c)
        best, best_t = candidates[0], float("inf")
        for tile in candidates:
            jobs = [c.job(accumulate=False) for c in copies]
            ctx.pt_run_chunk(jobs, n_steps, precision, tile, stream)        # warm-up (code objects, caches)
            dt = float("inf")
            for _ in range(repeats):
                t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0.record(stream)
                ctx.pt_run_chunk(jobs, n_steps, precision, tile, stream)
                t1.record(stream)
                t1.synchronize()
                dt = min(dt, t0.elapsed_time(t1))
            if dt < best_t:
                best, best_t = tile, dt
    return best


class Server:
    """Continuous-batching PT server (see the module docstring)."""

    def __init__(self, chunk=2000, nn_chunk=None, precision="mixed", nn_precision="mixed", mem_budget=None,
                 tile=0, autotune=True, nce_model=None, ce_seed=0, keep_occ=False, verify=True, device=0):
        import torch
        from ._lib import Context
        if not torch.cuda.is_available():
            raise RuntimeError("the served PT loop needs a ROCm GPU (libneuralce_hip.so); none is visible")
        self.chunk = int(chunk)
        self.nn_chunk = int(nn_chunk or chunk)
        self.precision = {"fp64": 0, "mixed": 1}[precision]
        self.nn_precision = {"fp64": 0, "mixed": 1, "bf16": 2}[nn_precision]
        self.device = device
        self.ctx = Context.get(device)
        free, total = torch.cuda.mem_get_info(device)
        self.mem_budget = int(mem_budget) if mem_budget else int(min(288e9, 0.9 * free))
        self.tile = int(tile)
        self.autotune = autotune and tile == 0
        self.nce_model = nce_model          # callable(n) -> NNModel, for "nce" jobs
        self.ce_seed = ce_seed
        self.keep_occ = keep_occ
        self.verify = verify                # CE jobs: full recompute at retirement, reported as JobResult.drift
        self._ce_models = {}
        self._nn_models = {}
        self._tile_cache = {}
        self.stream_ce = torch.cuda.Stream(device)
        self.stream_nn = torch.cuda.Stream(device)

    def _gce(self, n):
        if n not in self._ce_models:
            from .ce.classical_ce_gpu import GPUClusterExpansion
            from .synthetic import synthetic_ce, fcc_cell
            at = fcc_cell(n, 0.5, 0)
            self._ce_models[n] = GPUClusterExpansion(synthetic_ce(self.ce_seed), at, energy_scale=float(len(at)),
                                                     device=self.device)
        return self._ce_models[n]

    def _nn(self, n):
        if n not in self._nn_models:
            if self.nce_model is None:
                raise ValueError("NeuralCE jobs need Server(nce_model=...)")
            self._nn_models[n] = self.nce_model(n)
        return self._nn_models[n]

    def prepare(self, specs):
        """build (and keep) the cell models the requests will need, before serving them: a server holds the models
        of the cell shapes it serves, so their host-side construction is not part of a request's latency."""
        for n in sorted({s.n for s in specs if s.kind == "ce"}):
            self._gce(n)
        for n in sorted({s.n for s in specs if s.kind != "ce"}):
            self._nn(n)

    def share_models(self, other):
        """use the cell models another server (same device) has built."""
        self._ce_models = other._ce_models
        self._nn_models = other._nn_models

    def _bytes(self, spec):
        if spec.kind == "ce":
            return _CEJob.bytes_needed(spec)
        return _NCEJob.bytes_needed(spec, self._nn(spec.n), self.nn_precision)

    def _tile_for(self, ce_jobs):
        if not self.autotune:
            return self.tile
        import torch
        n_items = sum(j.state.R for j in ce_jobs)
        n_cu = torch.cuda.get_device_properties(self.device).multi_processor_count
        # replicas per CU, precision and the smallest cell (window overlaps within a group grow as the cell shrinks)
        key = (min(4, (n_items + n_cu - 1) // n_cu), self.precision,
               int(math.log2(max(1, min(j.N for j in ce_jobs)))))
        if key not in self._tile_cache:
            self._tile_cache[key] = autotune_tile(self.ctx, [j.state for j in ce_jobs], self.precision,
                                                  stream=self.stream_ce)
        return self._tile_cache[key]

    def run(self, specs: List[JobSpec], log=None):
        """Serve the request stream; returns (results sorted by job id, stats)."""
        import torch
        from collections import deque
        from .ce.pt_mcmc_ce import anchor_interval_chunks
        pending = sorted(specs, key=lambda s: (s.arrival, s.job_id))
        for s in pending:
            ch = self.chunk if s.kind == "ce" else self.nn_chunk
            if s.n_steps % ch or s.n_equilib % ch or s.n_equilib >= s.n_steps:
                raise ValueError(f"job {s.job_id}: n_steps {s.n_steps} and n_equilib {s.n_equilib} must be whole "
                                 f"chunks of {ch}, with production steps left")
            if self._bytes(s) > self.mem_budget:
                raise ValueError(f"job {s.job_id} needs {self._bytes(s)} bytes, more than the budget")
        active_ce, active_nn = [], []
        results = []
        mem_used = 0
        attempts = 0
        t0 = time.perf_counter()
        n_ticks = 0
        nn_inflight = None                  # event after the NeuralCE chunk in flight on its stream
        ce_inflight = deque()               # (event, jobs that finish with this tick)
        cycle = None                        # captured tick of the current CE set
        old_cycles = []                     # destroyed once no launch is in flight (their frees would wait on it)
        lookahead = 2                       # CE ticks the host may run ahead of the device

        def finish(job, stream):
            nonlocal mem_used
            with torch.cuda.stream(stream):
                E, cv, acc, occ = (job.result(self.keep_occ, self.verify) if job.spec.kind == "ce"
                                   else job.result(self.keep_occ))
            fin = time.perf_counter() - t0
            results.append(JobResult(job.spec.job_id, job.spec.kind, job.N, job.temps.astype(np.float64), E, cv,
                                     job.steps_done * job.state.R, acc, job.spec.arrival, job.admitted, fin, occ,
                                     job.drift))
            mem_used -= job.mem
            if log:
                log(f"job {job.spec.job_id} ({job.spec.kind}, {job.N} sites, {job.state.R} T) done at {fin:.2f} s, "
                    f"latency {fin - job.spec.arrival:.2f} s")

        def drain_ce(keep):
            while len(ce_inflight) > keep:
                ev, done = ce_inflight.popleft()
                ev.synchronize()
                for j in done:
                    finish(j, self.stream_ce)

        while pending or active_ce or active_nn or ce_inflight:
            now = time.perf_counter() - t0
            changed = False
            # admission in arrival order while the memory fits; a NeuralCE job joins at the next NeuralCE launch
            while pending and pending[0].arrival <= now and mem_used + self._bytes(pending[0]) <= self.mem_budget:
                spec = pending[0]
                if spec.kind != "ce" and nn_inflight is not None:
                    break
                pending.pop(0)
                stream = self.stream_ce if spec.kind == "ce" else self.stream_nn
                with torch.cuda.stream(stream):
                    job = (_CEJob(spec, self._gce(spec.n), self.chunk) if spec.kind == "ce"
                           else _NCEJob(spec, self._nn(spec.n), self.nn_chunk, self.nn_precision))
                job.admitted = time.perf_counter() - t0
                job.mem = self._bytes(spec)
                if spec.kind == "ce":
                    job.anchor_every = anchor_interval_chunks(job.N, self.chunk, self.precision,
                                                              job.state.gce.fp32_step_error)
                mem_used += job.mem
                if spec.kind == "ce":
                    active_ce.append(job)
                    changed = True
                else:
                    active_nn.append(job)
            if not active_ce and not active_nn and not ce_inflight:
                if pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0))))
                continue
            # the NeuralCE jobs: one launch on their stream, left running while the CE ticks continue
            if nn_inflight is None and active_nn:
                with torch.cuda.stream(self.stream_nn):
                    self.ctx.nn_run_chunk([j.state.job() for j in active_nn], self.nn_chunk, self.nn_precision)
                    nn_inflight = torch.cuda.Event()
                    nn_inflight.record(self.stream_nn)
            # one CE tick: the captured cycle of the current set (chunk, moment merge, exchange)
            if active_ce:
                if changed or cycle is None:
                    if cycle is not None:
                        old_cycles.append(cycle)
                    tile = self._tile_for(active_ce)
                    # smallest cells first: their replicas take the longest per chunk (window overlaps, near
                    # pairs), so they are dispatched first and the larger cells' workgroups fill in behind them
                    order = sorted(active_ce, key=lambda j: j.N)
                    with torch.cuda.stream(self.stream_ce):
                        cycle = self.ctx.pt_cycle(
                            [j.state.job(accumulate=not j.equilibrating) for j in order], self.chunk, 1,
                            self.precision, tile, [j.state.exchange_job(None, reorder_last=True) for j in order])
                cycle.launch(1, self.stream_ce)
                n_ticks += 1
                ended, rebuild = [], False
                for j in active_ce:
                    j.steps_done += self.chunk
                    j.n_ticks += 1
                    attempts += self.chunk * j.state.R
                    if j.anchor_every and j.n_ticks % j.anchor_every == 0 and j.steps_done < j.spec.n_steps:
                        with torch.cuda.stream(self.stream_ce):
                            j.state.anchor()            # FP64 re-anchor between two replays, stream-ordered
                    if j.steps_done >= j.spec.n_steps:
                        ended.append(j)
                    elif j.equilibrating and j.steps_done >= j.spec.n_equilib:
                        j.equilibrating = False         # moments start from this tick's end
                        with torch.cuda.stream(self.stream_ce):
                            j.state.reset_moments()
                        rebuild = True
                ev = torch.cuda.Event()
                ev.record(self.stream_ce)
                ce_inflight.append((ev, ended))
                for j in ended:
                    active_ce.remove(j)
                if ended or rebuild:
                    if cycle is not None:
                        old_cycles.append(cycle)
                    cycle = None
                drain_ce(lookahead)
            elif ce_inflight:
                drain_ce(0)
            elif nn_inflight is not None:
                nn_inflight.synchronize()           # nothing else to run meanwhile
            # the NeuralCE launch, once finished: moments, rebuilds, exchange and retirements
            if nn_inflight is not None and nn_inflight.query():
                nn_inflight = None
                with torch.cuda.stream(self.stream_nn):
                    for j in active_nn:
                        j.after_chunk()
                        attempts += self.nn_chunk * j.state.R
                n_ticks += 1
                for j in [j for j in active_nn if j.steps_done >= j.spec.n_steps]:
                    self.stream_nn.synchronize()
                    finish(j, self.stream_nn)
                    active_nn.remove(j)
            if old_cycles and nn_inflight is None and not ce_inflight:
                for cy in old_cycles:
                    cy.close()
                old_cycles.clear()
        torch.cuda.synchronize(self.device)
        for cy in old_cycles + ([cycle] if cycle is not None else []):
            cy.close()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in results])
        stats = dict(wall_s=wall, attempts=int(attempts), attempts_per_s=attempts / wall, n_jobs=len(results),
                     n_ticks=n_ticks, latency_p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                     latency_p99_s=float(np.percentile(lat, 99)) if len(lat) else None)
        return sorted(results, key=lambda r: r.job_id), stats


def job_stream(n_jobs, rate, seed=0, n_min=4, n_max=32, t_min=16, t_max=64, n_steps=20000, chunk=2000,
               nce_every=0):
    """Synthetic request stream: Poisson arrivals at `rate` per second, cells log-uniform in sites between the n_min
    and n_max cubes, Pt fraction uniform in {0.1, ..., 0.9}, temperatures uniform in [t_min, t_max]; every
    `nce_every`-th request a NeuralCE job (0: none).  Seeded."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        lo, hi = math.log(4 * n_min ** 3), math.log(4 * n_max ** 3)
        n = int(round((math.exp(rng.uniform(lo, hi)) / 4) ** (1.0 / 3.0)))
        n = min(max(n, n_min), n_max)
        x = float(rng.integers(1, 10)) / 10.0
        nt = int(rng.integers(t_min, t_max + 1))
        kind = "nce" if (nce_every and i % nce_every == nce_every - 1) else "ce"
        steps = int(math.ceil(n_steps / chunk) * chunk)
        out.append(JobSpec(job_id=i, n=n, x=x, n_temps=nt, n_steps=steps, seed=1000 + i, arrival=t, kind=kind))
    return out


def run_sequential(specs, precision="fp64", chunk=2000, tile=0, keep_occ=False, nce_model=None, models_from=None,
                   verify=True):
    """The plain reference form: one job at a time (FP64 by default), same kernels, same per-job schedule;
    models_from: a server whose cell models to reuse."""
    srv = Server(chunk=chunk, precision=precision, tile=tile, autotune=False, keep_occ=keep_occ, nce_model=nce_model,
                 verify=verify)
    if models_from is not None:
        srv.share_models(models_from)
    res, wall, att = [], 0.0, 0
    for s in sorted(specs, key=lambda s: s.job_id):
        s0 = JobSpec(**{**s.__dict__, "arrival": 0.0})
        r, st = srv.run([s0])
        res += r
        wall += st["wall_s"]
        att += st["attempts"]
    return res, dict(wall_s=wall, attempts=att, attempts_per_s=att / wall if wall else None)
