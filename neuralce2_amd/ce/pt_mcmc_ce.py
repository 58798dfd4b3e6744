"""Parallel-tempering swap Monte Carlo for the cluster expansion on the gfx950 chunk kernel.

Drop-in for neuralce/ce/pt_mcmc_ce.py: ``MCMCConfig`` (:37-104), ``make_temperature_ladder`` (:223-229),
``_build_interleaved_ladders`` (:232-260), ``MCMCResult`` (:587-598) and ``run_pt_mcmc`` (:601-1028) with the same
chunk / exchange arithmetic, the same checkpoint keys and the same output files.  The step loop is one
``nce_pt_run_chunk`` launch per chunk (every replica of the job, every step of the chunk, then the FP64 moment
merge on the device); the temperature-swap exchange is ``nce_pt_exchange`` at chunk boundaries, its counter kept on
the device.

Differences from the reference, by design:
  * random streams: counter-based Philox keyed by (seed, replica, global step) instead of JAX threefry keys split per
    slot, so trajectories are reproducible per replica independent of batching (a given --seed does not reproduce
    the JAX trajectory; parity is checked with injected streams);
  * FP64 accumulators of E and E^2 per temperature slot over every production step (heat capacity without
    subsampling), written next to the reference's outputs;
  * precision: 'fp64' (plain form) or 'mixed' (FP32 swap energy change, FP64 totals, exchange and accumulators).
"""
from __future__ import annotations

import math
import os
import sys
import time
import faulthandler
import signal
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_KB_EV_PER_K = 8.617333e-5          # CE path constant (pt_mcmc_ce.py:220)


@dataclass
class MCMCConfig:
    ce_path: str = ''
    supercell_matrix: list = field(default_factory=list)
    T_min: float = 100.0
    T_max: float = 2000.0
    n_replicas: int = 64
    n_steps: int = 100000
    n_equilib: int = 10000
    sample_interval: int = 100
    exchange_interval: int = 10
    chunk_size: int = 1000
    spin_orbits_pkl: str = ''
    spin_states: dict = field(default_factory=dict)
    spin_move_prob: float = 0.0
    initial_composition: dict = field(default_factory=dict)
    swap_pairs: Optional[list] = None
    initial_sigma: Optional[np.ndarray] = None
    n_ladders: int = 1
    seed: int = 42
    output_dir: str = 'mcmc_results'
    energy_mode: str = 'full'
    checkpoint_every: int = 0
    T_grid: Optional[np.ndarray] = None
    run_id: str = ''
    precision: str = 'fp64'            # 'fp64' | 'mixed'
    tile: int = 0                      # waves per replica (0 = default)


@dataclass
class MCMCResult:
    energies: np.ndarray
    temperatures: np.ndarray
    swap_accept_rate: float
    flip_accept_rate: float
    exchange_accept_rate: float
    sigma_final: np.ndarray
    sigma_min: np.ndarray
    E_min: float
    spins_final: Optional[np.ndarray]
    wall_time: float
    E_mean: Optional[np.ndarray] = None        # per temperature slot, every production step
    E_var: Optional[np.ndarray] = None
    attempts_per_s: float = 0.0


def make_temperature_ladder(T_min, T_max, n_replicas):
    """geomspace ladder; temperatures and betas cast to float32 as the reference does (pt_mcmc_ce.py:223-229)."""
    temps = np.array([T_min]) if n_replicas == 1 else np.geomspace(T_min, T_max, n_replicas)
    betas = 1.0 / (_KB_EV_PER_K * temps)
    return temps.astype(np.float32), betas.astype(np.float32)


def _build_interleaved_ladders(T_min, T_max, n_per_ladder, n_ladders):
    """replica i -> ladder i // n_per_ladder, position i % n_per_ladder, T = global geomspace[ladder + pos * n_ladders]
    (pt_mcmc_ce.py:232-260)."""
    n_total = n_per_ladder * n_ladders
    t_global = np.array([T_min], np.float64) if n_total == 1 else np.geomspace(T_min, T_max, n_total)
    i = np.arange(n_total)
    ladder_id = (i // n_per_ladder).astype(np.int32); ladder_pos = (i % n_per_ladder).astype(np.int32)
    temps = t_global[ladder_id.astype(np.int64) + ladder_pos.astype(np.int64) * n_ladders].astype(np.float32)
    betas = (1.0 / (_KB_EV_PER_K * temps)).astype(np.float32)
    return temps, betas, ladder_id, ladder_pos


def make_initial_state(seed, gce, initial_composition, n_replicas):
    """random configuration per active sublattice at the requested fractions (uniform if absent), one per replica
    (reference make_initial_state pt_mcmc_ce.py:266-330 draws with a numpy Generator seeded from the JAX key)."""
    rng = np.random.default_rng(seed)
    N = gce.n_atoms
    out = np.zeros((n_replicas, N), np.int8)
    for r in range(n_replicas):
        sig = np.zeros(N, np.int8)
        for sl in gce.sublattices:
            idx = sl['indices']; ns = sl['n_species']; name = sl['name']
            fr = initial_composition.get(name) if initial_composition else None
            fr = np.full(ns, 1.0 / ns) if fr is None else np.asarray(fr, float) / np.sum(fr)
            counts = np.floor(fr * len(idx)).astype(int); counts[0] += len(idx) - counts.sum()
            arr = np.repeat(np.arange(ns, dtype=np.int8), counts)
            rng.shuffle(arr)
            sig[idx] = arr
        out[r] = sig
    return out


class PTJobState:
    """Device-resident state of one PT job: occupations (R, N) int8, FP64 energies, ladder, counters, accumulators.

    The per-slot FP64 moments (count, mean, sum of squared deviations of the total energy over every accumulated
    step) and the exchange counter live on the device: a chunk launch folds its moments in and an exchange advances
    the counter, so consecutive chunks need no host round trip.  The same object serves the single-job driver below
    and the multi-job scheduler (neuralce2_amd.serve)."""

    def __init__(self, gce, sigma0, betas_slot, ladder_id=None, seed=0, sample_interval=1, samples_per_chunk=0,
                 job_id=0):
        import torch
        from .._lib import PTJob, ExchangeJob
        dev = gce.device
        self.gce = gce
        self.R = R = len(betas_slot)
        self.N = gce.n_atoms
        self.energy = torch.empty(R, dtype=torch.float64, device=dev)
        if torch.is_tensor(sigma0):                   # one configuration already on the device (taken as given)
            if tuple(sigma0.shape) != (self.N,):
                raise ValueError(f"initial sigma shape {tuple(sigma0.shape)} != ({self.N},)")
            sig = None
        else:
            sig = np.asarray(sigma0)
            if sig.shape != (R, self.N) and sig.shape != (self.N,):
                raise ValueError(f"initial sigma shape {sig.shape} != ({R}, {self.N}) or ({self.N},)")
            if sig.min() < 0 or sig.max() > 1:
                raise ValueError("species index out of range for a binary cluster expansion")
        if sig is None or sig.ndim == 1:
            # every replica starts from the same configuration: one row uploaded and replicated on the device, its
            # FP64 energy computed once
            row = (sigma0.to(device=dev, dtype=torch.int8) if sig is None
                   else torch.tensor(np.ascontiguousarray(sig, np.int8), device=dev)).unsqueeze(0)
            self.occ = row.expand(R, self.N).contiguous()
            gce.ctx.ce_energy(gce.model, row, self.energy[:1])
            self.energy[1:] = self.energy[:1]
        else:
            self.occ = torch.tensor(np.ascontiguousarray(sig, np.int8), device=dev)
            gce.ctx.ce_energy(gce.model, self.occ, self.energy)
        self.beta_slot = torch.tensor(np.asarray(betas_slot, np.float64), device=dev)
        self.beta_rep = self.beta_slot.clone()
        self.ladder_id = None if ladder_id is None else torch.tensor(np.asarray(ladder_id, np.int32), device=dev)
        self.rep_of_slot = torch.arange(R, dtype=torch.int32, device=dev)
        self.slot_of_rep = torch.arange(R, dtype=torch.int32, device=dev)
        self.step = torch.zeros(R, dtype=torch.int64, device=dev)
        self.n_accept = torch.zeros(R, dtype=torch.int64, device=dev)
        self.e_min = self.energy.clone()
        self.e_ref = self.energy.clone()
        self.s1 = torch.zeros(R, dtype=torch.float64, device=dev)
        self.s2 = torch.zeros(R, dtype=torch.float64, device=dev)
        self.acc_n = torch.zeros(R, dtype=torch.float64, device=dev)
        self.acc_mean = torch.zeros(R, dtype=torch.float64, device=dev)
        self.acc_m2 = torch.zeros(R, dtype=torch.float64, device=dev)
        self.n_swaps = torch.zeros(1, dtype=torch.int64, device=dev)
        self.ex_count = torch.zeros(1, dtype=torch.int64, device=dev)
        self.drift = torch.zeros((), dtype=torch.float64, device=dev)   # largest |full - running| at a re-anchor
        self.sample_interval = int(sample_interval)
        self.traj = torch.zeros((max(samples_per_chunk, 1), R), dtype=torch.float64, device=dev) if samples_per_chunk else None
        self.seed = int(seed) & ((1 << 64) - 1)
        self.job_id = job_id
        self._PTJob, self._ExchangeJob = PTJob, ExchangeJob

    def job(self, accumulate=False, inj=None, merge=True):
        """chunk-launch descriptor; accumulate adds every step's energy to the chunk moments, merge (with
        accumulate) folds them into the per-slot accumulators on the device after the chunk."""
        from .._lib import ptr
        j = self._PTJob()
        j.model = self.gce.model; j.n_replicas = self.R; j.n_sites = self.N
        j.occ = ptr(self.occ); j.energy = ptr(self.energy); j.beta = ptr(self.beta_rep)
        j.slot_of_rep = ptr(self.slot_of_rep); j.step = ptr(self.step); j.seed = self.seed
        j.n_accept = ptr(self.n_accept); j.e_min = ptr(self.e_min)
        j.traj = ptr(self.traj) if self.traj is not None else None
        j.sample_interval = self.sample_interval
        j.accumulate = int(bool(accumulate)); j.e_ref = ptr(self.e_ref); j.s1 = ptr(self.s1); j.s2 = ptr(self.s2)
        if accumulate and merge:
            j.acc_n = ptr(self.acc_n); j.acc_mean = ptr(self.acc_mean); j.acc_m2 = ptr(self.acc_m2)
        if inj is not None:
            self._inj = inj
            j.inj_ia = ptr(inj[0]); j.inj_ib = ptr(inj[1]); j.inj_u = ptr(inj[2])
        return j

    def exchange_job(self, ex_index=None, reorder_last=True, inj_parity=-1, inj_u=None):
        """exchange descriptor; ex_index None takes the device counter (exchange number = count + 1)."""
        from .._lib import ptr
        e = self._ExchangeJob()
        e.n_replicas = self.R; e.inj_parity = int(inj_parity)
        e.energy = ptr(self.energy); e.rep_of_slot = ptr(self.rep_of_slot); e.slot_of_rep = ptr(self.slot_of_rep)
        e.beta_slot = ptr(self.beta_slot); e.ladder_id = ptr(self.ladder_id); e.beta_rep = ptr(self.beta_rep)
        e.seed = self.seed ^ 0x9E3779B97F4A7C15; e.n_swaps = ptr(self.n_swaps)
        if ex_index is None:
            e.ex_count = ptr(self.ex_count)
        else:
            e.ex_index = int(ex_index)
        if inj_u is not None:
            self._inj_u = inj_u
        e.inj_u = ptr(inj_u)
        e.traj_last = ptr(self.traj[-1]) if (reorder_last and self.traj is not None) else None
        return e

    def anchor(self):
        """FP64 re-anchor at a chunk boundary: the running totals are replaced by a full recompute (and the moments'
        reference with them); the largest correction so far is kept on the device in `drift`."""
        import torch
        full = torch.empty_like(self.energy)
        self.gce.ctx.ce_energy(self.gce.model, self.occ, full)
        torch.maximum(self.drift, (full - self.energy).abs().max(), out=self.drift)
        self.energy.copy_(full)
        self.e_ref.copy_(full)

    def reset_moments(self):
        """start the accumulated moments afresh (end of equilibration)."""
        self.acc_n.zero_(); self.acc_mean.zero_(); self.acc_m2.zero_()
        self.s1.zero_(); self.s2.zero_()
        self.e_ref.copy_(self.energy)

    def moments(self):
        """per temperature slot: mean and variance (population) of the total energy over the accumulated steps."""
        n = self.acc_n.cpu().numpy()
        return self.acc_mean.cpu().numpy(), self.acc_m2.cpu().numpy() / np.maximum(n, 1.0)


# ── the single-job driver ───────────────────────────────────────────────────────────────────────────────────────
@dataclass
class _Schedule:
    """Chunk / sampling / exchange lengths of a run after reconciling them (reference pt_mcmc_ce.py:646-674)."""
    chunk: int
    sample_interval: int
    exchange_interval: int
    n_eq_chunks: int
    n_prod_chunks: int

    @classmethod
    def of(cls, cfg, n_total):
        chunk, sample, exch = int(cfg.chunk_size), int(cfg.sample_interval), int(cfg.exchange_interval)
        if n_total > 1 and exch % chunk:
            if exch < chunk:                  # exchanges more often than a chunk: the chunk shrinks to the interval
                chunk = exch
                if chunk % sample:
                    sample = math.gcd(chunk, sample)
                print(f"  exchange interval {exch} is shorter than chunk_size: chunks of {chunk} steps, samples "
                      f"every {sample}", flush=True)
            else:                             # rarer: the interval rounds to whole chunks
                rounded = int(round(exch / chunk)) * chunk
                print(f"  exchange interval {exch} is not a whole number of {chunk}-step chunks: using {rounded}",
                      flush=True)
                exch = rounded
        n_prod = cfg.n_steps - cfg.n_equilib
        for what, steps in (("n_equilib", cfg.n_equilib), ("n_steps - n_equilib", n_prod)):
            if steps % chunk:
                raise ValueError(f"{what} = {steps} is not a whole number of chunks of {chunk} steps")
        if chunk % sample:
            raise ValueError(f"sample_interval {sample} does not divide chunk_size {chunk}")
        return cls(chunk, sample, exch, cfg.n_equilib // chunk, n_prod // chunk)


def anchor_interval_chunks(n_sites, chunk, precision, step_error):
    """Chunks between FP64 re-anchors of the running totals (0: none).  The plain FP64 form needs none.  In the mixed
    form each accepted swap adds the rounding error of its FP32 energy change, of scale `step_error` (eV; the
    model's GPUClusterExpansion.fp32_step_error), and these random-walk in the FP64 total.  Keeping three standard
    deviations of that walk under 1e-9 eV per site after K steps at acceptance <= 1/2 needs
    K <= 2 (1e-9 N / (3 step_error))^2; the re-anchor falls on the first chunk end past K (every chunk for small
    cells or models with large coefficients, never within a run for the 8/6/5 synthetic model at 10^6 sites)."""
    if precision == 0 or step_error <= 0:
        return 0
    K = 2.0 * (1e-9 * float(n_sites) / (3.0 * float(step_error))) ** 2
    return max(1, int(min(K / float(chunk), 2 ** 40)))


def _ladder_of(cfg):
    """temperatures, betas (float32 as the reference stores them), ladder id / position, n per ladder, n ladders."""
    n_ladders = max(int(cfg.n_ladders), 1)
    T_grid = getattr(cfg, "T_grid", None)
    if T_grid is not None:
        if n_ladders != 1:
            raise ValueError("cfg.T_grid is one ladder; n_ladders must be 1")
        T = np.asarray(T_grid, dtype=np.float64).reshape(-1)
        n = len(T)
        return (T.astype(np.float32), (1.0 / (_KB_EV_PER_K * T)).astype(np.float32), np.zeros(n, np.int32),
                np.arange(n, dtype=np.int32), n, 1)
    n_per = int(cfg.n_replicas)
    if n_ladders == 1:
        temps, betas = make_temperature_ladder(cfg.T_min, cfg.T_max, n_per)
        return temps, betas, np.zeros(n_per, np.int32), np.arange(n_per, dtype=np.int32), n_per, 1
    temps, betas, lid, lpos = _build_interleaved_ladders(cfg.T_min, cfg.T_max, n_per, n_ladders)
    return temps, betas, lid, lpos, n_per, n_ladders


def _checkpoint_want(cfg, sch, n_total, n_atoms):
    return dict(n_steps=int(cfg.n_steps), chunk=sch.chunk, n_total=int(n_total), n_atoms=int(n_atoms),
                sample_interval=sch.sample_interval, seed=int(cfg.seed), run_id=str(getattr(cfg, "run_id", "") or ""))


def _save_checkpoint(path, c_done, si, samples, st, best, want, temps):
    """the reference's checkpoint keys (pt_mcmc_ce.py:902-911: chunk, si, energy_samples, key, sigma, rep_of_slot,
    beta_r, cfg_*, spins, E, sigma_min, E_min, swap_cum, flip_cum, ex_cum, step) plus this build's device state:
    per-replica step counters, accepted counts, slot moments and the exchange counter."""
    tmp = path + ".tmp.npz"
    np.savez(tmp, chunk=np.int64(c_done), si=np.int64(si), energy_samples=samples[:si],
             key=np.array([want["seed"]], np.int64), sigma=st.occ.cpu().numpy(),
             rep_of_slot=st.rep_of_slot.cpu().numpy(), beta_r=st.beta_rep.cpu().numpy().astype(np.float32),
             cfg_temps=np.asarray(temps), spins=np.zeros(0), E=st.energy.cpu().numpy(), sigma_min=best[1],
             E_min=np.float64(best[0]), swap_cum=np.int64(st.n_accept.sum().item()), flip_cum=np.int64(0),
             ex_cum=np.int64(st.n_swaps.item()), step=np.int32(int(st.step[0].item())),
             step_rep=st.step.cpu().numpy(), n_accept=st.n_accept.cpu().numpy(),
             acc_n=st.acc_n.cpu().numpy(), acc_mean=st.acc_mean.cpu().numpy(), acc_m2=st.acc_m2.cpu().numpy(),
             e_ref=st.e_ref.cpu().numpy(), ex_count=st.ex_count.cpu().numpy(),
             **{"cfg_" + k: v for k, v in want.items()})
    os.replace(tmp, path)


def _restore_checkpoint(path, st, want, temps):
    """load a checkpoint into the device state; refuse one written by a run with other settings."""
    import torch
    ck = np.load(path, allow_pickle=False)
    have = {k: (str(ck["cfg_" + k]) if isinstance(v, str) else int(ck["cfg_" + k])) for k, v in want.items()
            if "cfg_" + k in ck.files}
    if "run_id" not in have:                 # written without a run fingerprint: nothing to compare it with
        have["run_id"] = want["run_id"]
        print(f"  note: {path} carries no run_id; resuming without checking the model and start structures",
              flush=True)
    diff = sorted(k for k in want if have.get(k) != want[k])
    if not np.allclose(np.asarray(ck["cfg_temps"], np.float64), np.asarray(temps, np.float64), rtol=0, atol=1e-6):
        diff.append("temperatures")
    if diff:
        raise RuntimeError(f"{path} was written by another run (it differs in {', '.join(diff)}); remove the file "
                           f"to start this run from the beginning")

    def put(t, a):
        t.copy_(torch.as_tensor(np.ascontiguousarray(a)).to(device=t.device, dtype=t.dtype))
    put(st.occ, ck["sigma"]); put(st.energy, ck["E"]); put(st.rep_of_slot, ck["rep_of_slot"])
    slot = np.empty(st.R, np.int32); slot[ck["rep_of_slot"]] = np.arange(st.R)
    put(st.slot_of_rep, slot); put(st.beta_rep, ck["beta_r"].astype(np.float64))
    put(st.step, ck["step_rep"]); put(st.n_accept, ck["n_accept"]); st.n_swaps.fill_(int(ck["ex_cum"]))
    put(st.acc_n, ck["acc_n"]); put(st.acc_mean, ck["acc_mean"]); put(st.acc_m2, ck["acc_m2"])
    put(st.e_ref, ck["e_ref"]); put(st.ex_count, ck["ex_count"])
    st.e_min.copy_(st.energy); st.s1.zero_(); st.s2.zero_()
    return int(ck["chunk"]), int(ck["si"]), ck["energy_samples"], (float(ck["E_min"]), ck["sigma_min"])


def run_pt_mcmc(gce, cfg):
    """Run PT-MCMC (reference pt_mcmc_ce.py:601-1028): equilibration and production chunks, an exchange after every
    chunk whose last step is a multiple of the exchange interval, subsampled energies by temperature slot, the
    lowest-energy configuration of production, checkpoint / resume, energies.npz and mcmc_result.npz."""
    os.makedirs(cfg.output_dir, exist_ok=True)
    energy_mode = getattr(cfg, "energy_mode", "full")
    if energy_mode not in ("full", "delta"):
        raise ValueError(f"cfg.energy_mode must be 'full' or 'delta'; got {energy_mode!r}")
    if bool(cfg.spin_orbits_pkl) and bool(cfg.spin_states):
        raise NotImplementedError("spin cluster expansions are not part of the gfx950 swap path")
    precision = {"fp64": 0, "mixed": 1}[getattr(cfg, "precision", "fp64")]
    tile = int(getattr(cfg, "tile", 0))
    temps, betas, ladder_id, ladder_pos, n_per_ladder, n_ladders = _ladder_of(cfg)
    n_total = len(temps)
    sch = _Schedule.of(cfg, n_total)
    sl_species = [sl['chemical_symbols'] for sl in gce.sublattices]
    if cfg.swap_pairs:
        bad = [tuple(p) for p in cfg.swap_pairs if not any(set(p) <= set(sp) for sp in sl_species)]
        if bad:
            raise ValueError(f"swap_pairs {bad} mix species of different sublattices {sl_species}")
    try:
        faulthandler.register(signal.SIGUSR1, all_threads=True)
    except (AttributeError, ValueError, RuntimeError):
        pass
    print(f"PT-MCMC on the GPU: {n_total} temperatures {float(np.min(temps)):.0f}-{float(np.max(temps)):.0f} K"
          + (f" in {n_ladders} interleaved ladders" if n_ladders > 1 else "")
          + (" (T_grid)" if getattr(cfg, "T_grid", None) is not None else "")
          + f", {gce.n_atoms} sites, sublattices {sl_species}")
    print(f"  {cfg.n_steps:,} steps = {sch.n_eq_chunks} equilibration + {sch.n_prod_chunks} production chunks of "
          f"{sch.chunk:,}; exchange every {sch.exchange_interval}, sample every {sch.sample_interval}; "
          f"energy mode {energy_mode}, precision {cfg.precision}", flush=True)

    if cfg.initial_sigma is not None:
        sig0 = np.asarray(cfg.initial_sigma)
        if sig0.shape not in ((gce.n_atoms,), (n_total, gce.n_atoms)):
            raise ValueError(f"initial_sigma has shape {sig0.shape}; expected ({gce.n_atoms},) or "
                             f"({n_total}, {gce.n_atoms})")
    else:
        sig0 = make_initial_state(cfg.seed, gce, cfg.initial_composition, n_total)
    per_chunk = sch.chunk // sch.sample_interval
    st = PTJobState(gce, sig0, betas.astype(np.float64), ladder_id if n_ladders > 1 else None, seed=cfg.seed,
                    sample_interval=sch.sample_interval, samples_per_chunk=per_chunk)
    samples = np.zeros((sch.n_prod_chunks * per_chunk, n_total), dtype=np.float64)
    si = 0
    c_prod0 = 0
    best = (float(st.energy.min()), st.occ[int(st.energy.argmin())].cpu().numpy())
    ck_every = int(getattr(cfg, "checkpoint_every", 0) or 0)
    ck_path = os.path.join(cfg.output_dir, "checkpoint.npz")
    want = _checkpoint_want(cfg, sch, n_total, gce.n_atoms)
    if ck_every and os.path.exists(ck_path):
        c_prod0, si, done, best = _restore_checkpoint(ck_path, st, want, temps)
        samples[:si] = done
        print(f"  continuing from {ck_path}: {c_prod0} of {sch.n_prod_chunks} production chunks done", flush=True)
    ctx = gce.ctx
    n_chunks_all = sch.n_eq_chunks + sch.n_prod_chunks
    todo = (0 if c_prod0 else sch.n_eq_chunks) + sch.n_prod_chunks - c_prod0
    report_every = max(1, n_chunks_all // 50)
    t0 = time.time()
    t_launch = 0.0
    n_run = 0

    anchor_every = anchor_interval_chunks(gce.n_atoms, sch.chunk, precision, gce.fp32_step_error)

    def chunk(accumulate):
        """one chunk of every replica, then the exchange if the global step count reached an exchange point, then
        the FP64 re-anchor of the mixed form when it is due."""
        nonlocal t_launch
        tk = time.time()
        ctx.pt_run_chunk([st.job(accumulate=accumulate)], sch.chunk, precision, tile)
        step_now = int(st.step[0].item())                       # waits for the chunk
        t_launch += time.time() - tk
        if n_total > 1 and step_now % sch.exchange_interval == 0:
            ctx.pt_exchange([st.exchange_job(None, reorder_last=True)])
        if anchor_every and (step_now // sch.chunk) % anchor_every == 0:
            st.anchor()
        return step_now

    def report(phase, k, figure):
        nonlocal n_run
        n_run += 1
        if k % report_every and k != n_chunks_all and not (phase == "equilibration" and k == sch.n_eq_chunks):
            return
        el = time.time() - t0
        print(f"  {phase:<13s} chunk {k:>5d}/{n_chunks_all}  {figure}  {el:7.1f} s elapsed, "
              f"{el / n_run * (todo - n_run):7.1f} s to go", flush=True)

    if c_prod0 == 0:
        for c in range(sch.n_eq_chunks):
            chunk(False)
            report("equilibration", c + 1, f"lowest E now {float(st.energy.min()):12.4f}")
        if sch.n_eq_chunks:
            i0 = int(st.energy.argmin())
            best = (float(st.energy[i0]), st.occ[i0].cpu().numpy())
            print("  equilibration done; the lowest energy is tracked from production on", flush=True)
        st.e_min.copy_(st.energy)
        st.reset_moments()
    for c in range(c_prod0, sch.n_prod_chunks):
        chunk(True)
        samples[si:si + per_chunk] = st.traj.cpu().numpy()
        si += per_chunk
        lows = st.e_min.cpu().numpy()
        if lows.min() < best[0]:                            # chunk granularity, as pt_mcmc_ce.py:811-815
            best = (float(lows.min()), st.occ[int(st.energy.argmin())].cpu().numpy())
        if ck_every and (c + 1) % ck_every == 0 and c + 1 < sch.n_prod_chunks:
            _save_checkpoint(ck_path, c + 1, si, samples, st, best, want, temps)
        report("production", sch.n_eq_chunks + c + 1, f"lowest E {best[0]:12.4f}")
    wall = time.time() - t0
    attempts = cfg.n_steps * n_total
    ex_pairs = (cfg.n_steps // sch.exchange_interval) * (n_per_ladder // 2) * n_ladders
    swap_rate = float(st.n_accept.sum().item()) / max(1, attempts)
    ex_rate = float(st.n_swaps.item()) / max(1, ex_pairs)
    rate = n_total * sch.chunk * todo / max(t_launch, 1e-12)
    E_min = best[0]
    print(f"\nfinished in {wall:.1f} s, {rate:,.0f} swap attempts/s in the chunk launches; acceptance: swaps "
          f"{swap_rate:.4f}, exchanges {ex_rate:.4f}; lowest production energy {E_min:.4f}"
          + (f"; largest re-anchor correction {float(st.drift.item()) / gce.n_atoms:.2e} eV/site" if anchor_every
             else ""), flush=True)
    E_mean, E_var = st.moments()
    temps_np = np.asarray(temps)
    out = lambda name: os.path.join(cfg.output_dir, name)          # noqa: E731
    np.savez(out("energies.npz"), energies=samples[:si], temperatures=temps_np, ladder_id=np.asarray(ladder_id),
             n_atoms=np.int32(gce.n_atoms), swap_accept_rate=swap_rate, exchange_accept_rate=ex_rate,
             flip_accept_rate=0.0, E_min=E_min, wall_time=wall, E_mean=E_mean, E_var=E_var,
             n_accum=st.acc_n.cpu().numpy())
    print(f"  {out('energies.npz')}: {si} samples x {n_total} temperatures", flush=True)
    ip = np.asarray(gce.inv_perm)
    by_slot = st.occ.cpu().numpy()[st.rep_of_slot.cpu().numpy()].astype(np.int32)
    result = MCMCResult(energies=samples[:si], temperatures=temps_np, swap_accept_rate=swap_rate,
                        flip_accept_rate=0.0, exchange_accept_rate=ex_rate,
                        sigma_final=np.ascontiguousarray(by_slot[:, ip]),
                        sigma_min=np.asarray(best[1]).astype(np.int32)[ip], E_min=E_min, spins_final=None,
                        wall_time=wall, E_mean=E_mean, E_var=E_var, attempts_per_s=rate)
    np.savez(out("mcmc_result.npz"), energies=result.energies, temperatures=result.temperatures,
             ladder_id=np.asarray(ladder_id), ladder_pos=np.asarray(ladder_pos), n_ladders=np.int32(n_ladders),
             sigma_final=result.sigma_final, sigma_min=result.sigma_min, E_min=result.E_min, spins_final=np.array([]),
             swap_accept_rate=result.swap_accept_rate, flip_accept_rate=0.0,
             exchange_accept_rate=result.exchange_accept_rate, wall_time=result.wall_time)
    print(f"  {out('mcmc_result.npz')}: final and lowest-energy configurations", flush=True)
    if ck_every and os.path.exists(ck_path):
        os.remove(ck_path)
    return result
