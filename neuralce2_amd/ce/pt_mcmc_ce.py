"""Parallel-tempering swap Monte Carlo for the cluster expansion on the gfx950 chunk kernel.

Drop-in for neuralce/ce/pt_mcmc_ce.py: ``MCMCConfig`` (:37-104), ``make_temperature_ladder`` (:223-229),
``_build_interleaved_ladders`` (:232-260), ``MCMCResult`` (:587-598) and ``run_pt_mcmc`` (:601-1028) with the same
chunk / exchange arithmetic, the same checkpoint keys and the same output files.  The step loop is one
``nce_pt_run_chunk`` launch per chunk (every replica of the job, every step of the chunk); the temperature-swap
exchange is ``nce_pt_exchange`` at chunk boundaries.

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

    The same object serves the single-job driver below and the multi-job scheduler (neuralce2_amd.serve)."""

    def __init__(self, gce, sigma0, betas_slot, ladder_id=None, seed=0, sample_interval=1, samples_per_chunk=0,
                 job_id=0):
        import torch
        from .._lib import PTJob, ExchangeJob
        dev = gce.device
        self.gce = gce
        self.R = R = len(betas_slot)
        self.N = gce.n_atoms
        sig = np.asarray(sigma0)
        if sig.ndim == 1:
            sig = np.broadcast_to(sig, (R, self.N))
        if sig.shape != (R, self.N):
            raise ValueError(f"initial sigma shape {sig.shape} != ({R}, {self.N})")
        if sig.min() < 0 or sig.max() > 1:
            raise ValueError("species index out of range for a binary cluster expansion")
        self.occ = torch.tensor(np.ascontiguousarray(sig, np.int8), device=dev)
        self.energy = torch.empty(R, dtype=torch.float64, device=dev)
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
        self.n_swaps = torch.zeros(1, dtype=torch.int64, device=dev)
        self.sample_interval = int(sample_interval)
        self.traj = torch.zeros((max(samples_per_chunk, 1), R), dtype=torch.float64, device=dev) if samples_per_chunk else None
        self.seed = int(seed) & ((1 << 64) - 1)
        self.job_id = job_id
        # FP64 per-slot accumulators on the host (Chan's parallel merge of per-chunk moments)
        self.acc_n = np.zeros(R, np.float64); self.acc_mean = np.zeros(R); self.acc_m2 = np.zeros(R)
        self._PTJob, self._ExchangeJob = PTJob, ExchangeJob

    def job(self, accumulate=False, inj=None):
        from .._lib import ptr
        j = self._PTJob()
        j.model = self.gce.model; j.n_replicas = self.R; j.n_sites = self.N
        j.occ = ptr(self.occ); j.energy = ptr(self.energy); j.beta = ptr(self.beta_rep)
        j.slot_of_rep = ptr(self.slot_of_rep); j.step = ptr(self.step); j.seed = self.seed
        j.n_accept = ptr(self.n_accept); j.e_min = ptr(self.e_min)
        j.traj = ptr(self.traj) if self.traj is not None else None
        j.sample_interval = self.sample_interval
        j.accumulate = int(bool(accumulate)); j.e_ref = ptr(self.e_ref); j.s1 = ptr(self.s1); j.s2 = ptr(self.s2)
        if inj is not None:
            self._inj = inj
            j.inj_ia = ptr(inj[0]); j.inj_ib = ptr(inj[1]); j.inj_u = ptr(inj[2])
        return j

    def exchange_job(self, ex_index, reorder_last=True, inj_parity=-1, inj_u=None):
        from .._lib import ptr
        e = self._ExchangeJob()
        e.n_replicas = self.R; e.inj_parity = int(inj_parity)
        e.energy = ptr(self.energy); e.rep_of_slot = ptr(self.rep_of_slot); e.slot_of_rep = ptr(self.slot_of_rep)
        e.beta_slot = ptr(self.beta_slot); e.ladder_id = ptr(self.ladder_id); e.beta_rep = ptr(self.beta_rep)
        e.seed = self.seed ^ 0x9E3779B97F4A7C15; e.ex_index = int(ex_index); e.n_swaps = ptr(self.n_swaps)
        if inj_u is not None:
            self._inj_u = inj_u
        e.inj_u = ptr(inj_u)
        e.traj_last = ptr(self.traj[-1]) if (reorder_last and self.traj is not None) else None
        return e

    def merge_accumulators(self, n_steps, slot_of_rep_np):
        """fold the chunk's per-replica (s1, s2) around e_ref into the per-slot FP64 moments; reset for the next chunk."""
        s1 = self.s1.cpu().numpy(); s2 = self.s2.cpu().numpy(); ref = self.e_ref.cpu().numpy()
        n = float(n_steps)
        mean_c = ref + s1 / n
        m2_c = np.maximum(s2 - s1 * s1 / n, 0.0)
        for r in range(self.R):
            t = int(slot_of_rep_np[r])
            na = self.acc_n[t]; nb = n; tot = na + nb
            delta = mean_c[r] - self.acc_mean[t]
            self.acc_mean[t] += delta * nb / tot
            self.acc_m2[t] += m2_c[r] + delta * delta * na * nb / tot
            self.acc_n[t] = tot
        self.s1.zero_(); self.s2.zero_()
        self.e_ref.copy_(self.energy)

    def moments(self):
        n = np.maximum(self.acc_n, 1.0)
        return self.acc_mean.copy(), self.acc_m2 / n


def run_pt_mcmc(gce, cfg):
    """Run PT-MCMC (reference pt_mcmc_ce.py:601-1028): equilibration and production chunks, exchange at chunk
    boundaries whose step is a multiple of the exchange interval, subsampled energies by temperature slot,
    checkpoint / resume, energies.npz and mcmc_result.npz."""
    import torch
    os.makedirs(cfg.output_dir, exist_ok=True)
    has_spin = bool(cfg.spin_orbits_pkl) and bool(cfg.spin_states)
    energy_mode = getattr(cfg, "energy_mode", "full")
    if energy_mode not in ("full", "delta"):
        raise ValueError(f"cfg.energy_mode must be 'full' or 'delta'; got {energy_mode!r}")
    if has_spin:
        raise NotImplementedError("spin cluster expansions are not part of the gfx950 swap path")
    precision = {"fp64": 0, "mixed": 1}[getattr(cfg, "precision", "fp64")]
    n_ladders = max(int(cfg.n_ladders), 1)
    n_per_ladder = int(cfg.n_replicas)
    n_total = n_per_ladder * n_ladders
    T_grid = getattr(cfg, "T_grid", None)
    if T_grid is not None:
        if n_ladders != 1:
            raise ValueError("cfg.T_grid is one ladder; n_ladders must be 1")
        T_grid = np.asarray(T_grid, dtype=np.float64).reshape(-1)
        n_per_ladder = n_total = int(len(T_grid))
        temps = T_grid.astype(np.float32)
        betas = (1.0 / (_KB_EV_PER_K * T_grid)).astype(np.float32)
        ladder_id = np.zeros(n_total, np.int32); ladder_pos = np.arange(n_total, dtype=np.int32)
    elif n_ladders == 1:
        temps, betas = make_temperature_ladder(cfg.T_min, cfg.T_max, n_per_ladder)
        ladder_id = np.zeros(n_total, np.int32); ladder_pos = np.arange(n_total, dtype=np.int32)
    else:
        temps, betas, ladder_id, ladder_pos = _build_interleaved_ladders(cfg.T_min, cfg.T_max, n_per_ladder, n_ladders)

    chunk_size = int(cfg.chunk_size)
    sample_interval = int(cfg.sample_interval)
    exch_interval = int(cfg.exchange_interval)
    if n_total > 1 and exch_interval % chunk_size != 0:
        if exch_interval < chunk_size:
            new_chunk = exch_interval
            if new_chunk % sample_interval != 0:
                sample_interval = math.gcd(new_chunk, sample_interval)
            print(f"  chunk_size {chunk_size} -> {new_chunk} (= exchange interval) so exchanges fall on chunk "
                  f"boundaries; sample_interval {sample_interval}", flush=True)
            chunk_size = new_chunk
        else:
            new_ex = int(round(exch_interval / chunk_size)) * chunk_size
            print(f"  exchange_interval {exch_interval} -> {new_ex} (whole chunks of {chunk_size})", flush=True)
            exch_interval = new_ex
    n_production = cfg.n_steps - cfg.n_equilib
    if cfg.n_equilib % chunk_size != 0:
        raise ValueError(f"n_equilib ({cfg.n_equilib}) must be a multiple of chunk_size ({chunk_size})")
    if n_production % chunk_size != 0:
        raise ValueError(f"(n_steps − n_equilib) = {n_production} must be a multiple of chunk_size ({chunk_size})")
    if chunk_size % sample_interval != 0:
        raise ValueError(f"chunk_size ({chunk_size}) must be a multiple of sample_interval ({sample_interval})")

    print(f"PT-MCMC: {n_total} replicas, T=[{float(np.min(temps)):.0f}, {float(np.max(temps)):.0f}] K"
          + (" (explicit list)" if T_grid is not None else "") + (f", {n_ladders} ladders" if n_ladders > 1 else ""))
    print(f"  n_steps={cfg.n_steps:,}, n_equilib={cfg.n_equilib:,}, chunk_size={chunk_size:,}, exchange_interval="
          f"{exch_interval}, sample_interval={sample_interval}, energy_mode={energy_mode}, precision="
          f"{cfg.precision}", flush=True)
    sl_species = [sl['chemical_symbols'] for sl in gce.sublattices]
    print(f"  sublattices: {sl_species}")
    if cfg.swap_pairs:
        sl_sets = [set(sp) for sp in sl_species]
        bad = [(a, b) for (a, b) in cfg.swap_pairs if not any({a, b} <= s for s in sl_sets)]
        if bad:
            raise ValueError(f"swap_pairs {bad} do not lie within any single sublattice (sublattices: {sl_species}).")
    try:
        faulthandler.register(signal.SIGUSR1, all_threads=True)
    except (AttributeError, ValueError, RuntimeError):
        pass

    if cfg.initial_sigma is not None:
        init_sigma = np.asarray(cfg.initial_sigma)
        if init_sigma.shape not in ((gce.n_atoms,), (n_total, gce.n_atoms)):
            raise ValueError(f"initial_sigma shape {init_sigma.shape} must be (n_atoms,)=({gce.n_atoms},) "
                             f"or (n_replicas, n_atoms)=({n_total}, {gce.n_atoms})")
    else:
        init_sigma = make_initial_state(cfg.seed, gce, cfg.initial_composition, n_total)
    samples_per_chunk = chunk_size // sample_interval
    st = PTJobState(gce, init_sigma, betas.astype(np.float64), ladder_id if n_ladders > 1 else None,
                    seed=cfg.seed, sample_interval=sample_interval, samples_per_chunk=samples_per_chunk)
    n_chunks_eq = cfg.n_equilib // chunk_size
    n_chunks_prod = n_production // chunk_size
    n_chunks_tot = n_chunks_eq + n_chunks_prod
    print(f"\nRunning {cfg.n_steps:,} steps  ({n_chunks_eq} equilib + {n_chunks_prod} production chunks)")
    t0 = time.time()
    progress_every = max(1, n_chunks_tot // 50)
    energy_samples = np.zeros((n_chunks_prod * samples_per_chunk, n_total), dtype=np.float64)
    si = 0; c0 = 0; n_ex_calls = 0
    E_min = float(st.energy.min()); sigma_min = st.occ[int(st.energy.argmin())].cpu().numpy()
    ck_every = int(getattr(cfg, "checkpoint_every", 0) or 0)
    ck_path = os.path.join(cfg.output_dir, "checkpoint.npz")
    if ck_every and os.path.exists(ck_path):
        ck = np.load(ck_path, allow_pickle=False)
        want = dict(n_steps=cfg.n_steps, chunk=chunk_size, n_total=n_total, n_atoms=gce.n_atoms,
                    sample_interval=sample_interval, seed=cfg.seed)
        have = {k: int(ck["cfg_" + k]) for k in want}
        if "cfg_run_id" in ck.files:
            want["run_id"] = str(getattr(cfg, "run_id", "") or ""); have["run_id"] = str(ck["cfg_run_id"])
        if not np.allclose(np.asarray(ck["cfg_temps"]), np.asarray(temps), rtol=0, atol=1e-6):
            have["temps"] = "differ"; want["temps"] = "same"
        if have != want:
            raise RuntimeError(f"checkpoint {ck_path} belongs to a different run (checkpoint {have} vs this run "
                               f"{want}); delete it to start fresh")
        c0 = int(ck["chunk"]); si = int(ck["si"]); energy_samples[:si] = ck["energy_samples"]
        dev = gce.device
        st.occ.copy_(torch.from_numpy(np.ascontiguousarray(ck["sigma"], np.int8)).to(dev))
        st.energy.copy_(torch.from_numpy(ck["E"]).to(dev))
        st.rep_of_slot.copy_(torch.from_numpy(ck["rep_of_slot"].astype(np.int32)).to(dev))
        sor = np.empty(n_total, np.int32); sor[ck["rep_of_slot"]] = np.arange(n_total)
        st.slot_of_rep.copy_(torch.from_numpy(sor).to(dev))
        st.beta_rep.copy_(torch.from_numpy(ck["beta_r"].astype(np.float64)).to(dev))
        st.step.copy_(torch.from_numpy(ck["step_rep"].astype(np.int64)).to(dev))
        st.n_accept.copy_(torch.from_numpy(ck["n_accept"].astype(np.int64)).to(dev))
        st.n_swaps.fill_(int(ck["ex_cum"]))
        st.e_ref.copy_(st.energy); st.e_min.copy_(st.energy)
        sigma_min = ck["sigma_min"]; E_min = float(ck["E_min"])
        st.acc_n[:] = ck["acc_n"]; st.acc_mean[:] = ck["acc_mean"]; st.acc_m2[:] = ck["acc_m2"]
        n_ex_calls = int(ck["n_ex_calls"])
        print(f"  resumed from {ck_path}: production chunk {c0}/{n_chunks_prod}, {si} samples", flush=True)
    chunks_this_run = (n_chunks_eq if c0 == 0 else 0) + (n_chunks_prod - c0)
    ctx = gce.ctx
    t_kernel = 0.0

    def one_chunk(accumulate):
        nonlocal n_ex_calls, t_kernel
        slot_before = st.slot_of_rep.cpu().numpy() if accumulate else None
        tk = time.time()
        ctx.pt_run_chunk([st.job(accumulate=accumulate)], chunk_size, precision, int(getattr(cfg, "tile", 0)))
        step_now = int(st.step[0].item())          # synchronises
        t_kernel += time.time() - tk
        if accumulate:
            st.merge_accumulators(chunk_size, slot_before)
        if n_total > 1 and step_now % exch_interval == 0:
            ctx.pt_exchange([st.exchange_job(step_now // exch_interval, reorder_last=True)])
            n_ex_calls += 1
        return step_now

    if c0 == 0:
        for c in range(n_chunks_eq):
            one_chunk(False)
            if (c + 1) % progress_every == 0 or c + 1 == n_chunks_eq:
                el = time.time() - t0; eta = el / (c + 1) * (chunks_this_run - (c + 1))
                print(f"  [equil  {c+1:>5d}/{n_chunks_tot}]  E_min(curr)={float(st.energy.min()):>10.2f}  "
                      f"elapsed={el:>7.1f}s  ETA={eta:>7.1f}s", flush=True)
        idx0 = int(st.energy.argmin()); E_min = float(st.energy[idx0]); sigma_min = st.occ[idx0].cpu().numpy()
        st.e_min.copy_(st.energy); st.e_ref.copy_(st.energy)
        if n_chunks_eq > 0:
            print("  → Equilibration complete; reset E_min for production.")
    for c in range(c0, n_chunks_prod):
        one_chunk(True)
        E_sub = st.traj.cpu().numpy()
        energy_samples[si:si + samples_per_chunk] = E_sub; si += samples_per_chunk
        emin_r = st.e_min.cpu().numpy()
        if emin_r.min() < E_min:                          # chunk-granular sigma_min, as pt_mcmc_ce.py:811-815
            E_min = float(emin_r.min())
            sigma_min = st.occ[int(st.energy.argmin())].cpu().numpy()
        if ck_every and (c + 1) % ck_every == 0 and c + 1 < n_chunks_prod:
            np.savez(ck_path + ".tmp.npz", chunk=c + 1, si=si, energy_samples=energy_samples[:si],
                     key=np.array([cfg.seed], np.int64), sigma=st.occ.cpu().numpy(),
                     rep_of_slot=st.rep_of_slot.cpu().numpy(), beta_r=st.beta_rep.cpu().numpy().astype(np.float32),
                     cfg_n_steps=cfg.n_steps, cfg_chunk=chunk_size, cfg_n_total=n_total, cfg_n_atoms=gce.n_atoms,
                     cfg_sample_interval=sample_interval, cfg_seed=cfg.seed, cfg_temps=np.asarray(temps),
                     cfg_run_id=str(getattr(cfg, "run_id", "") or ""), spins=np.zeros(0),
                     E=st.energy.cpu().numpy(), sigma_min=sigma_min, E_min=np.float64(E_min),
                     swap_cum=np.int64(st.n_accept.sum().item()), flip_cum=np.int64(0),
                     ex_cum=np.int64(st.n_swaps.item()), step=np.int32(int(st.step[0].item())),
                     step_rep=st.step.cpu().numpy(), n_accept=st.n_accept.cpu().numpy(), acc_n=st.acc_n,
                     acc_mean=st.acc_mean, acc_m2=st.acc_m2, n_ex_calls=np.int64(n_ex_calls))
            os.replace(ck_path + ".tmp.npz", ck_path)
        global_chunk = n_chunks_eq + c + 1
        if global_chunk % progress_every == 0 or c + 1 == n_chunks_prod:
            el = time.time() - t0
            done = (n_chunks_eq if c0 == 0 else 0) + (c + 1 - c0)
            eta = el / done * (chunks_this_run - done) if done else 0.0
            print(f"  [prod   {global_chunk:>5d}/{n_chunks_tot}]  E_min(run)={E_min:>10.4f}  elapsed={el:>7.1f}s  "
                  f"ETA={eta:>7.1f}s", flush=True)
    wall_time = time.time() - t0
    n_swap_attempts = cfg.n_steps * n_total
    pairs_per_step = (n_per_ladder // 2) * n_ladders
    n_ex_attempts = (cfg.n_steps // exch_interval) * pairs_per_step
    swap_accept_rate = float(st.n_accept.sum().item()) / max(1, n_swap_attempts)
    exchange_accept_rate = float(st.n_swaps.item()) / max(1, n_ex_attempts)
    attempts_per_s = n_total * chunk_size * chunks_this_run / max(t_kernel, 1e-12)
    print(f"\nDone in {wall_time:.1f}s  ({attempts_per_s:,.0f} swap attempts/s in the chunk kernel)")
    print(f"  swap acceptance     : {swap_accept_rate:.4f}")
    print(f"  exchange acceptance : {exchange_accept_rate:.4f}")
    print(f"  E_min (production)  : {E_min:.4f}", flush=True)
    temps_np = np.asarray(temps)
    E_mean, E_var = st.moments()
    np.savez(os.path.join(cfg.output_dir, 'energies.npz'), energies=energy_samples[:si], temperatures=temps_np,
             ladder_id=np.asarray(ladder_id), n_atoms=np.int32(gce.n_atoms), swap_accept_rate=swap_accept_rate,
             exchange_accept_rate=exchange_accept_rate, flip_accept_rate=0.0, E_min=E_min, wall_time=wall_time,
             E_mean=E_mean, E_var=E_var, n_accum=st.acc_n)
    print(f"  wrote {os.path.join(cfg.output_dir, 'energies.npz')} ({si} samples x {n_total} T)", flush=True)
    ros = st.rep_of_slot.cpu().numpy()
    sigma_final = st.occ.cpu().numpy()[ros].astype(np.int32)
    ip = np.asarray(gce.inv_perm)
    sigma_final = np.ascontiguousarray(sigma_final)[:, ip]
    sigma_min_o = np.asarray(sigma_min).astype(np.int32)[ip]
    result = MCMCResult(energies=energy_samples[:si], temperatures=temps_np, swap_accept_rate=swap_accept_rate,
                        flip_accept_rate=0.0, exchange_accept_rate=exchange_accept_rate, sigma_final=sigma_final,
                        sigma_min=sigma_min_o, E_min=E_min, spins_final=None, wall_time=wall_time,
                        E_mean=E_mean, E_var=E_var, attempts_per_s=attempts_per_s)
    np.savez(os.path.join(cfg.output_dir, 'mcmc_result.npz'), energies=result.energies,
             temperatures=result.temperatures, ladder_id=np.asarray(ladder_id), ladder_pos=np.asarray(ladder_pos),
             n_ladders=np.int32(n_ladders), sigma_final=result.sigma_final, sigma_min=result.sigma_min,
             E_min=result.E_min, spins_final=np.array([]), swap_accept_rate=result.swap_accept_rate,
             flip_accept_rate=0.0, exchange_accept_rate=result.exchange_accept_rate, wall_time=result.wall_time)
    print(f"  wrote {os.path.join(cfg.output_dir, 'mcmc_result.npz')}", flush=True)
    if ck_every and os.path.exists(ck_path):
        os.remove(ck_path)
    return result
