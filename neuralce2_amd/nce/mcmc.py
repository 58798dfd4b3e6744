"""Parallel-tempering MC for the NeuralCE model (reference: neuralce/nce/mcmc.py).

``_run_replicas`` follows mcmc.py:288-422: chunks of swap steps for every replica (one ``nce_nn_run_chunk`` launch),
a full recompute of caches and energies every ``anchor_every`` chunks with the drift logged, and at the end of every
chunk ``max(chunk // exchange_every, 1)`` rounds of neighbour configuration swaps accepted with
d = (b_i - b_{i+1})(E_i - E_{i+1}), d >= 0 or u < exp(d) in FP64 on the host (mcmc.py:370-384; configurations and
caches move, O(R N)).  Outputs: energies.npz (E_chunk, acc_chunk, drift, drift_eV, traj, betas, temps),
final_sigma.npy in input atom order.  Random numbers: Philox per (seed, replica, step) on the device, numpy
Generator(seed) for the exchange.
"""
from __future__ import annotations

import os
import time

import numpy as np

KB = 8.617333262e-5


def _log(*a):
    print(*a, flush=True)


class NNJobState:
    """Device state of R replicas of one NeuralCE model."""

    def __init__(self, nn, sig0, betas, seed=0, precision=1, sample_interval=1, samples_per_chunk=0):
        import torch
        self.nn = nn
        self.R = R = len(betas)
        sig = np.asarray(sig0)
        if sig.ndim == 1:
            sig = np.broadcast_to(sig, (R, nn.N))
        self.occ = torch.tensor(np.ascontiguousarray(sig, np.int8), device=nn.device)
        self.precision = precision
        self.cache, self.scratch = nn.alloc(R, precision)
        self.energy = torch.empty(R, dtype=torch.float64, device=nn.device)
        self.anchor()
        self.beta = torch.tensor(np.asarray(betas, np.float64), device=nn.device)
        self.step = torch.zeros(R, dtype=torch.int64, device=nn.device)
        self.n_accept = torch.zeros(R, dtype=torch.int64, device=nn.device)
        self.sample_interval = int(sample_interval)
        self.traj = (torch.zeros((samples_per_chunk, R), dtype=torch.float64, device=nn.device)
                     if samples_per_chunk else None)
        self.seed = int(seed)

    def anchor(self):
        """full recompute of caches and energies; returns the largest change of the carried energies."""
        import torch
        E_old = self.energy.clone()
        E = torch.empty_like(self.energy)
        self.nn.ctx.nn_full_cache(self.nn.model, self.occ, self.cache, E, self.precision)
        self.energy.copy_(E)
        return E_old, E

    def job(self, inj=None):
        from .._lib import NNJob, ptr
        j = NNJob()
        j.model = self.nn.model; j.n_replicas = self.R; j.n_sites = self.nn.N
        j.occ = ptr(self.occ); j.cache = ptr(self.cache); j.scratch = ptr(self.scratch)
        j.energy = ptr(self.energy); j.beta = ptr(self.beta); j.step = ptr(self.step); j.seed = self.seed
        j.n_accept = ptr(self.n_accept); j.traj = ptr(self.traj) if self.traj is not None else None
        j.sample_interval = self.sample_interval
        if inj is not None:
            self._inj = inj
            j.inj_ia = ptr(inj[0]); j.inj_ib = ptr(inj[1]); j.inj_u = ptr(inj[2])
        return j

    def run(self, n_steps, inj=None):
        self.nn.ctx.nn_run_chunk([self.job(inj)], n_steps, self.precision)

    def permute(self, order):
        import torch
        idx = torch.as_tensor(np.asarray(order), device=self.occ.device)
        for name in ("occ", "cache", "energy"):
            t = getattr(self, name)
            t.copy_(t[idx])


def _run_replicas(mc, betas, n_steps, seed, chunk, log_every, sample_interval, out_dir, tag, exchange_every=0,
                  anchor_every=None, init_npz=None, checkpoint_every=0):
    nn = mc["nn"]
    R = len(betas)
    betas = np.asarray(betas, np.float32).astype(np.float64)          # the reference keeps beta in float32
    sig0 = mc["sig0"] if init_npz is None else init_npz
    n_chunks = max(n_steps // chunk, 1)
    if chunk % sample_interval != 0:
        raise ValueError(f"chunk ({chunk}) must be a multiple of sample_interval ({sample_interval})")
    st = NNJobState(nn, sig0, betas, seed=seed, precision=mc["precision"], sample_interval=sample_interval,
                    samples_per_chunk=chunk // sample_interval)
    if anchor_every is None:
        anchor_every = max(1, int(np.ceil(nn.N / chunk)))
    rng = np.random.default_rng(seed)
    E_chunk, acc_chunk, drift_log, traj = [], [], [], []
    n_ex_try = n_ex_acc = 0
    drift = 0.0
    _log(f"[{tag}] {R} replicas, {n_chunks} chunks of {chunk} steps; full re-anchor every {anchor_every} chunk(s)")
    t0 = time.time()
    acc_prev = np.zeros(R, np.int64)
    for c in range(n_chunks):
        st.run(chunk)
        if (c + 1) % anchor_every == 0 or c == n_chunks - 1:
            E_old, E_new = st.anchor()
            drift = float((E_new - E_old).abs().max().item())
        E = st.energy.cpu().numpy()
        E_eV = E * nn.ys + nn.ym
        acc = st.n_accept.cpu().numpy()
        E_chunk.append(E_eV); acc_chunk.append((acc - acc_prev) / chunk); acc_prev = acc
        drift_log.append(drift)
        traj.append(st.traj.cpu().numpy().T * nn.ys + nn.ym)
        if exchange_every > 0 and R > 1:
            for _ in range(max(chunk // exchange_every, 1)):
                start = int(rng.integers(0, 2))
                u = rng.random(R)
                E_np = st.energy.cpu().numpy() * nn.ys * nn.N
                order = list(range(R))
                for i in range(start, R - 1, 2):
                    n_ex_try += 1
                    d = (betas[i] - betas[i + 1]) * (E_np[i] - E_np[i + 1])
                    if d >= 0 or u[i] < np.exp(d):
                        order[i], order[i + 1] = order[i + 1], order[i]; n_ex_acc += 1
                if order != list(range(R)):
                    st.permute(order)
        if log_every and ((c + 1) % max(log_every // chunk, 1) == 0 or c == n_chunks - 1):
            _log(f"[{tag}] chunk {c+1}/{n_chunks}  E(eV/atom) mean {E_eV.mean():.6f} [{E_eV.min():.6f},"
                 f"{E_eV.max():.6f}]  acc {np.mean(acc_chunk[-1]):.3f}  drift {drift:.1e}  "
                 f"{R * chunk * (c + 1) / (time.time() - t0):,.0f} steps/s" + (f"  exch {n_ex_acc}/{n_ex_try}"
                                                                              if exchange_every else ""))
    wall = time.time() - t0
    os.makedirs(out_dir, exist_ok=True)
    E_chunk = np.array(E_chunk); acc_chunk = np.array(acc_chunk); traj = np.concatenate(traj, axis=1)
    np.savez(f"{out_dir}/energies.npz", E_chunk=E_chunk, acc_chunk=acc_chunk, drift=np.array(drift_log),
             drift_eV=np.array(drift_log) * nn.ys, traj=traj, betas=betas, temps=1.0 / (KB * betas))
    sig_np = st.occ.cpu().numpy()[:, mc["inv_perm"]]
    np.save(f"{out_dir}/final_sigma.npy", sig_np)
    return dict(steps_per_s=float(R * chunk * n_chunks / wall), n_replicas=R, n_steps=int(n_steps), chunk=int(chunk),
                E_final=[float(x) for x in E_chunk[-1]], acc_mean=float(acc_chunk.mean()),
                drift_max=float(np.max(drift_log)), exchange_acc=(n_ex_acc / n_ex_try if n_ex_try else None),
                model_depth=int(mc["depth"]), n_atoms=int(nn.N))


def _prepare(model_path, init_cif, supercell=(1, 1, 1), seed_idx=0, dtype="float32", seed=0, **_):
    from ..lattice import read_cif
    from .mc_model import load_model, setup
    mdl = load_model(model_path)
    atoms = read_cif(init_cif)
    sc = tuple(int(x) for x in supercell)
    if sc != (1, 1, 1):
        atoms = atoms.repeat(sc)
        if os.environ.get("CE_CNN_SUPERCELL_PERIODIC", "0") != "1":
            Z = atoms.numbers.copy()
            atoms.numbers = np.random.default_rng(seed).permutation(Z)
    t0 = time.time()
    mc = setup(mdl, atoms, seed_idx=seed_idx, dtype=dtype)
    _log(f"[ce_cnn] model {os.path.basename(os.path.dirname(model_path))}: depth {mc['depth']}, {mc['n_atoms']} atoms, "
         f"setup {time.time()-t0:.1f}s, E0 = {mc['energy'](mc['sig0']):.6f} eV/atom")
    mc["mdl"] = mdl; mc["atoms"] = atoms
    return mc


def run_pt_mc(model_path, init_cif, n_steps=10000, T_min=100.0, T_max=2000.0, n_replicas=32, T_grid=None,
              exchange_every=100, seed=0, log_every=0, out_dir=".", chunk=1000, sample_interval=100,
              supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None, init_npz=None,
              checkpoint_every=0, **_):
    """reference: neuralce/nce/mcmc.py:466-475 (geomspace ladder, beta = 1 / (kB T), kB = 8.617333262e-5)."""
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    temps = np.asarray(T_grid, float) if T_grid is not None else np.geomspace(T_min, T_max, n_replicas)
    betas = [1.0 / (KB * t) for t in temps]
    sig0 = None
    if init_npz:
        from ..mc.init_structures import load_init_npz
        from ..lattice import ATOMIC_NUMBERS
        nums, _ = load_init_npz(init_npz, mc["mdl"]["primitive"], mc["atoms"], temps)
        lut = np.full(256, -1, np.int64)
        for k, sp in enumerate(mc["nn"].geo["species"]):
            lut[ATOMIC_NUMBERS[sp]] = k
        sig = lut[nums.astype(np.int64)]
        if (sig < 0).any():
            raise ValueError("init_npz holds a species outside the model's")
        sig0 = sig[:, mc["perm"]].astype(np.int8)
    res = _run_replicas(mc, betas, n_steps, seed, min(chunk, n_steps), log_every, sample_interval, out_dir,
                        "ce_cnn PT", exchange_every=exchange_every, anchor_every=anchor_every, init_npz=sig0)
    res.update(temps=[float(t) for t in temps], exchange_every=int(exchange_every))
    return res


def run_mc(model_path, init_cif, n_steps=10000, temperature=300.0, n_replicas=1, chunk=1000, sample_interval=100,
           seed=0, log_every=0, out_dir=".", supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None,
           **_):
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    res = _run_replicas(mc, [1.0 / (KB * temperature)] * n_replicas, n_steps, seed, min(chunk, n_steps), log_every,
                        sample_interval, out_dir, f"ce_cnn T={temperature:.0f}", anchor_every=anchor_every)
    res.update(temperature=float(temperature))
    return res


def run_multiT_mc(model_path, init_cif, temps, n_steps=10000, chunk=1000, sample_interval=100, seed=0, log_every=0,
                  out_dir=".", supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None, **_):
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    res = _run_replicas(mc, [1.0 / (KB * t) for t in temps], n_steps, seed, min(chunk, n_steps), log_every,
                        sample_interval, out_dir, "ce_cnn multiT", anchor_every=anchor_every)
    res.update(temps=[float(t) for t in temps])
    return res
