"""Parallel-tempering MC for the NeuralCE model (reference: neuralce/nce/mcmc.py).

``_run_replicas`` keeps the contract of mcmc.py:288-422: chunks of swap steps for every replica (one
``nce_nn_run_chunk`` launch), a full recompute of caches and energies every ``anchor_every`` chunks (FP32 / BF16
drift control; the drift is logged), neighbour exchanges of configurations between chunks accepted with
d = (b_i - b_{i+1})(E_i - E_{i+1}), d >= 0 or u < exp(d) in FP64 (mcmc.py:370-384), checkpoint / resume through
``checkpoint.npz`` (mcmc.py:322-399) and the output files energies.npz (E_chunk, acc_chunk, drift, drift_eV, traj,
betas, temps), final_sigma.npy (input atom order) and, for cells up to CE_CNN_CIF_MAX sites, one cif per replica.

Random numbers are counter-based: the device draws proposals from Philox keyed by (seed, replica, step), the host
draws the exchange parity and uniforms from Philox keyed by (seed, exchange count).  A run's state is therefore its
occupations, step counters and exchange count, which is what a checkpoint holds; a resumed run continues the same
streams.
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
        sig = None if torch.is_tensor(sig0) else np.asarray(sig0)   # a device tensor: one configuration
        self.precision = precision
        self.cache, self.scratch = nn.alloc(R, precision)
        self.energy = torch.empty(R, dtype=torch.float64, device=nn.device)
        if sig is None or sig.ndim == 1:
            # every replica starts from the same configuration: one row uploaded and replicated on the device, its
            # caches and energy computed once and copied to the other replicas
            row = (sig0.to(device=nn.device, dtype=torch.int8) if sig is None
                   else torch.tensor(np.ascontiguousarray(sig, np.int8), device=nn.device)).unsqueeze(0)
            self.occ = row.expand(R, nn.N).contiguous()
            nn.ctx.nn_full_cache(nn.model, row, self.cache[:1], self.energy[:1], precision)
            self.cache[1:] = self.cache[:1]
            self.energy[1:] = self.energy[:1]
        else:
            self.occ = torch.tensor(np.ascontiguousarray(sig, np.int8), device=nn.device)
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


_CK_NAME = "checkpoint.npz"


def _fingerprint(nn, sig):
    """16 hex digits identifying the network parameters, the standardisation and the starting occupations."""
    import hashlib
    h = hashlib.sha256()
    P = nn.P

    def feed(x):
        h.update(np.ascontiguousarray(np.asarray(x, np.float64)).tobytes())
    feed(P["lin"]); feed(nn.mu); feed(nn.sd); feed(P["head"]["b"])
    if "w" in P["head"]:
        feed(P["head"]["w"])
    for L in P["layers"]:
        for k in sorted(L):
            h.update(k.encode()); feed(L[k])
    h.update(np.ascontiguousarray(np.asarray(sig), np.int8).tobytes())
    return h.hexdigest()[:16]


def exchange_draws(seed, ex_index, R):
    """parity (0/1) and R uniforms of exchange number `ex_index`: Philox keyed by the run seed, counter
    (ex_index, i, 'EX')."""
    from ..philox import philox4x32_10, uniform53
    key = (int(seed) & 0xFFFFFFFF, ((int(seed) >> 32) & 0xFFFFFFFF) ^ 0x45584348)
    lo, hi = int(ex_index) & 0xFFFFFFFF, (int(ex_index) >> 32) & 0xFFFFFFFF
    ctr = np.zeros((R + 1, 4), np.uint64)
    ctr[:, 0] = lo; ctr[:, 1] = hi; ctr[:, 2] = np.arange(R + 1); ctr[:, 3] = 0x4558
    w = philox4x32_10(ctr, key)
    return int(w[R, 0] & 1), uniform53(w[:R])


def ladder_swaps(E_total, betas, parity, u):
    """neighbour pairs (i, i+1) from `parity` on, accepted when d = (b_i - b_{i+1})(E_i - E_{i+1}) >= 0 or
    u_i < exp(d) (reference mcmc.py:370-384).  Returns the new order of the replicas and the (tried, accepted)
    counts."""
    R = len(E_total)
    order = np.arange(R)
    tried = acc = 0
    for i in range(parity, R - 1, 2):
        tried += 1
        d = (betas[i] - betas[i + 1]) * (E_total[i] - E_total[i + 1])
        if d >= 0 or u[i] < np.exp(d):
            order[i], order[i + 1] = order[i + 1], order[i]
            acc += 1
    return order, tried, acc


class _Record:
    """what a run accumulates chunk by chunk (the energies.npz fields) plus the counters a resume needs."""

    def __init__(self, R):
        self.E_chunk, self.acc_chunk, self.drift_log, self.traj = [], [], [], []
        self.n_ex = self.n_ex_try = self.n_ex_acc = 0
        self.drift = 0.0
        self.acc_prev = np.zeros(R, np.int64)

    def traj_array(self, R):
        return np.concatenate(self.traj, axis=1) if self.traj else np.zeros((R, 0))


def _save_checkpoint(path, c_done, st, rec, want, betas):
    """reference keys (mcmc.py:389-399): chunk, sig, spin, key, drift, n_ex_try, n_ex_acc, E_chunk, acc_chunk,
    drift_log, traj, snaps, snap_chunks, spin_snaps, cfg_betas and cfg_*; `key` holds this build's stream state (the
    replicas' step counters and the exchange count) and n_accept the accepted-move counters.  spin and spin_snaps
    are empty: spin models are refused at setup."""
    R = st.R
    tmp = path + ".tmp.npz"
    np.savez(tmp, chunk=np.int64(c_done), sig=st.occ.cpu().numpy(), spin=np.zeros(0),
             key=np.concatenate([st.step.cpu().numpy(), [rec.n_ex]]).astype(np.int64),
             n_accept=st.n_accept.cpu().numpy(), drift=np.float64(rec.drift), n_ex_try=np.int64(rec.n_ex_try),
             n_ex_acc=np.int64(rec.n_ex_acc), E_chunk=np.array(rec.E_chunk).reshape(-1, R),
             acc_chunk=np.array(rec.acc_chunk).reshape(-1, R), drift_log=np.array(rec.drift_log, np.float64),
             traj=rec.traj_array(R), snaps=np.zeros((0, R, st.nn.N), np.int8), snap_chunks=np.zeros(0, np.int64),
             spin_snaps=np.zeros((0, R, st.nn.N), np.float16), cfg_betas=np.asarray(betas, np.float64), **{"cfg_" + k: v for k, v in want.items()})
    os.replace(tmp, path)


def _load_checkpoint(path, want, betas):
    ck = np.load(path, allow_pickle=False)
    have = {k: (str(ck["cfg_" + k]) if isinstance(v, str) else int(ck["cfg_" + k])) for k, v in want.items()
            if "cfg_" + k in ck.files}
    same_T = np.allclose(np.asarray(ck["cfg_betas"], np.float64), np.asarray(betas, np.float64), rtol=0, atol=1e-6)
    if have != want or not same_T:
        diff = sorted(k for k in want if have.get(k) != want[k]) + ([] if same_T else ["temperatures"])
        raise RuntimeError(f"{path} was written by another run (it differs in {', '.join(diff)}); remove the file "
                           f"to start this run from the beginning")
    return ck


def _run_replicas(mc, betas, n_steps, seed, chunk, log_every, sample_interval, out_dir, tag, exchange_every=0,
                  anchor_every=None, init_npz=None, checkpoint_every=0):
    """R replicas of the NeuralCE model, replica r at betas[r]; neighbour exchanges when exchange_every > 0.
    anchor_every: chunks between full recomputes (default ceil(N / chunk), about one per sweep).
    init_npz: starting occupations (R, N) in canonical order (default: the setup's configuration everywhere).
    checkpoint_every: chunks between checkpoints, rounded up to a multiple of anchor_every so a checkpoint follows a
    full recompute and a resumed run repeats the uninterrupted one."""
    nn = mc["nn"]
    R = len(betas)
    betas = np.asarray(betas, np.float32).astype(np.float64)          # the reference keeps beta in float32
    if chunk % sample_interval != 0:
        raise ValueError(f"chunk ({chunk}) must be a multiple of sample_interval ({sample_interval})")
    n_chunks = max(n_steps // chunk, 1)
    if anchor_every is None:
        anchor_every = max(1, int(np.ceil(nn.N / chunk)))
    sig0 = np.broadcast_to(np.asarray(mc["sig0"], np.int8), (R, nn.N)) if init_npz is None else np.asarray(init_npz)
    ck_every = int(checkpoint_every or 0)
    if ck_every:
        ck_every = -(-ck_every // anchor_every) * anchor_every
    ck_path = os.path.join(out_dir, _CK_NAME)
    want = dict(n_steps=int(n_steps), chunk=int(chunk), R=int(R), n_atoms=int(nn.N),
                sample_interval=int(sample_interval), seed=int(seed), anchor_every=int(anchor_every),
                exchange_every=int(exchange_every), model=_fingerprint(nn, sig0))
    rec = _Record(R)
    c_first = 0
    resume = None
    if ck_every and os.path.exists(ck_path):
        resume = _load_checkpoint(ck_path, want, betas)
        c_first = int(resume["chunk"])
        sig0 = resume["sig"]
        rec.E_chunk = list(resume["E_chunk"]); rec.acc_chunk = list(resume["acc_chunk"])
        rec.drift_log = [float(x) for x in resume["drift_log"]]
        rec.traj = [resume["traj"]] if resume["traj"].shape[1] else []
        rec.n_ex = int(resume["key"][-1]); rec.n_ex_try = int(resume["n_ex_try"])
        rec.n_ex_acc = int(resume["n_ex_acc"]); rec.drift = float(resume["drift"])
    st = NNJobState(nn, sig0, betas, seed=seed, precision=mc["precision"], sample_interval=sample_interval,
                    samples_per_chunk=chunk // sample_interval)
    if resume is not None:
        import torch
        st.step.copy_(torch.as_tensor(resume["key"][:-1], dtype=torch.int64))
        st.n_accept.copy_(torch.as_tensor(resume["n_accept"], dtype=torch.int64))
        rec.acc_prev = resume["n_accept"].astype(np.int64)
        _log(f"[{tag}] continuing from {ck_path} after chunk {c_first} of {n_chunks}")
    ex_rounds = max(chunk // exchange_every, 1) if (exchange_every > 0 and R > 1) else 0
    _log(f"[{tag}] {R} replicas x {n_chunks} chunks of {chunk} steps on {nn.N} sites; caches rebuilt every "
         f"{anchor_every} chunk(s)" + (f", {ex_rounds} exchange round(s) per chunk" if ex_rounds else "")
         + (f", checkpoint every {ck_every} chunks" if ck_every else ""))
    log_chunks = max(log_every // chunk, 1) if log_every else 0
    t0 = time.time()
    for c in range(c_first, n_chunks):
        st.run(chunk)
        # (a BF16 replica run over several workgroups returns NaN if its workgroups lost each other; the check
        # comes before a full recompute could replace the energy)
        if not bool(st.energy.isfinite().all().item()):
            raise RuntimeError(f"[{tag}] chunk {c}: non-finite replica energies after the local-update kernel")
        last = c == n_chunks - 1
        if (c + 1) % anchor_every == 0 or last:
            E_run, E_full = st.anchor()
            rec.drift = float((E_full - E_run).abs().max().item())
        E_eV = st.energy.cpu().numpy() * nn.ys + nn.ym
        n_acc = st.n_accept.cpu().numpy()
        rec.E_chunk.append(E_eV)
        rec.acc_chunk.append((n_acc - rec.acc_prev) / chunk)
        rec.acc_prev = n_acc
        rec.drift_log.append(rec.drift)
        rec.traj.append(st.traj.cpu().numpy().T * nn.ys + nn.ym)
        for _ in range(ex_rounds):
            parity, u = exchange_draws(seed, rec.n_ex, R)
            rec.n_ex += 1
            order, tried, acc = ladder_swaps(st.energy.cpu().numpy() * nn.de_scale, betas, parity, u)
            rec.n_ex_try += tried; rec.n_ex_acc += acc
            if acc:
                st.permute(order)
        if log_chunks and ((c + 1) % log_chunks == 0 or last):
            rate = R * chunk * (c + 1 - c_first) / max(time.time() - t0, 1e-9)
            _log(f"[{tag}] {c + 1}/{n_chunks}: E {E_eV.mean():.6f} eV/atom ({E_eV.min():.6f} .. {E_eV.max():.6f}), "
                 f"accepted {np.mean(rec.acc_chunk[-1]):.3f}, rebuild drift {rec.drift:.1e}, {rate:,.0f} attempts/s"
                 + (f", exchanges {rec.n_ex_acc}/{rec.n_ex_try}" if ex_rounds else ""))
        if ck_every and (c + 1) % ck_every == 0 and not last:
            os.makedirs(out_dir, exist_ok=True)
            rec.traj = [rec.traj_array(R)]
            _save_checkpoint(ck_path, c + 1, st, rec, want, betas)
    wall = time.time() - t0
    os.makedirs(out_dir, exist_ok=True)
    E_chunk = np.array(rec.E_chunk); acc_chunk = np.array(rec.acc_chunk); drift = np.array(rec.drift_log)
    np.savez(os.path.join(out_dir, "energies.npz"), E_chunk=E_chunk, acc_chunk=acc_chunk, drift=drift,
             drift_eV=drift * nn.ys, traj=rec.traj_array(R), betas=betas, temps=1.0 / (KB * betas))
    sig_in = st.occ.cpu().numpy()[:, mc["inv_perm"]]
    np.save(os.path.join(out_dir, "final_sigma.npy"), sig_in)
    if nn.N <= int(os.environ.get("CE_CNN_CIF_MAX", "200000")) and mc.get("atoms") is not None:
        _write_cifs(mc, sig_in, 1.0 / (KB * betas), out_dir)
    if ck_every and os.path.exists(ck_path):
        os.remove(ck_path)
    return dict(steps_per_s=float(R * chunk * (n_chunks - c_first) / max(wall, 1e-9)), n_replicas=R,
                n_steps=int(n_steps), chunk=int(chunk), E_final=[float(x) for x in E_chunk[-1]],
                acc_mean=float(acc_chunk.mean()), drift_max=float(np.max(drift)),
                exchange_acc=(rec.n_ex_acc / rec.n_ex_try if rec.n_ex_try else None), model_depth=int(mc["depth"]),
                n_atoms=int(nn.N), resumed_from_chunk=int(c_first))


def _write_cifs(mc, sig_in, temps, out_dir):
    """final_T<T>_r<r>.cif per replica: the input structure with its active sites decorated by the final species."""
    from ..lattice import ATOMIC_NUMBERS, write_cif
    Z_of = np.array([ATOMIC_NUMBERS[s] for s in mc["nn"].geo["species"]], np.int64)
    for r in range(sig_in.shape[0]):
        at = mc["atoms"].copy()
        at.numbers = Z_of[np.asarray(sig_in[r], np.int64)]
        write_cif(os.path.join(out_dir, f"final_T{temps[r]:.0f}_r{r}.cif"), at)


def _prepare(model_path, init_cif, supercell=(1, 1, 1), seed_idx=0, dtype="float32", seed=0, **_):
    """model + starting cell -> the setup() dict (reference mcmc.py _prepare): the cif is tiled by `supercell`; the
    tiled decoration is shuffled with the run seed unless CE_CNN_SUPERCELL_PERIODIC=1 keeps it periodic."""
    from ..lattice import read_cif
    from .mc_model import load_model, setup
    mdl = load_model(model_path)
    atoms = read_cif(init_cif)
    reps = tuple(int(x) for x in supercell)
    if reps != (1, 1, 1):
        atoms = atoms.repeat(reps)
        if os.environ.get("CE_CNN_SUPERCELL_PERIODIC", "0") != "1":
            atoms.numbers = np.random.default_rng(seed).permutation(atoms.numbers)
    t0 = time.time()
    mc = setup(mdl, atoms, seed_idx=seed_idx, dtype=dtype)
    mc["mdl"] = mdl
    mc["atoms"] = atoms
    _log(f"NeuralCE {os.path.basename(os.path.dirname(os.path.abspath(model_path)))}: {mc['n_atoms']} sites, "
         f"{mc['depth']} hidden layer(s), tables built in {time.time() - t0:.1f} s; starting energy "
         f"{mc['energy'](mc['sig0']):.6f} eV/atom")
    return mc


def _start_from_npz(mc, path, temps):
    """per-replica starting occupations (R, N) in canonical order from an --init_npz file."""
    from ..lattice import ATOMIC_NUMBERS
    from ..mc.init_structures import load_init_npz
    numbers, _ = load_init_npz(path, mc["mdl"]["primitive"], mc["atoms"], temps)
    index_of = {ATOMIC_NUMBERS[s]: k for k, s in enumerate(mc["nn"].geo["species"])}
    bad = sorted(set(np.unique(numbers).tolist()) - set(index_of))
    if bad:
        raise ValueError(f"{path}: atomic numbers {bad} are not species of the model")
    lut = np.zeros(256, np.int8)
    for z, k in index_of.items():
        lut[z] = k
    return lut[numbers.astype(np.int64)][:, mc["perm"]]


def _drive(mc, temps, tag, n_steps, seed, chunk, log_every, sample_interval, out_dir, exchange_every=0,
           anchor_every=None, init_npz=None, checkpoint_every=0):
    temps = np.asarray(temps, float)
    start = _start_from_npz(mc, init_npz, temps) if init_npz else None
    if init_npz:
        _log(f"[{tag}] replica starting structures from {init_npz}")
    return _run_replicas(mc, 1.0 / (KB * temps), n_steps, seed, min(chunk, n_steps), log_every, sample_interval,
                         out_dir, tag, exchange_every=exchange_every, anchor_every=anchor_every, init_npz=start,
                         checkpoint_every=checkpoint_every)


def run_pt_mc(model_path, init_cif, n_steps=10000, T_min=100.0, T_max=2000.0, n_replicas=32, T_grid=None,
              exchange_every=100, seed=0, log_every=0, out_dir=".", chunk=1000, sample_interval=100,
              supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None, init_npz=None,
              checkpoint_every=0, **_):
    """parallel tempering on a geometric ladder (or T_grid); beta = 1 / (kB T), kB = 8.617333262e-5 (reference
    mcmc.py:466-475)."""
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    temps = np.asarray(T_grid, float) if T_grid is not None else np.geomspace(T_min, T_max, n_replicas)
    res = _drive(mc, temps, "nce PT", n_steps, seed, chunk, log_every, sample_interval, out_dir,
                 exchange_every=exchange_every, anchor_every=anchor_every, init_npz=init_npz,
                 checkpoint_every=checkpoint_every)
    res.update(temps=[float(t) for t in temps], exchange_every=int(exchange_every))
    return res


def run_mc(model_path, init_cif, n_steps=10000, temperature=300.0, n_replicas=1, chunk=1000, sample_interval=100,
           seed=0, log_every=0, out_dir=".", supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None,
           init_npz=None, checkpoint_every=0, **_):
    """n_replicas independent chains at one temperature."""
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    res = _drive(mc, [float(temperature)] * int(n_replicas), f"nce T={temperature:.0f}", n_steps, seed, chunk,
                 log_every, sample_interval, out_dir, anchor_every=anchor_every, init_npz=init_npz,
                 checkpoint_every=checkpoint_every)
    res.update(temperature=float(temperature))
    return res


def run_multiT_mc(model_path, init_cif, temps, n_steps=10000, chunk=1000, sample_interval=100, seed=0, log_every=0,
                  out_dir=".", supercell=(1, 1, 1), seed_idx=0, dtype="float32", anchor_every=None, init_npz=None,
                  checkpoint_every=0, **_):
    """one independent chain per temperature, no exchange."""
    mc = _prepare(model_path, init_cif, supercell, seed_idx, dtype, seed=seed)
    res = _drive(mc, temps, "nce multiT", n_steps, seed, chunk, log_every, sample_interval, out_dir,
                 anchor_every=anchor_every, init_npz=init_npz, checkpoint_every=checkpoint_every)
    res.update(temps=[float(t) for t in temps])
    return res
