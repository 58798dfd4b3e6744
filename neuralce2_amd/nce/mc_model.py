"""NeuralCE MC machinery on the gfx950 kernels (reference: neuralce/nce/mc_model.py:60-228).

``setup(mdl, atoms, seed_idx, dtype)`` returns the reference's keys (energy, full_cache, perm / inv_perm, sig0, ys, ym,
n_atoms, propose, ...) backed by libneuralce_hip.so: the layer-0 template, shells and network are uploaded once, caches
and local updates are HIP kernels.  Only the implicit (template-only) table mode exists; sites are in canonical order.
"""
from __future__ import annotations

import numpy as np

from ..clusterspace import ClusterSpace
from .gpu_layer0_implicit import build_geometry
from .model_io import check_supported, ensemble_params, expand_params, load_model  # noqa: F401


def silu_network(c0, width, depth, n_shells, seed=0, glorot=True):
    """Parameters of the SiLU graph-convolution variant (Glorot-uniform weights from `seed`): per layer self and
    shell maps, biases 0; per-site readout lin (layer 0) and w (last layer)."""
    rng = np.random.default_rng(seed)

    def glorot_u(fan_in, fan_out, shape):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, size=shape)
    layers = []
    cin = c0
    for _ in range(depth):
        layers.append(dict(self=glorot_u(cin, width, (cin, width)),
                           shell=glorot_u(cin, width, (n_shells, cin, width)), b=np.zeros(width)))
        cin = width
    return dict(lin=glorot_u(c0, 1, (c0,)), layers=layers, head=dict(w=glorot_u(cin, 1, (cin,)), b=np.array(0.0)))


def default_cache_mode(P):
    """The reference's choice of cache layout (nce/mc_model.py:202-206): 'h0' (every site's C0 layer-0 channels
    cached) unless the first layer's projections, (P (x2 with a gate) + 1) x D values per site, are fewer, then
    'proj'.  Here 'proj' drops the layer-0 rows from the cache: the local update never reads them, since it caches
    every hidden layer's pre-activations and changes the first by the projected layer-0 change."""
    if not P["layers"]:
        return "h0"
    L1 = P["layers"][0]
    C0 = len(np.ravel(P["lin"]))
    D = int(np.shape(L1["self"])[1])
    n_sh = int(np.shape(L1["shell"])[0])
    return "h0" if C0 <= (n_sh * (2 if "gate" in L1 else 1) + 1) * D else "proj"


class NNModel:
    """A NeuralCE model uploaded for one cell: geometry, parameters, sizes."""

    def __init__(self, cs, atoms, P, mu, sd, shell_edges, act="gate", readout="mean", ys=1.0, ym=0.0, device=0,
                 cache_mode=None):
        import torch
        from .._lib import Context
        if not torch.cuda.is_available():
            raise RuntimeError("the NeuralCE MC path needs a ROCm GPU (libneuralce_hip.so); none is visible")
        check_supported(P)
        self.cs = cs
        self.P, self.mu, self.sd = P, mu, sd
        self.depth = len(P["layers"])
        self.geo = build_geometry(cs, atoms, shell_edges, self.depth)
        self.N = self.geo["N"]; self.nA = self.geo["nA"]
        self.ys, self.ym = float(ys), float(ym)
        self.readout = readout
        self.device = torch.device("cuda", device)
        self.ctx = Context.get(device)
        de_scale = self.ys * self.N if readout == "mean" else 1.0
        self.cache_mode = cache_mode or default_cache_mode(P)
        if self.cache_mode not in ("h0", "proj"):
            raise ValueError(f"cache_mode must be 'h0' or 'proj', not {self.cache_mode!r}")
        self.model, self.cache_elems, self.scratch_elems = self.ctx.nn_model_create(
            self.geo, P, mu, sd, act=act, readout=readout, de_scale=de_scale, cache_mode=self.cache_mode)
        self.de_scale = de_scale

    def alloc(self, R, precision):
        import torch
        dt = torch.float64 if precision == 0 else torch.float32
        cache = torch.empty((R, self.cache_elems), dtype=dt, device=self.device)
        scratch = torch.empty((R, self.scratch_elems), dtype=dt, device=self.device)
        return cache, scratch

    def full_cache(self, occ, precision=1):
        """occ (R, N) int8 device tensor -> (cache, energies in model units)."""
        import torch
        cache, _ = self.alloc(occ.shape[0], precision)
        E = torch.empty(occ.shape[0], dtype=torch.float64, device=self.device)
        self.ctx.nn_full_cache(self.model, occ, cache, E, precision)
        return cache, E

    def occ(self, sig):
        import torch
        s = torch.as_tensor(np.asarray(sig)).to(device=self.device, dtype=torch.int8)
        if s.dim() == 1:
            s = s[None]
        return s.contiguous()


def setup(mdl, atoms, seed_idx=0, dtype="float32", table=None, device=0, cache_mode=None):
    """Build the MC machinery for `atoms` (an ideal-lattice supercell of the model's primitive); see the module doc.
    seed_idx < 0 selects the seed ensemble (mean energy over every saved parameter set), run as one network whose
    width is the sum of the seeds' widths (model_io.ensemble_params).  cache_mode 'h0' / 'proj' / None (the
    reference's choice by per-site bytes, default_cache_mode), as setup_implicit's (nce/mc_model.py:175,202-206)."""
    if int(mdl.get("spin", 0) or 0):
        raise NotImplementedError("spin models (spin channels in layer 0, spin moves; reference nce/spin_features.py) "
                                  "are not supported by the gfx950 local update")
    cs = ClusterSpace(mdl["primitive"], list(mdl["cutoffs"]), mdl["chemical_symbols"])
    if seed_idx is not None and int(seed_idx) < 0:
        P, mu, sd = ensemble_params(mdl)
    else:
        P, mu, sd = expand_params(mdl, 0 if seed_idx is None else int(seed_idx))
    nn = NNModel(cs, atoms, P, mu, sd, list(mdl["shell_edges"]), act="gate", readout="mean",
                 ys=mdl["ys"], ym=mdl["ym"], device=device, cache_mode=cache_mode)
    prec = {"float64": 0, None: 0, "bfloat16": 2, "bf16": 2}.get(dtype, 1)
    geo = nn.geo

    def energy(sig, spin_state=None):
        """eV/atom, FP64 full recompute (the reference's gate path)."""
        _, E = nn.full_cache(nn.occ(sig), precision=0)
        return float(E[0].item()) * nn.ys + nn.ym

    def full_cache(sig, spin_state=None):
        """device state of one configuration (occupations, caches, running energy in model units), the `cache` that
        local_delta and commit take (local_update_implicit.py:89-104)."""
        from .mcmc import NNJobState
        return NNJobState(nn, np.asarray(sig), [1.0], precision=prec)

    def local_delta_full(sig, flips):
        """energy change (model units) of applying `flips` [[site, species], ...], from two full recomputes."""
        s0 = np.asarray(sig).copy(); s1 = s0.copy()
        for site, sp in np.asarray(flips):
            s1[int(site)] = int(sp)
        _, E = nn.full_cache(nn.occ(np.stack([s0, s1])), precision=0)
        return float(E[1].item() - E[0].item())

    def local_delta(cache, sig, flips, spin_state=None, sflips=None):
        """(dE in model units, patch, sig1) of the swap flips = [[i, sigma_j], [j, sigma_i]] by the local update on
        the GPU, nothing committed (local_update_implicit.py:117-183).  `cache` is a full_cache(sig) state."""
        import torch
        if sflips is not None:
            raise NotImplementedError("spin flips are not part of the canonical swap path")
        fl = np.asarray(flips, np.int64).reshape(2, 2)
        i, j = int(fl[0, 0]), int(fl[1, 0])
        sig1 = np.asarray(sig).copy()
        sig1[i], sig1[j] = fl[0, 1], fl[1, 1]
        if not (int(fl[0, 1]) == int(np.asarray(sig)[j]) and int(fl[1, 1]) == int(np.asarray(sig)[i])):
            raise ValueError("flips must be the swap [[i, sigma[j]], [j, sigma[i]]]")
        inj = (torch.tensor([[i]], device=nn.device), torch.tensor([[j]], device=nn.device),
               torch.zeros((1, 1), dtype=torch.float64, device=nn.device))
        dE = torch.empty(1, dtype=torch.float64, device=nn.device)
        nn.ctx.nn_delta([cache.job(inj)], dE, prec)
        return float(dE.item()), (i, j), sig1

    def commit(cache, patch, accept):
        """apply the swap `patch` to the caches, occupations and running energy when accept (a forced-accept step
        of the same proposal; local_update_implicit.py:185-194)."""
        import torch
        if not accept:
            return cache
        i, j = patch
        inj = (torch.tensor([[i]], device=nn.device), torch.tensor([[j]], device=nn.device),
               torch.zeros((1, 1), dtype=torch.float64, device=nn.device))
        cache.run(1, inj)
        return cache

    def propose(key):
        """(i, j): two sites uniform over the cell from the counter-based stream of `key` (an int)."""
        from .._lib import philox4x32_10
        x = [int(v) for v in philox4x32_10([int(key) & 0xffffffff, (int(key) >> 32) & 0xffffffff, 0, 0], [0x5EED, 0])]
        return (x[0] | (x[1] << 32)) % nn.N, (x[2] | (x[3] << 32)) % nn.N

    return dict(cs=cs, P=P, mu=mu, sd=sd, depth=nn.depth, ys=nn.ys, ym=nn.ym, n_atoms=nn.N, n_active=nn.nA,
                sig0=geo["sig"], energy=energy, full_cache=full_cache, local_delta_full=local_delta_full,
                local_delta=local_delta, commit=commit, propose=propose,
                to_eV_per_atom=lambda E_std: E_std * nn.ys + nn.ym, dE_eV_total=lambda dE: dE * nn.ys * nn.N,
                implicit=True, geo=geo, perm=geo["perm"], inv_perm=geo["inv_perm"], nn=nn, dtype=dtype,
                precision=prec, cache_mode=nn.cache_mode)
