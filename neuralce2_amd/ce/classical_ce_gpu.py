"""``GPUClusterExpansion`` on the gfx950 kernels (reference: neuralce/ce/classical_ce_gpu.py:496-593).

Only the template-only ("implicit") table mode exists here: memory is the state plus a template of a few hundred
rows, sites are in canonical order (``perm`` / ``inv_perm`` / ``encode``), and the energy and swap energy change are
HIP kernels called through libneuralce_hip.so.  ``table_mode`` values 'tiled' and 'legacy' of the reference select
O(N) tables with the same mathematics; they are accepted and served by the implicit tables.
"""
from __future__ import annotations

import numpy as np

from .ce_gpu_implicit import CETemplate, build_ce_gpu_implicit, encode_canonical
from ..clusterspace import point_function_table


def extract_point_functions(cs):
    """pf[sl, bf, species] (reference classical_ce_gpu.py:80-141, which probes icet): bf 0 = 1, bf 1 = theta_0."""
    n_sl = int(np.max(cs.basis_to_sl)) + 1
    M = int(np.max(cs.M_basis))
    pf = np.zeros((n_sl, max(M, 1), max(M, 1)))
    pf[:, 0, :] = 1.0
    if M > 1:
        tab = point_function_table(M)
        for sl in range(n_sl):
            pf[sl] = tab
    return pf.astype(np.float32), n_sl


_TEMPLATE_CACHE = {}


def _template(ce):
    key = (id(ce._cluster_space), ce.parameters.tobytes())
    t = _TEMPLATE_CACHE.get(key)
    if t is None:
        t = _TEMPLATE_CACHE[key] = CETemplate(ce._cluster_space, ce.parameters)
    return t


class GPUClusterExpansion:
    def __init__(self, ce, supercell, fractional_position_tolerance=None, table_mode=None, energy_scale=1.0,
                 device=0):
        import torch
        from .._lib import Context
        self.ce = ce
        self.supercell = supercell
        self.implicit = True
        self.table_mode = table_mode or "implicit"
        self.energy_scale = float(energy_scale)
        self.info = info = build_ce_gpu_implicit(ce, supercell, energy_scale=energy_scale, template=_template(ce))
        self.tp = info
        self.perm = info['perm']; self.inv_perm = info['inv_perm']; self.site_to_sl = info['site_to_sl']
        self._n_atoms = info['n_atoms']; self._n_instances = info['n_instances']; self._max_size = info['max_size']
        nB = info['n_basis']
        self.sublattices = []
        for sl in info['act_sl']:
            bs = np.where(info['basis_to_sl'] == sl)[0]
            idx = (np.arange(info['n_cells'])[:, None] * nB + bs[None]).reshape(-1).astype(np.int64)
            sp = list(info['sl_species'][sl])
            self.sublattices.append({'indices': np.sort(idx), 'n_species': len(sp), 'chemical_symbols': sp,
                                     'name': "".join(sp)})
        if not torch.cuda.is_available():
            raise RuntimeError("GPUClusterExpansion needs a ROCm GPU (libneuralce_hip.so); none is visible")
        self.device = torch.device("cuda", device)
        self.ctx = Context.get(device)
        self.model = self.ctx.ce_model_create(info)
        # rounding scale of one FP32 swap energy change: unit roundoff x factor x the largest sum of |coefficient|
        # over a window, for the two windows of a swap (the re-anchor schedule of the mixed form is set from it)
        coef = np.abs(np.asarray(info['template'].ent_coef, np.float64))
        self.fp32_step_error = 2.0 ** -24 * abs(float(info['factor'])) * 2.0 * float(coef.sum(axis=-1).max())
        self.spin_tables = None
        self._has_spin = False

    # ── properties (classical_ce_gpu.py:562-579) ──
    @property
    def n_atoms(self):
        return self._n_atoms

    @property
    def n_instances(self):
        return self._n_instances

    @property
    def max_size(self):
        return self._max_size

    @property
    def has_spin(self):
        return False

    def encode(self, struct):
        return encode_canonical(self.info, self.ce._cluster_space, struct)

    # ── energies ──
    def _occ(self, sigma):
        import torch
        s = torch.as_tensor(np.asarray(sigma) if not hasattr(sigma, "device") else sigma)
        s = s.to(device=self.device, dtype=torch.int8)
        if s.dim() == 1:
            s = s[None]
        if s.shape[1] != self.n_atoms:
            raise ValueError(f"sigma has {s.shape[1]} sites, the cell has {self.n_atoms}")
        if int(s.min()) < 0 or int(s.max()) > 1:
            raise ValueError("species index out of range for a binary cluster expansion")
        return s.contiguous()

    def batched_energy(self, sigma_batch):
        """(R, N) -> (R,) float64 total energies (GPU)."""
        import torch
        occ = self._occ(sigma_batch)
        E = torch.empty(occ.shape[0], dtype=torch.float64, device=self.device)
        self.ctx.ce_energy(self.model, occ, E)
        return E.cpu().numpy()

    def energy(self, sigma):
        return float(self.batched_energy(sigma)[0])

    def energy_a(self, tp, sigma):
        return self.energy(sigma)

    def delta_energy_a(self, tp, i_a, i_b, sigma, s_a_new, s_b_new, precision=0):
        """E(after) - E(before) of the swap of sites i_a and i_b (canonical order)."""
        import torch
        occ = self._occ(sigma)
        sa, sb = int(occ[0, int(i_a)]), int(occ[0, int(i_b)])
        if (int(s_a_new), int(s_b_new)) != (sb, sa):
            raise NotImplementedError("only swap moves (s_a_new = sigma[i_b], s_b_new = sigma[i_a]) are supported")
        ia = torch.tensor([int(i_a)], dtype=torch.int64, device=self.device)
        ib = torch.tensor([int(i_b)], dtype=torch.int64, device=self.device)
        dE = torch.empty(1, dtype=torch.float64, device=self.device)
        self.ctx.ce_delta(self.model, occ, ia, ib, dE, precision)
        return float(dE.cpu()[0])

    def __repr__(self):
        return (f"GPUClusterExpansion(n_atoms={self.n_atoms}, n_instances={self.n_instances}, "
                f"max_size={self.max_size}, sublattices={len(self.sublattices)})")
