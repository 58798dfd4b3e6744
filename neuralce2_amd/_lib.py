"""ctypes binding of libneuralce_hip.so (declarations in include/neuralce_hip.h).

The shared library is the only compute path: importing this module on a machine where the library is missing or
cannot be loaded raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NCE_HIP_LIB", os.path.join(_HERE, "libneuralce_hip.so"))

PREC_FP64 = 0
PREC_MIXED = 1
PREC_BF16 = 2          # NeuralCE: FP32 caches, graph-convolution projections on BF16 MFMA

_p = C.c_void_p


class ModelDesc(C.Structure):
    _fields_ = [("n_basis", C.c_int32), ("max_order", C.c_int32), ("n_entries", C.c_int32),
                ("max_window", C.c_int32), ("n_rows", C.c_int32), ("n_sublattices", C.c_int32),
                ("n_active", C.c_int32), ("max_active_basis", C.c_int32), ("d", C.c_int32 * 3),
                ("n_cells", C.c_int64), ("factor", C.c_double), ("zerolet", C.c_double),
                ("entries_per_basis", _p), ("window_per_basis", _p), ("entry_slot", _p), ("entry_coef", _p),
                ("window", _p), ("rows", _p), ("row_bf", _p), ("row_coef", _p), ("pf", _p), ("basis_to_sl", _p),
                ("active_basis", _p), ("active_nbasis", _p), ("active_cum", _p)]


class PTJob(C.Structure):
    _fields_ = [("model", C.c_int32), ("n_replicas", C.c_int32), ("n_sites", C.c_int64), ("occ", _p),
                ("energy", _p), ("beta", _p), ("slot_of_rep", _p), ("step", _p), ("seed", C.c_uint64),
                ("n_accept", _p), ("e_min", _p), ("traj", _p), ("sample_interval", C.c_int32),
                ("accumulate", C.c_int32), ("e_ref", _p), ("s1", _p), ("s2", _p), ("inj_ia", _p), ("inj_ib", _p),
                ("inj_u", _p), ("acc_n", _p), ("acc_mean", _p), ("acc_m2", _p)]


class ExchangeJob(C.Structure):
    _fields_ = [("n_replicas", C.c_int32), ("inj_parity", C.c_int32), ("energy", _p), ("rep_of_slot", _p),
                ("slot_of_rep", _p), ("beta_slot", _p), ("ladder_id", _p), ("beta_rep", _p), ("seed", C.c_uint64),
                ("ex_index", C.c_uint64), ("n_swaps", _p), ("inj_u", _p), ("traj_last", _p), ("ex_count", _p)]


class NNModelDesc(C.Structure):
    _fields_ = [("d", C.c_int32 * 3), ("n_cells", C.c_int64), ("n_window", C.c_int32), ("window", _p),
                ("n_entries", C.c_int32), ("entry_slot", _p), ("entry_col", _p), ("entry_coef", _p),
                ("n_shells", C.c_int32), ("shell_count", _p), ("partners", _p), ("c0", C.c_int32), ("mu", _p),
                ("sd", _p), ("lin", _p), ("n_layers", C.c_int32), ("layer_cin", _p), ("layer_d", _p),
                ("w_self", _p), ("w_shell", _p), ("w_gate", _p), ("b", _p), ("gb", _p), ("activation", C.c_int32),
                ("head_w", _p), ("head_b", C.c_double), ("readout", C.c_int32), ("de_scale", C.c_double),
                ("ball_cap", C.c_int32 * 8), ("cache_mode", C.c_int32)]


class NNJob(C.Structure):
    _fields_ = [("model", C.c_int32), ("n_replicas", C.c_int32), ("n_sites", C.c_int64), ("occ", _p), ("cache", _p),
                ("scratch", _p), ("energy", _p), ("beta", _p), ("step", _p), ("seed", C.c_uint64),
                ("n_accept", _p), ("traj", _p), ("sample_interval", C.c_int32), ("pad", C.c_int32),
                ("inj_ia", _p), ("inj_ib", _p), ("inj_u", _p)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          f"or neuralce2_amd/csrc/build.sh")
    lib = C.CDLL(LIB_PATH)
    lib.nce_last_error.restype = C.c_char_p
    lib.nce_ctx_create.argtypes = [C.c_int32, C.POINTER(_p)]
    lib.nce_ctx_destroy.argtypes = [_p]
    lib.nce_ctx_destroy.restype = None
    lib.nce_ce_model_create.argtypes = [_p, C.POINTER(ModelDesc), C.POINTER(C.c_int32)]
    lib.nce_ce_energy.argtypes = [_p, C.c_int32, _p, C.c_int32, C.c_int64, _p, _p]
    lib.nce_ce_energy_noalloc.argtypes = [_p, C.c_int32, _p, C.c_int32, C.c_int64, _p, _p]
    lib.nce_ce_delta.argtypes = [_p, C.c_int32, _p, C.c_int32, C.c_int64, _p, _p, _p, C.c_int32, _p]
    lib.nce_pt_run_chunk.argtypes = [_p, _p, C.c_int32, C.c_int64, C.c_int32, C.c_int32, _p]
    lib.nce_pt_exchange.argtypes = [_p, _p, C.c_int32, _p]
    lib.nce_pt_cycle_create.argtypes = [_p, _p, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32, _p, C.c_int32,
                                        C.POINTER(_p)]
    lib.nce_pt_cycle_launch.argtypes = [_p, C.c_int32, _p]
    lib.nce_pt_cycle_destroy.argtypes = [_p]
    lib.nce_pt_cycle_destroy.restype = None
    lib.nce_nn_model_create.argtypes = [_p, C.POINTER(NNModelDesc), C.POINTER(C.c_int32)]
    lib.nce_nn_sizes.argtypes = [_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.nce_nn_full_cache.argtypes = [_p, C.c_int32, _p, C.c_int32, C.c_int64, _p, _p, C.c_int32, _p]
    lib.nce_nn_run_chunk.argtypes = [_p, _p, C.c_int32, C.c_int64, C.c_int32, _p]
    lib.nce_nn_delta.argtypes = [_p, _p, C.c_int32, _p, C.c_int32, _p]
    lib.nce_philox4x32_10.argtypes = [_p, _p, _p]
    lib.nce_philox4x32_10.restype = None
    return lib


lib = _load()

EXPORTED = ("nce_last_error", "nce_ctx_create", "nce_ctx_destroy", "nce_ce_model_create", "nce_ce_energy",
            "nce_ce_energy_noalloc",
            "nce_ce_delta", "nce_pt_run_chunk", "nce_pt_exchange", "nce_pt_cycle_create", "nce_pt_cycle_launch",
            "nce_pt_cycle_destroy", "nce_philox4x32_10", "nce_nn_model_create", "nce_nn_sizes", "nce_nn_full_cache",
            "nce_nn_run_chunk", "nce_nn_delta")


class NCEError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise NCEError(f"libneuralce_hip error {rc}: {lib.nce_last_error().decode()}")


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32); k = np.ascontiguousarray(key, np.uint32); o = np.zeros(4, np.uint32)
    lib.nce_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def ptr(t):
    """device (or host) address of a torch tensor / numpy array, 0 for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Context:
    """One device context (nce_ctx); owns uploaded models."""

    _by_device = {}

    def __init__(self, device=0):
        h = _p()
        check(lib.nce_ctx_create(int(device), C.byref(h)))
        self.h = h
        self.device = int(device)
        self._keep = []

    @classmethod
    def get(cls, device=0):
        if device not in cls._by_device:
            cls._by_device[device] = cls(device)
        return cls._by_device[device]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib.nce_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def ce_model_create(self, info):
        """upload the model of a CE cell (dict from ce_gpu_implicit.build_ce_gpu_implicit)."""
        tpl = info['template']
        arrs = dict(
            entries_per_basis=np.ascontiguousarray(tpl.J_b, np.int32),
            window_per_basis=np.ascontiguousarray(tpl.W_b, np.int32),
            entry_slot=np.ascontiguousarray(tpl.ent_slot, np.uint16),
            entry_coef=np.ascontiguousarray(tpl.ent_coef, np.float64),
            window=np.ascontiguousarray(info['win'], np.int32),
            rows=np.ascontiguousarray(np.concatenate([tpl.T_b[..., None], info['T_uo']], -1), np.int32),
            row_bf=np.ascontiguousarray(tpl.T_bf, np.int32),
            row_coef=np.ascontiguousarray(tpl.T_coef, np.float64),
            pf=np.ascontiguousarray(tpl.pf, np.float64),
            basis_to_sl=np.ascontiguousarray(tpl.basis_to_sl, np.int32),
            active_basis=np.ascontiguousarray(tpl.act_b, np.int32),
            active_nbasis=np.ascontiguousarray(tpl.act_nb, np.int32),
            active_cum=np.ascontiguousarray(np.cumsum(tpl.act_p), np.float64),
        )
        d = ModelDesc()
        d.n_basis = tpl.nB; d.max_order = tpl.mo; d.n_entries = tpl.J; d.max_window = tpl.Wmax
        d.n_rows = tpl.Rt; d.n_sublattices = tpl.pf.shape[0]; d.n_active = len(tpl.act_nb)
        d.max_active_basis = tpl.act_b.shape[1]
        for k in range(3):
            d.d[k] = int(info['d'][k])
        d.n_cells = int(info['n_cells']); d.factor = float(info['factor']); d.zerolet = float(info['zerolet'])
        for k, a in arrs.items():
            setattr(d, k, a.ctypes.data)
        mid = C.c_int32()
        check(lib.nce_ce_model_create(self.h, C.byref(d), C.byref(mid)))
        return int(mid.value)

    def ce_energy(self, model, occ, energy, stream=None):
        R, N = occ.shape
        check(lib.nce_ce_energy(self.h, model, ptr(occ), R, N, ptr(energy), stream_ptr(stream)))

    def ce_energy_noalloc(self, model, occ, energy, stream=None):
        """nce_ce_energy with the partial sums in the library's fixed per-GPU scratch (no allocation)."""
        R, N = occ.shape
        check(lib.nce_ce_energy_noalloc(self.h, model, ptr(occ), R, N, ptr(energy), stream_ptr(stream)))

    def ce_delta(self, model, occ, ia, ib, dE, precision=PREC_FP64, stream=None):
        R, N = occ.shape
        check(lib.nce_ce_delta(self.h, model, ptr(occ), R, N, ptr(ia), ptr(ib), ptr(dE), int(precision),
                               stream_ptr(stream)))

    def pt_run_chunk(self, jobs, n_steps, precision=PREC_FP64, tile=0, stream=None):
        arr = (PTJob * len(jobs))(*jobs)
        check(lib.nce_pt_run_chunk(self.h, C.addressof(arr), len(jobs), int(n_steps), int(precision), int(tile),
                                   stream_ptr(stream)))

    def pt_exchange(self, jobs, stream=None):
        arr = (ExchangeJob * len(jobs))(*jobs)
        check(lib.nce_pt_exchange(self.h, C.addressof(arr), len(jobs), stream_ptr(stream)))

    def pt_cycle(self, jobs, n_steps, chunks_per_cycle, precision, tile, ex_jobs):
        """capture `chunks_per_cycle` chunks (+ moment merges) and the exchange of `ex_jobs` as one HIP graph."""
        return Cycle(self, jobs, n_steps, chunks_per_cycle, precision, tile, ex_jobs)

    # ── NeuralCE ──
    def nn_model_create(self, geo, P, mu, sd, act="gate", readout="mean", de_scale=1.0, cache_mode="h0"):
        """upload a NeuralCE model: geometry dict (gpu_layer0_implicit.build_geometry) and expanded parameters;
        cache_mode 'h0' keeps the layer-0 rows in each replica's cache, 'proj' does not."""
        layers = P["layers"]
        keep = {}

        def A(x, dt):
            a = np.ascontiguousarray(x, dt); keep[id(a)] = a; return a
        d = NNModelDesc()
        for k in range(3):
            d.d[k] = int(geo["d"][k])
        d.n_cells = int(geo["n_cells"])
        d.n_window = len(geo["window"]); d.window = A(geo["window"], np.int32).ctypes.data
        d.n_entries = len(geo["ent_col"]); d.entry_slot = A(geo["ent_slot"], np.uint16).ctypes.data
        d.entry_col = A(geo["ent_col"], np.int32).ctypes.data; d.entry_coef = A(geo["ent_coef"], np.float64).ctypes.data
        d.n_shells = len(geo["shell_count"]); d.shell_count = A(geo["shell_count"], np.int32).ctypes.data
        d.partners = A(geo["partners"], np.int32).ctypes.data if len(geo["partners"]) else None
        d.c0 = len(P["lin"]); d.mu = A(mu, np.float64).ctypes.data; d.sd = A(sd, np.float64).ctypes.data
        d.lin = A(P["lin"], np.float64).ctypes.data
        d.n_layers = len(layers)
        if layers:
            d.layer_cin = A([L["self"].shape[0] for L in layers], np.int32).ctypes.data
            d.layer_d = A([L["self"].shape[1] for L in layers], np.int32).ctypes.data
            d.w_self = A(np.concatenate([np.ravel(L["self"]) for L in layers]), np.float64).ctypes.data
            d.w_shell = A(np.concatenate([np.ravel(L["shell"]) for L in layers]), np.float64).ctypes.data
            d.b = A(np.concatenate([np.ravel(L["b"]) for L in layers]), np.float64).ctypes.data
            if "gate" in layers[0]:
                d.w_gate = A(np.concatenate([np.ravel(L["gate"]) for L in layers]), np.float64).ctypes.data
                d.gb = A(np.concatenate([np.ravel(L["gb"]) for L in layers]), np.float64).ctypes.data
            d.head_w = A(P["head"]["w"], np.float64).ctypes.data
        d.activation = {"gate": 0, "tanh": 0, "silu": 1}[act]
        d.head_b = float(P["head"]["b"]); d.readout = {"mean": 0, "sum": 1}[readout]; d.de_scale = float(de_scale)
        for l, cap in enumerate(geo["ball_cap"]):
            d.ball_cap[l] = int(cap)
        d.cache_mode = {"h0": 0, "proj": 1}[cache_mode]
        mid = C.c_int32()
        check(lib.nce_nn_model_create(self.h, C.byref(d), C.byref(mid)))
        ce, se = C.c_int64(), C.c_int64()
        check(lib.nce_nn_sizes(self.h, mid, C.byref(ce), C.byref(se)))
        return int(mid.value), int(ce.value), int(se.value)

    def nn_full_cache(self, model, occ, cache, energy, precision=PREC_FP64, stream=None):
        R, N = occ.shape
        check(lib.nce_nn_full_cache(self.h, model, ptr(occ), R, N, ptr(cache), ptr(energy), int(precision),
                                    stream_ptr(stream)))

    def nn_run_chunk(self, jobs, n_steps, precision=PREC_MIXED, stream=None):
        arr = (NNJob * len(jobs))(*jobs)
        check(lib.nce_nn_run_chunk(self.h, C.addressof(arr), len(jobs), int(n_steps), int(precision),
                                   stream_ptr(stream)))

    def nn_delta(self, jobs, dE, precision=PREC_MIXED, stream=None):
        """energy change of the injected proposal of every replica, nothing committed (dE: device float64)."""
        arr = (NNJob * len(jobs))(*jobs)
        check(lib.nce_nn_delta(self.h, C.addressof(arr), len(jobs), ptr(dE), int(precision), stream_ptr(stream)))


class Cycle:
    """A captured exchange cycle (nce_pt_cycle): the job tables live on the device, a launch replays the graph."""

    def __init__(self, ctx, jobs, n_steps, chunks_per_cycle, precision, tile, ex_jobs):
        arr = (PTJob * len(jobs))(*jobs)
        ex = (ExchangeJob * max(len(ex_jobs), 1))(*ex_jobs)
        h = _p()
        check(lib.nce_pt_cycle_create(ctx.h, C.addressof(arr), len(jobs), int(n_steps), int(chunks_per_cycle),
                                      int(precision), int(tile), C.addressof(ex), len(ex_jobs), C.byref(h)))
        self.h = h
        self.ctx = ctx                     # destroyed before the context

    def launch(self, n_cycles=1, stream=None):
        check(lib.nce_pt_cycle_launch(self.h, int(n_cycles), stream_ptr(stream)))

    def close(self):
        if getattr(self, "h", None):
            lib.nce_pt_cycle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
