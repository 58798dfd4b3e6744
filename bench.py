#!/usr/bin/env python3
"""Headline benchmark: MC swap attempts/s summed over all replicas, parallel-tempering CE swap Monte Carlo.

Workload (default, the 1,048,576-site configuration): fcc 64 x 64 x 64 conventional cells = 1,048,576 sites, Ni/Pt
with Pt fraction x = 0.25 at random from a seed, cluster space pairs <= 8 A / triplets <= 6 A / quadruplets <= 5 A,
synthetic ECIs N(0, (20 meV / r^2)^2), 256 replicas on a geometric ladder 300-1400 K, int8 occupations, FP32 swap
energy change with FP64 running totals / exchange / accumulators ("mixed"), exchange every 5000 steps.

One bench step = one chunk of --chunk MC steps of every replica (one nce_pt_run_chunk launch) followed by the
temperature-swap exchange (one nce_pt_exchange launch).  Inputs are resident in HBM before the timed region.

The line also carries
  roofline      the chunk kernel's algorithmic HBM bytes (2 covering windows of 141 int8 sites read + 2 sites written
                = 284 B per attempt) over its average launch duration (HIP events on the launch stream), against
                MI355X's 8 TB/s;
  cpu_baseline  the FP64 C oracle (same proposal / swap energy change / Metropolis) timed on this host's cores on a
                bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_ALG_PER_ATTEMPT = 284           # bytes: two 141-site int8 windows + two written sites (8/6/5 fcc space)
KB = 8.617333e-5
PUBLISHED_H100 = 4_605_830.0      # reference: CE attempts/s at N = 1,048,576, R = 256 on H100 (pure step loop)
# HBM bytes per swap attempt moved by the chunk kernel, from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per
# dispatch (factors from a 1 GiB device copy), for the workload and precision they were measured on, with the
# profile summary that holds the counts (the counters cannot run inside the timed region)
MEASURED_TRAFFIC_PER_ATTEMPT = {("ce_1m", "mixed"): 3461.0, ("nce_ptni_16k", "mixed"): 82060.0,
                                ("nce_16k", "bf16"): 2.28e6}
TRAFFIC_PROFILE = {("ce_1m", "mixed"): "profiles/ce_chunk_spec_kernel_fcc64cubic_1m_r256.md",
                   ("nce_ptni_16k", "mixed"): "profiles/nn_chunk_kernel_ptni_fcc16cubic_r32.md",
                   ("nce_16k", "bf16"): "profiles/nn_chunk_kernel_silu3x64_fcc16cubic_r64_bf16.md"}

CONFIGS = {
    "ce_1m": dict(n=64, x=0.25, R=256, Tmin=300.0, Tmax=1400.0, exchange=5000, precision="mixed"),
    "ce_2k": dict(n=8, x=0.5, R=32, Tmin=300.0, Tmax=1400.0, exchange=5000, precision="fp64"),
    "ce_131k": dict(n=32, x=0.25, R=256, Tmin=300.0, Tmax=1400.0, exchange=5000, precision="mixed"),
    "ce_10m": dict(n=136, x=0.25, R=256, Tmin=300.0, Tmax=1400.0, exchange=5000, precision="mixed"),
    # NeuralCE: 3 SiLU graph-conv layers of width 64 over the 12 nearest neighbours on 40 cluster channels
    "nce_16k": dict(n=16, x=0.5, R=64, Tmin=300.0, Tmax=1400.0, exchange=5000, precision="bf16", nce="silu"),
    # NeuralCE: the reference's shipped PtNi model (depth 1, width 16, 3 shells) at 16,384 sites, 32 replicas
    "nce_ptni_16k": dict(n=16, x=0.5, R=32, Tmin=700.0, Tmax=1100.0, exchange=5000, precision="mixed", nce="ptni"),
    "nce_ptni_1m": dict(n=64, x=0.5, R=32, Tmin=700.0, Tmax=1100.0, exchange=5000, precision="mixed", nce="ptni"),
    "nce_ptni_10m": dict(n=136, x=0.5, R=8, Tmin=700.0, Tmax=1100.0, exchange=5000, precision="mixed", nce="ptni"),
}
# reference NeuralCE local-update attempts/s on H100 at the same N and R (results/bench_stage/gpu_conv_*.json)
PUBLISHED_H100_NCE = {"nce_ptni_16k": 159_250.0, "nce_ptni_1m": 84_585.0, "nce_ptni_10m": 14_615.0}


def nce_bench(args, cfg):
    """NeuralCE local-update chunk kernel; roofline on MFMA-eligible FLOPs of the convolution layers."""
    import torch
    from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    from neuralce2_amd.nce.mcmc import NNJobState
    at = fcc_cell(cfg["n"], cfg["x"], args.seed)
    if cfg["nce"] == "silu":
        cs = fcc_cluster_space((9.0, 6.0, 5.0))
        P = silu_network(len(cs), 64, 3, 1, seed=args.seed)
        mu, sd = np.zeros(len(cs)), np.ones(len(cs))
        nn = NNModel(cs, at, P, mu, sd, [0.0, 3.3], act="silu", readout="sum")
    else:
        from neuralce2_amd.nce.model_io import load_model, expand_params
        from neuralce2_amd.clusterspace import ClusterSpace
        mdl = load_model(os.path.join(ROOT, "tests", "golden", "nipt_v3_nomax_sel.pkl"))
        cs = ClusterSpace(mdl["primitive"], mdl["cutoffs"], mdl["chemical_symbols"])
        P, mu, sd = expand_params(mdl, 0)
        nn = NNModel(cs, at, P, mu, sd, mdl["shell_edges"], ys=mdl["ys"], ym=mdl["ym"])
    R = cfg["R"]
    temps = np.geomspace(cfg["Tmin"], cfg["Tmax"], R)
    betas = (1.0 / (8.617333262e-5 * temps)).astype(np.float32).astype(np.float64)
    prec = {"fp64": 0, "mixed": 1, "bf16": 2}[cfg["precision"]]
    sig = nn.geo["sig"].astype(np.int8)
    st = NNJobState(nn, sig, betas, seed=args.seed, precision=prec)
    chunk = args.chunk
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        st.run(chunk)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for c in range(args.steps):
        evs[c][0].record(stream)
        st.run(chunk)
        evs[c][1].record(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    k_ms = np.array([a.elapsed_time(b) for a, b in evs])
    value = R * chunk * args.steps / dt
    # algorithmic FLOPs per attempt of the graph convolutions (the incremental local update): per hidden layer l the
    # change of layer l-1 on ball_{l-1} (union of the two flips' balls, from the geometry) projected through the
    # concatenated [self | shell | gate] maps, 2 FLOP per multiply-add; a swap of two equal species (probability
    # 1 - 2x(1-x) for a random pair) does no work
    geo = nn.geo
    caps = [min(2 * sz, nn.N) for sz in geo["ball_sizes"]]
    n_sh = len(geo["shell_count"])
    fl = 0.0
    for l, Lr in enumerate(P["layers"]):
        cin, d = Lr["self"].shape
        fl += 2.0 * caps[l] * cin * d * (1 + n_sh + (n_sh if "gate" in Lr else 0))
    fl *= 2.0 * cfg["x"] * (1.0 - cfg["x"])
    achieved = fl * R * chunk / (k_ms.mean() * 1e-3) / 1e12
    peak = 2500.0 if prec == 2 else (157.3 if prec == 1 else 78.6)   # dense bf16 MFMA / fp32 / fp64
    # running totals against a full recompute of the caches after the timed chunks (drift of the local update)
    E_run = st.energy.clone()
    st.anchor()
    drift = float((st.energy - E_run).abs().max().item()) * (nn.ys if cfg["nce"] == "ptni" else 1.0)
    acc = float(st.n_accept.sum().item()) / (R * chunk * (args.warmup + args.steps))
    return dict(value=value, dt=dt, k_ms=k_ms, achieved=achieved, peak=peak, acc=acc, R=R, N=nn.N,
                flops_attempt=fl, prec=prec, drift=drift)


def served_bench(args):
    """Served request stream: Poisson arrivals of C_v(T) requests, continuous batching on one GPU, against the plain
    FP64 one-job-at-a-time form on the same requests.  The cell models of the stream's sizes are built before the
    clock starts (a server keeps the models of the shapes it serves); both forms use them."""
    from neuralce2_amd.serve import Server, job_stream, run_sequential
    specs = job_stream(args.jobs, args.rate, seed=args.seed, n_min=4, n_max=args.n_max, t_min=16,
                       t_max=args.t_max, n_steps=args.job_steps, chunk=args.chunk)
    srv = Server(chunk=args.chunk, precision=args.precision or "mixed", verify=False)
    srv.prepare(specs)
    srv.run(specs[:2])                                   # warm-up: code objects, tile cache
    res, st = srv.run(specs)
    base = None
    if not args.no_cpu:                                  # the plain form is the comparison; skipped with --no-cpu
        _, base = run_sequential(specs, precision="fp64", chunk=args.chunk, models_from=srv, verify=False)
    return res, st, base, specs


def batched_bench(args):
    """Batched uneven jobs: every request present at t = 0, fcc n^3 cubic cells (n in 4..32, log-uniform in sites),
    x in 0.1..0.9, 16-64 replicas, CE and NeuralCE jobs 3:1 (the NeuralCE jobs on the 3-layer SiLU network), all
    run together by the continuous-batching server: one chunk launch for every CE job, one for every NeuralCE job."""
    from neuralce2_amd.serve import Server, job_stream
    from neuralce2_amd.synthetic import fcc_cluster_space, fcc_cell
    from neuralce2_amd.nce.mc_model import NNModel, silu_network
    cs = fcc_cluster_space((9.0, 6.0, 5.0))
    P = silu_network(len(cs), 64, 3, 1, seed=args.seed)
    mu, sd = np.zeros(len(cs)), np.ones(len(cs))

    def nce_model(n):
        return NNModel(cs, fcc_cell(n, 0.5, 0), P, mu, sd, [0.0, 3.3], act="silu", readout="sum")

    specs = job_stream(args.jobs, 0.0, seed=args.seed, n_min=4, n_max=args.n_max, t_min=16, t_max=64,
                       n_steps=args.job_steps, chunk=args.chunk, nce_every=4)
    srv = Server(chunk=args.chunk, precision=args.precision or "mixed", nn_precision="bf16", nce_model=nce_model)
    srv.run([s for s in specs if s.kind == "ce"][:1] + [s for s in specs if s.kind == "nce"][:1])   # warm-up
    res, st = srv.run(specs)
    return res, st, specs


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed chunks")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chunk", type=int, default=5000, help="MC steps per replica per chunk")
    ap.add_argument("--config", default="ce_1m", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed", "bf16"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU time of the cpu_baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--served", action="store_true", help="served request stream (p50/p99 job latency)")
    ap.add_argument("--batched", action="store_true",
                    help="uneven CE + NeuralCE jobs (3:1), all present at once (use --jobs 512 --job-steps N)")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--rate", type=float, default=100.0, help="requests per second (Poisson)")
    ap.add_argument("--job-steps", type=int, default=100000)
    ap.add_argument("--n-max", type=int, default=32, help="largest cell: fcc n^3 cubic cells (136: 10,061,824 sites)")
    ap.add_argument("--t-max", type=int, default=64, help="most temperatures of a request")
    ap.add_argument("--dump-moments", default=None,
                    help="write the per-temperature <E> and C_v of the timed run (npz) for precision comparisons")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 1 GiB (the known-size kernel that scales rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes)")
    return ap.parse_args()


def calibrate_copy():
    import torch
    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 3}), flush=True)


# the reference's published icet/mchammer CPU sampler at the benchmark's cell size (a different host):
# results/bench_stage/mchammer.json, n_atoms 1,048,576 (steps_per_s_manual; one process, 900 K, 64^3 cubic cells)
PUBLISHED_MCHAMMER = {"value": 24.503007275504032, "unit": "swap steps/s (one chain)", "n_sites": 1048576,
                      "host": "ccc0472.campuscluster.illinois.edu (the reference's CPU node, not this box)",
                      "source": "results/bench_stage/mchammer.json"}


def cpu_baseline(info, occ0, betas, target_s, seed):
    """The FP64 oracle on this host's cores on a bounded sample (all replicas, n steps sized to ~target_s), plus a
    single-thread mchammer-style chain (partner always of the other species) on one replica of the same cell, and the
    reference's published mchammer figure for this cell size.  Threads: the cores this process may use, at most 16:
    the box allots 16 CPU cores to a one-GPU job, whatever the machine has."""
    from oracle.ce_oracle import OracleCE
    o = OracleCE(info)
    cores = min(len(os.sched_getaffinity(0)), 16)
    R = occ0.shape[0]
    occ = np.ascontiguousarray(occ0.copy())
    E0 = o.energy(occ[0])
    E = np.array([E0] * R)
    t = time.perf_counter()
    o.run_chunk(occ, E, betas, 4, seed=seed, step0=10**9, n_threads=cores)
    per_step = max((time.perf_counter() - t) / 4, 1e-6)
    n = int(max(4, min(100000, target_s / per_step)))
    t = time.perf_counter()
    o.run_chunk(occ, E, betas, n, seed=seed, step0=2 * 10**9, n_threads=cores)
    dt = time.perf_counter() - t
    one = np.ascontiguousarray(occ0[0].copy())
    n1 = 20000
    t1 = time.perf_counter()
    o.mchammer_swaps(one, E0, float(betas[0]), n1, seed=seed)
    dt1 = time.perf_counter() - t1
    n1 = int(max(n1, min(2_000_000, n1 * 0.25 * target_s / max(dt1, 1e-9))))
    t1 = time.perf_counter()
    o.mchammer_swaps(one, E0, float(betas[0]), n1, seed=seed + 1)
    dt1 = time.perf_counter() - t1
    return dict(value=R * n / dt, unit="MC swap attempts/s", cores=cores, kind="port",
                sample=f"{R} replicas x {n} steps of the same cell and ladder, FP64 oracle, OpenMP over replicas on "
                       f"{cores} threads ({dt:.1f} s)",
                single_thread_mchammer_style={"value": n1 / dt1, "unit": "swap steps/s (one chain, one thread)",
                                              "sample": f"{n1} swaps of one replica, partner of the other species, "
                                                        f"FP64 local energy change ({dt1:.1f} s)"},
                published_mchammer=PUBLISHED_MCHAMMER)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is measured per GPU", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batched:
        res, st, specs = batched_bench(args)
        n_ce = sum(1 for s in specs if s.kind == "ce")
        line = {"metric": "MC swap attempts/s (all replicas, batched uneven jobs)", "value": st["attempts_per_s"],
                "unit": "attempts/s", "n_gpus": 1, "steps": st["n_ticks"], "warmup": 1,
                "ms_per_step": st["wall_s"] / max(st["n_ticks"], 1) * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32 dE / bf16 mfma NeuralCE / f64 totals",
                "data": "synthetic batch (seeded)",
                "config": {"workload": f"{len(specs)} PT jobs at once ({n_ce} CE, {len(specs) - n_ce} NeuralCE 3-layer "
                                       f"SiLU), fcc n^3 cubic cells log-uniform 256-{4 * args.n_max ** 3} sites, "
                                       f"x in 0.1..0.9, 16-64 replicas, {specs[0].n_steps} steps each, chunk "
                                       f"{args.chunk}"},
                "latency_p50_s": st["latency_p50_s"], "latency_p99_s": st["latency_p99_s"], "wall_s": st["wall_s"],
                "attempts": st["attempts"], "roofline": None, "cpu_baseline": None}
        for kind in ("ce", "nce"):
            lat = np.array([r.latency for r in res if r.kind == kind])
            if len(lat):
                line[f"latency_{kind}_p50_s"] = float(np.percentile(lat, 50))
                line[f"latency_{kind}_p99_s"] = float(np.percentile(lat, 99))
        print(json.dumps(line), flush=True)
        return
    if args.served:
        res, st, base, specs = served_bench(args)
        line = {"metric": "MC swap attempts/s (all replicas, served stream); p50/p99 PT-job latency",
                "value": st["attempts_per_s"], "unit": "attempts/s", "n_gpus": 1, "steps": st["n_ticks"],
                "warmup": 1, "ms_per_step": st["wall_s"] / max(st["n_ticks"], 1) * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32 dE / f64 totals",
                "data": "synthetic request stream (seeded)",
                "config": {"workload": f"{args.jobs} C_v(T) requests, Poisson {args.rate}/s, fcc n^3 cubic cells "
                                       f"log-uniform 256-{4 * args.n_max ** 3} sites, x in 0.1..0.9, 16-{args.t_max} "
                                       f"temperatures, {args.job_steps} steps each, chunk {args.chunk}"},
                "latency_p50_s": st["latency_p50_s"], "latency_p99_s": st["latency_p99_s"],
                "wall_s": st["wall_s"], "attempts": st["attempts"],
                "plain_fp64_sequential": base,
                "speedup_vs_plain_fp64_sequential": (st["attempts_per_s"] / base["attempts_per_s"]) if base else None,
                "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return
    import torch
    from neuralce2_amd.synthetic import synthetic_ce, fcc_cell
    from neuralce2_amd.ce.classical_ce_gpu import GPUClusterExpansion
    from neuralce2_amd.ce.pt_mcmc_ce import PTJobState
    from neuralce2_amd._lib import PREC_FP64, PREC_MIXED
    cfg = dict(CONFIGS[args.config])
    if args.precision:
        cfg["precision"] = args.precision
    if "nce" in cfg:
        res = nce_bench(args, cfg)
        line = {"metric": "MC swap attempts/s (all replicas)", "value": res["value"], "unit": "attempts/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["dt"] / args.steps * 1e3,
                "higher_is_better": True, "scaling": "weak",
                "vs_baseline": (res["value"] / PUBLISHED_H100_NCE[args.config]
                                if args.config in PUBLISHED_H100_NCE else None),
                "dtype": ["f64", "f32", "bf16 mfma / f32"][res["prec"]],
                "data": ("synthetic (seeded random occupations; the shipped PtNi model's weights)" if cfg["nce"] == "ptni"
                         else "synthetic (seeded random occupations and weights)"),
                "config": {"workload": f"NeuralCE PT swap MC ({cfg['nce']}), fcc {cfg['n']}^3 cubic = {res['N']} "
                                       f"sites, {res['R']} replicas", "chunk_steps": args.chunk},
                "roofline": {"bound": "mfma", "achieved": res["achieved"],
                             "peak": res["peak"],
                             "traffic": (MEASURED_TRAFFIC_PER_ATTEMPT[(args.config, cfg["precision"])] * res["R"]
                                         * args.chunk
                                         if (args.config, cfg["precision"]) in MEASURED_TRAFFIC_PER_ATTEMPT else None),
                             "traffic_from": TRAFFIC_PROFILE.get((args.config, cfg["precision"])),
                             "unit": "TFLOP/s", "frac": res["achieved"] / res["peak"],
                             "kernel": "nn_chunk_kernel", "kernel_ms_avg": float(res["k_ms"].mean()),
                             "flops_per_attempt": res["flops_attempt"]},
                "cpu_baseline": None, "acceptance": res["acc"],
                "drift_vs_full_recompute": res["drift"]}
        print(json.dumps(line), flush=True)
        return
    prec = PREC_MIXED if cfg["precision"] == "mixed" else PREC_FP64
    ce = synthetic_ce(args.seed)
    at = fcc_cell(cfg["n"], cfg["x"], args.seed)
    g = GPUClusterExpansion(ce, at, energy_scale=float(len(at)))
    R = cfg["R"]
    temps = np.geomspace(cfg["Tmin"], cfg["Tmax"], R).astype(np.float32)
    betas = (1.0 / (KB * temps.astype(np.float64))).astype(np.float32).astype(np.float64)
    sig = g.encode(at).astype(np.int8)
    chunk = args.chunk
    st = PTJobState(g, sig, betas, seed=args.seed, sample_interval=chunk, samples_per_chunk=1)
    stream = torch.cuda.current_stream()
    ex_every_chunks = max(1, cfg["exchange"] // chunk)

    def one(c, ev=None):
        if ev is not None:
            ev[0].record(stream)
        g.ctx.pt_run_chunk([st.job(accumulate=True)], chunk, prec, args.tile)
        if ev is not None:
            ev[1].record(stream)
        if (c + 1) % ex_every_chunks == 0:
            g.ctx.pt_exchange([st.exchange_job(c + 1, reorder_last=True)])

    for c in range(args.warmup):
        one(c)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(args.steps):
        one(args.warmup + c, evs[c])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    k_ms = np.array([a.elapsed_time(b) for a, b in evs])
    attempts = R * chunk * args.steps
    value = attempts / dt
    # outside the timed region: running totals against an FP64 full recompute, and the per-slot moments
    full = torch.empty_like(st.energy)
    g.ctx.ce_energy(g.model, st.occ, full)
    drift = float((full - st.energy).abs().max().item()) / len(at)
    if args.dump_moments:
        mean, var = st.moments()
        T64 = temps.astype(np.float64)
        np.savez(args.dump_moments, temps=T64, E_mean=mean / len(at), C_v=var / (KB * T64) ** 2 / len(at),
                 steps=chunk * (args.warmup + args.steps), precision=cfg["precision"], drift=drift)
    bytes_launch = R * chunk * B_ALG_PER_ATTEMPT
    achieved = bytes_launch / (k_ms.mean() * 1e-3) / 1e9
    line = {
        "metric": "MC swap attempts/s (all replicas)",
        "value": value,
        "unit": "attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_H100 if args.config == "ce_1m" else None,
        "dtype": "f32 dE / f64 totals" if prec == PREC_MIXED else "f64",
        "data": "synthetic (seeded random occupations and ECIs)",
        "config": {"workload": f"CE PT swap MC, fcc {cfg['n']}^3 cubic = {len(at)} sites, x={cfg['x']}, "
                               f"{R} replicas, T {cfg['Tmin']:.0f}-{cfg['Tmax']:.0f} K, cutoffs 8/6/5 A, "
                               f"exchange every {cfg['exchange']} steps",
                   "sites": len(at), "replicas": R, "chunk_steps": chunk, "precision": cfg["precision"],
                   "tile": args.tile},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (MEASURED_TRAFFIC_PER_ATTEMPT[(args.config, cfg["precision"])] * R * chunk
                                 if (args.config, cfg["precision"]) in MEASURED_TRAFFIC_PER_ATTEMPT else None),
                     "traffic_from": TRAFFIC_PROFILE.get((args.config, cfg["precision"])),
                     "kernel": "ce_chunk_spec_kernel", "kernel_ms_avg": float(k_ms.mean()),
                     "bytes_per_attempt_alg": B_ALG_PER_ATTEMPT},
        "cpu_baseline": None,
        "acceptance": float(st.n_accept.sum().item()) / (R * chunk * (args.warmup + args.steps)),
        "drift_vs_full_recompute": {"ev_per_site": drift, "steps_per_replica": chunk * (args.warmup + args.steps)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g.info, np.broadcast_to(sig, (R, len(sig))), betas,
                                            args.cpu_seconds, args.seed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
