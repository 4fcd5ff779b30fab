#!/usr/bin/env python3
"""Chain-sweeps/s of the projector-DQMC sweep on one MI355X.

Workload (default): 14x14 EPSOCZ lattice (Ns = 392, ne = 196), Lt = 1000,
dtau = 0.1, stabilization every 10 slices, coupling 0.64, omega = 1, fields
N(0,1) from seed 7, one chain, plain FP64.  One step = one chain-sweep: one
iteration of the reference's sweep loop (all slices x 4 bond families x all
bonds, wraps, QR stabilization every 10 slices, one measurement at Lt/2).
Directions alternate starting backward, as in the reference.

`batched` reports the batched form: --batch chains of the same workload
advanced together (pqmc_sweep_batch), their aggregate chain-sweeps/s and the
largest |G - G_plain| of the batch's first chain, which replays the plain
chain's fields and draws.

`mixed_batch` reports 128 independent chains of uneven lattice and Lt (32
each at L = 8, 10, 12, 14; Lt cycling through 400..1000; seeds 1000 + i)
advanced by pqmc_sweep_many, 16 at a time.

Prints one JSON line.  `roofline` is that of the dominant kernel class,
timed with HIP events on the chain's own stream over the timed region;
`cpu_baseline` runs the oracle (a single-threaded C restatement of the
reference algorithm) as the reference deploys itself: one chain per host
core on all cores of this process's share at once, each over one full sweep
of the workload; `value` there is the mean per-core rate.
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

PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix)
PEAK_HBM = 8.0e12    # MI355X HBM3E

CONFIGS = {
    # name: (L, Lt, ngroup, ne, seed)
    "c2": (14, 1000, 10, 196, 7),
    "c1": (4, 100, 10, 16, 1234),
}


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def sweep_flops(Ns, ne, Nb, Lt, ng, n_acc, d=4):
    """F_sweep of the survey's accounting: rank-4 updates + wraps + propagation + stabilization."""
    return (8 * d * Ns ** 2 * n_acc + Lt * Nb * (16 * d * d * Ns + 8 * d * d * ne)
            + (Lt / ng) * (32 * Ns * ne ** 2 + (16 / 3) * ne ** 3 + 8 * Ns ** 2 * ne))


def _cpu_chain(spec):
    """One oracle chain (a process of its own): init_g, then one full backward
    sweep with the measurement at Lt/2, timed without the init."""
    L, Lt, ng, ne, seed = spec
    t_start = time.time()
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    orc = O.Oracle(L, L, Lt, ng, ne)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    u = O.Ran2Stream(seed).draws(2 * m.Nb * Lt)
    t_init = time.time()
    t0 = time.perf_counter()
    orc.sweep(False, u, measure=True)
    return time.perf_counter() - t0, t_start, t_init


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(L, Lt, ng, ne, seed, cores):
    """The oracle as the reference deploys itself, one independent chain per
    core on `cores` host cores at once (one MPI rank per core), each timed
    over one full sweep of the workload; seeds seed + i."""
    import multiprocessing as mp
    # one thread per chain, BLAS included, as the reference's runs set
    # OPENBLAS_NUM_THREADS / MKL_NUM_THREADS = 1 per rank (submit.py:607-609);
    # the workers inherit this environment when they start
    saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in saved:
        os.environ[k] = "1"
    t0 = time.time()
    with mp.get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_chain, [(L, Lt, ng, ne, seed + i) for i in range(cores)])
    wall = time.time() - t0
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    dts = [r[0] for r in res]
    rates = [1.0 / d for d in dts]
    return {"value": float(np.mean(rates)), "unit": "chain-sweeps/s", "cores": cores, "kind": "port",
            "aggregate": float(np.sum(rates)), "cpu_model": _cpu_model(),
            "sweep_s": {"min": float(min(dts)), "max": float(max(dts))},
            "process_start_s": {"min": min(r[1] for r in res) - t0, "max": max(r[1] for r in res) - t0},
            "init_s": {"min": min(r[2] - r[1] for r in res), "max": max(r[2] - r[1] for r in res)},
            "sample": f"oracle (C, 1 thread per process), {cores} chains (seeds {seed}..{seed + cores - 1}) on "
                      f"{cores} cores at once, one full backward sweep each with the Lt/2 measurement "
                      f"(init not timed; {wall:.1f} s wall); value = mean per-core rate"}


def default_cpu_cores():
    """Host cores of this process's share: the affinity set, capped by the
    thread count the environment grants (OMP_NUM_THREADS) and by 16."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        n = min(n, int(omp))
    return max(1, min(n, 16))


def mixed_batch_leg(nchains, threads, helper_bonds, sweeps):
    """The mixed batch: nchains independent chains (32 each at L = 8, 10, 12,
    14; Lt cycling through 400..1000; seeds 1000 + i) advanced by
    pqmc_sweep_many, `threads` at a time.  Reports aggregate chain-sweeps/s
    and Lt=1000-equivalent sweeps/s (sum of Lt / 1000 per second)."""
    from pqmc_allinone_amd.host import Chain, Ran2, sweep_many
    from pqmc_allinone_amd.model import ModelParams, build_model, mixed_chain_specs, synthetic_inputs
    specs = mixed_chain_specs(nchains)
    models, chains, rngs = {}, [], []
    t0 = time.perf_counter()
    for L, Lt, seed in specs:
        if (L, Lt) not in models:
            models[(L, Lt)] = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=10, ne=L * L))
        m = models[(L, Lt)]
        x, sl = synthetic_inputs(m, seed)
        ch = Chain(m, helper_bonds=helper_bonds)
        ch.set_fields(x)
        ch.set_trial(sl)
        ch.init_g(1)
        chains.append(ch)
        rngs.append(Ran2(seed))
    t_init = time.perf_counter() - t0
    fw = True
    dt = 0.0
    for _ in range(sweeps):
        fw = not fw
        draws = [r.draws(2 * c.Nb * c.model.p.Lt) for r, c in zip(rngs, chains)]
        t0 = time.perf_counter()
        sweep_many(chains, [fw] * nchains, draws, measure=True, max_concurrent=threads)
        dt += time.perf_counter() - t0
    drift = max(c.counters()["err_fast_max"] for c in chains)
    for c in chains:
        c.close()
    lt_sum = sum(Lt for _, Lt, _ in specs)
    return {"chains": nchains, "threads": threads, "helper_bonds": helper_bonds, "sweeps": sweeps,
            "value": nchains * sweeps / dt, "unit": "chain-sweeps/s",
            "lt1000_sweeps_per_s": lt_sum / 1000.0 * sweeps / dt, "s_per_sweep": dt / sweeps,
            "init_s": t_init, "err_fast_max": drift, "dtype": "c128 (fp64)",
            "workload": "32 chains each at L = 8, 10, 12, 14 (Ns = 2 L^2, ne = L^2), Lt cycling 400/600/800/1000, "
                        "ngroup 10, seeds 1000 + i"}


def precision_leg(L, Lt, seed, modes=("fp32x3", "bf16x3")):
    """The mixed-precision B-string products against FP64 on one chain of
    L x L, Lt slices, fields and draws from `seed`: max|G - G_FP64| after
    init_g(1) and after one backward sweep with the same draws, how many of
    the sweep's accept decisions agree, and the largest relative difference
    of the Lt/2 measurement (entries above 1e-6 of the largest)."""
    from pqmc_allinone_amd.host import Chain, Ran2
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=10, ne=L * L))
    x, sl = synthetic_inputs(m, seed)
    u = Ran2(seed).draws(2 * m.Nb * Lt)

    def run(mode):
        ch = Chain(m)
        ch.set_precision(mode)
        ch.set_fields(x)
        ch.set_trial(sl)
        t0 = time.perf_counter()
        ch.init_g(1)
        g_init = ch.g_h
        t1 = time.perf_counter()
        meas = ch.sweep(False, u, measure=True)
        ch.synchronize()
        t2 = time.perf_counter()
        out = (g_init, ch.g_h, ch.accepts(), meas, t1 - t0, t2 - t1, ch.drift()[0])
        ch.close()
        return out

    ref = run("fp64")
    res = {"lattice": f"{L}x{L}", "Lt": Lt, "seed": seed, "fp64_sweep_s": ref[5]}
    big = np.abs(ref[3]) > 1e-6 * np.abs(ref[3]).max()
    for mode in modes:
        r = run(mode)
        same = r[2] == ref[2]
        res[mode] = {"g_err_init": float(np.abs(r[0] - ref[0]).max()),
                     "g_err_after_sweep": float(np.abs(r[1] - ref[1]).max()),
                     "accepts_agree": float(same.mean()),
                     "first_accept_difference": None if same.all() else int(np.argmin(same)),
                     "measurement_max_rel_diff": float((np.abs(r[3] - ref[3])[big] / np.abs(ref[3])[big]).max()),
                     "drift_g": r[6], "sweep_s": r[5]}
    return res


def batched_leg(m, x0, sl, seed, K, warmup, steps, block_bonds, helper_bonds, g_plain, precision="fp32x3"):
    """K chains of the workload, one batch sweep per step; chain 0 replays the
    plain chain (same fields, trial and draws), chains i > 0 take fields and
    draws from seed + i."""
    from pqmc_allinone_amd.host import Chain, Ran2, sweep_batch
    from pqmc_allinone_amd.model import synthetic_inputs
    chains, rngs = [], []
    for i in range(K):
        xi = x0 if i == 0 else synthetic_inputs(m, seed + i)[0]
        ch = Chain(m, block_bonds=block_bonds, helper_bonds=helper_bonds)
        ch.set_precision(precision)
        ch.set_fields(xi)
        ch.set_trial(sl)
        ch.init_g(1)
        chains.append(ch)
        rngs.append(Ran2(seed + i))
    nd = 2 * m.Nb * m.p.Lt
    fw = True
    for _ in range(warmup):
        fw = not fw
        sweep_batch(chains, [fw] * K, [r.draws(nd) for r in rngs], measure=True)
    for ch in chains:
        ch.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fw = not fw
        sweep_batch(chains, [fw] * K, [r.draws(nd) for r in rngs], measure=True)
    for ch in chains:
        ch.synchronize()
    dt = time.perf_counter() - t0
    g_err = float(np.abs(chains[0].g_h - g_plain).max())
    for ch in chains:
        ch.close()
    return {"chains": K, "helper_bonds": helper_bonds, "value": K * steps / dt, "unit": "chain-sweeps/s",
            "ms_per_step": 1e3 * dt / steps, "precision": precision,
            "dtype": "c128 (fp64); B-string products between checkpoints on FP32 MFMA with three-term split "
                     "operands" if precision == "fp32x3" else "c128 (fp64)",
            "max_abs_g_err_vs_plain": g_err}


def ft_leg(m, x, seed, warmup, steps):
    """The finite-temperature Green's function (proj = .false.) on the same
    lattice, fields and draws: one chain, sweeps with a measurement at Lt/2;
    G = (1 + B)^-1 from Ns x Ns strings, so each stabilization is heavier."""
    from pqmc_allinone_amd.host import Chain, Ran2
    ch = Chain(m, finite_temperature=True)
    ch.set_fields(x)
    t0 = time.perf_counter()
    ch.init_g(1)
    t_init = time.perf_counter() - t0
    rng = Ran2(seed)
    nd = 2 * m.Nb * m.p.Lt
    fw = True
    for _ in range(warmup):
        fw = not fw
        ch.sweep(fw, rng.draws(nd), measure=True)
    for k in range(3):
        ch.timing(k, True)
    ch.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fw = not fw
        ch.sweep(fw, rng.draws(nd), measure=True)
    ch.synchronize()
    dt = time.perf_counter() - t0
    kern = {}
    for k, name in enumerate(("family_pass", "wrap", "stabilization")):
        ms, n = ch.timing(k, False)
        kern[name] = {"ms_total": ms, "launches": n, "us_avg": 1e3 * ms / max(n, 1)}
    out = {"value": steps / dt, "unit": "chain-sweeps/s", "ms_per_step": 1e3 * dt / steps,
           "init_ms": 1e3 * t_init, "dtype": "c128 (fp64)", "kernels": kern,
           "ln_cw_re": ch.ln_cw.real}
    ch.close()
    return out


def grad_leg(m, x, sl, passes, nchains):
    """The field gradient d ln_cw / dx (pqmc_gradient) on the same lattice and
    fields: one chain alone, then nchains chains at once (one host thread
    each, their own streams), each pass from a fresh init_g(1), which is not
    timed.  A hybrid-Monte-Carlo force evaluation is one init_g plus one
    pass."""
    from concurrent.futures import ThreadPoolExecutor
    from pqmc_allinone_amd.host import Chain
    nf = m.Nb * m.p.Lt

    def chain(k):
        ch = Chain(m)
        ch.set_fields(x if k == 0 else x + 1e-3 * k)
        ch.set_trial(sl)
        return ch

    ch = chain(0)
    t_init, t_pass = [], []
    for _ in range(passes):
        t0 = time.perf_counter()
        ch.init_g(1)
        ch.synchronize()
        t1 = time.perf_counter()
        ch.gradient()
        t_pass.append(time.perf_counter() - t1)
        t_init.append(t1 - t0)
    ch.close()
    out = {"fields": nf, "ms_per_pass": 1e3 * min(t_pass), "init_ms": 1e3 * min(t_init),
           "fields_per_s": nf / min(t_pass), "dtype": "c128 (fp64)"}
    if nchains > 1:
        chs = [chain(k) for k in range(nchains)]
        with ThreadPoolExecutor(nchains) as pool:
            list(pool.map(lambda c: c.init_g(1), chs))
            for c in chs:
                c.synchronize()
            t0 = time.perf_counter()
            list(pool.map(lambda c: c.gradient(), chs))
            dt = time.perf_counter() - t0
        for c in chs:
            c.close()
        out["chains"] = {"n": nchains, "ms": 1e3 * dt, "fields_per_s": nchains * nf / dt}
    return out


def hmc_leg(m, x, sl, ntraj):
    """Hybrid-Monte-Carlo moves on one chain of the same workload: leapfrog
    trajectories of 10 steps of 0.05 with the Fourier-accelerated mass, each
    force one init_g(1) plus one gradient pass, the endpoint weighed by the
    exact full-determinant weight."""
    from pqmc_allinone_amd.host import Chain, HybridMC
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    hmc = HybridMC(ch, step=0.05, nsteps=10, seed=1)
    dts, dhs = [], []
    for _ in range(ntraj):
        t0 = time.perf_counter()
        hmc.update()
        dts.append(time.perf_counter() - t0)
        dhs.append(hmc.last_dH)
    ch.close()
    return {"trajectories": ntraj, "s_per_trajectory": float(np.mean(dts)), "dH": dhs,
            "accepted": hmc.accepted, "leapfrog": "10 x 0.05, Fourier-accelerated mass"}


def kernel_classes(reps):
    """The path's kernel classes other than the family pass, on the 14x14,
    Lt=1000 chain, for counter collection: init_g(1), then `reps` each of
    get_g_scratch (B-string propagation, blocked Householder QR, G from
    scratch), take_measurement and the standalone two-way wrap, then
    `reps` stabilizations with the string products in each reduced
    precision (FP32 split and BF16x3)."""
    from pqmc_allinone_amd.host import Chain
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs
    L, Lt, ng, ne, seed = CONFIGS["c2"]
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    for r in range(reps):
        ch.get_g_scratch(Lt - 10 * (r + 1) + 1, False)
        ch.take_measurement(Lt // 2)
        ch.right_evolve(Lt - r, 4)
    for mode in ("fp32x3", "bf16x3"):
        ch.set_precision(mode)
        for r in range(reps):
            ch.get_g_scratch(Lt - 10 * (r + 1) + 1, False)
    ch.set_precision("fp64")
    ch.synchronize()
    print(json.dumps({"kernel_classes": reps}))


def pass_phases():
    """Phase timers of the family pass (pqmc_update_profile) at 14x14,
    Lt=1000: one warm-up sweep, then one forward and one backward sweep with
    the timers on; per role the wall-clock microseconds per pass in each of
    the kernel's eight timer slots (their meaning per role is in
    k_pass.hip's PROF_ACC marks)."""
    import ctypes as C
    from pqmc_allinone_amd import _lib
    from pqmc_allinone_amd.host import Chain, Ran2
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs
    L, Lt, ng, ne, seed = CONFIGS["c2"]
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne, lam=0.64, omega=1.0))
    x, sl = synthetic_inputs(m, seed)
    ch = Chain(m)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    rng = Ran2(seed)
    nd = 2 * m.Nb * Lt
    ch.sweep(False, rng.draws(nd), measure=True)
    lib = _lib.lib()
    P = C.c_int32(0)
    out = {}
    for fwd in (True, False):
        _lib.check(lib.pqmc_update_profile(ch.h, 1, None, C.addressof(P)))
        ch.timing(0, True)
        ch.sweep(fwd, rng.draws(nd), measure=True)
        ch.synchronize()
        ms, n = ch.timing(0, False)
        G = P.value
        prof = np.zeros(8 * G)
        _lib.check(lib.pqmc_update_profile(ch.h, 0, prof.ctypes.data, C.addressof(P)))
        prof = prof.reshape(G, 8) / n
        H = G - 5
        rows = {"leader wave 0": prof[0], "helpers (mean)": prof[1:H + 1].mean(0), "leader wave 1": prof[H + 1],
                "leader wave 2": prof[H + 2], "leader wave 6": prof[H + 3], "leader wave 7": prof[H + 4]}
        out["forward" if fwd else "backward"] = {"us_per_pass": 1e3 * ms / n,
                                                 **{k: [round(float(v), 2) for v in r] for k, r in rows.items()}}
        log(f"pass phases, {'forward' if fwd else 'backward'} sweep: {1e3 * ms / n:.1f} us per pass")
    ch.close()
    print(json.dumps({"pass_phases": out}))


def copy_calibration(mib):
    """Elementwise copy kernel over `mib` MiB of f64: FETCH_SIZE and WRITE_SIZE of
    one dispatch should each read mib * 1024 KiB; the ratio scales the counters
    of the path's kernels to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.randn(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_kib": mib * 1024, "dispatches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--block-bonds", type=int, default=0)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="host cores (one oracle chain each) of the CPU baseline; 0 = this process's share, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--batch", type=int, default=9,
                    help="chains of the batched leg (0 or 1: no batched leg)")
    ap.add_argument("--batch-precision", default="fp32x3", choices=("fp64", "fp32x3", "bf16x3"),
                    help="B-string products of the batched leg's chains")
    ap.add_argument("--precision-lattices", default="14,20",
                    help="lattices of the precision leg (Lt=1000, seed 11; empty: no such leg)")
    ap.add_argument("--batch-helper-bonds", type=int, default=4,
                    help="bonds per column-owning workgroup in the batched leg's family passes (1, 2 or 4)")
    ap.add_argument("--ft-steps", type=int, default=2,
                    help="timed sweeps of the finite-temperature leg (0: no such leg)")
    ap.add_argument("--grad-passes", type=int, default=2,
                    help="field-gradient passes on one chain (0: skip the gradient leg)")
    ap.add_argument("--grad-chains", type=int, default=16, help="chains of the concurrent gradient pass")
    ap.add_argument("--hmc-trajectories", type=int, default=1, help="HMC trajectories (0: skip the HMC leg)")
    ap.add_argument("--mixed-chains", type=int, default=128,
                    help="chains of the mixed batch leg (32 per lattice of L = 8..14; 0: no such leg)")
    ap.add_argument("--mixed-threads", type=int, default=16, help="chains of the mixed batch swept at once")
    ap.add_argument("--mixed-sweeps", type=int, default=2, help="timed sweeps of the mixed batch")
    ap.add_argument("--kernel-classes", type=int, default=0, metavar="REPS",
                    help="only run REPS of each kernel class other than the family pass (counter collection)")
    ap.add_argument("--pass-phases", action="store_true",
                    help="only print the family pass's phase timers at 14x14, Lt=1000 (one sweep each way)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB) five times: "
                         "the known-size reference for the HBM byte counters")
    args = ap.parse_args()
    # one hardware queue per chain of the batched and gradient legs (read when
    # the HIP runtime starts, so before the first device call)
    nq = max(args.batch if args.batch > 1 else 0, args.grad_chains if args.grad_passes > 0 else 0,
             args.mixed_threads if args.mixed_chains > 0 else 0)
    if nq > 1:
        q = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        if q < nq + 2:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, nq + 3))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.kernel_classes:
        kernel_classes(args.kernel_classes)
        return
    if args.pass_phases:
        pass_phases()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); chains shard across GPUs with no "
              "exchange during sweeps, so multi-GPU throughput is N x the one-GPU value", file=sys.stderr)
        sys.exit(2)

    from pqmc_allinone_amd.host import Chain, Ran2
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs

    L, Lt, ng, ne, seed = CONFIGS[args.config]
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne, lam=0.64, omega=1.0))
    x, sl = synthetic_inputs(m, seed)
    log(f"plain chain {L}x{L} Lt={Lt}: init_g")
    ch = Chain(m, block_bonds=args.block_bonds)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
    log(f"plain chain: {args.warmup} warm-up and {args.steps} timed sweeps")
    rng = Ran2(seed)
    nd = 2 * m.Nb * Lt
    forward = True
    for _ in range(args.warmup):
        forward = not forward
        ch.sweep(forward, rng.draws(nd), measure=True)
    c0 = ch.counters()
    timing = not args.no_timing
    if timing:
        for k in range(3):
            ch.timing(k, True)
    ch.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        forward = not forward
        ch.sweep(forward, rng.draws(nd), measure=True)
    ch.synchronize()
    dt = time.perf_counter() - t0
    kern = {}
    if timing:
        for k, name in enumerate(("family_pass", "wrap", "stabilization")):
            ms, n = ch.timing(k, False)
            kern[name] = {"ms_total": ms, "launches": n, "us_avg": 1e3 * ms / max(n, 1)}
    c1 = ch.counters()
    n_acc = (c1["positive_accept"] - c0["positive_accept"]) / args.steps
    n_prop = n_acc + (c1["negative_accept"] - c0["negative_accept"]) / args.steps
    sps = args.steps / dt
    Ns, Nb, npla = m.Ns, m.Nb, m.npla
    F = sweep_flops(Ns, ne, Nb, Lt, ng, n_acc)
    roof = None
    if kern:
        dom = max(kern, key=lambda k: kern[k]["ms_total"])
        per = kern[dom]
        if dom == "family_pass":
            # one launch = one family at one slice: the rank-4 update of G for
            # every accepted bond (8 d Ns^2 flops each, d = 4) plus the fused
            # two-way wrap (2 * 8 d^2 Ns flops per bond)
            passes = per["launches"] / args.steps
            fl = (n_acc / passes) * 8 * 4 * Ns ** 2 + 2 * 8 * 16 * Ns * npla
        elif dom == "stabilization":
            fl = 32 * Ns * ne ** 2 + (16 / 3) * ne ** 3 + 8 * Ns ** 2 * ne
        else:
            fl = 2 * 8 * 16 * Ns * npla
        achieved = fl / (per["us_avg"] * 1e-6) / 1e12
        roof = {"kernel": dom, "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / (PEAK_FP64 / 1e12), "traffic": None,
                "flops_per_launch": fl}
    log(f"plain chain: {sps:.3f} chain-sweeps/s")
    batched = None
    if args.batch > 1:
        log(f"batched leg: {args.batch} chains")
        batched = batched_leg(m, x, sl, seed, args.batch, args.warmup, args.steps, args.block_bonds,
                              args.batch_helper_bonds, ch.g_h, args.batch_precision)
        batched["vs_plain"] = batched["value"] / sps
    ft = None
    if args.ft_steps > 0:
        log("finite-temperature leg")
        ft = ft_leg(m, x, seed, args.warmup, args.ft_steps)
    grad = None
    if args.grad_passes > 0:
        log("gradient leg")
        grad = grad_leg(m, x, sl, args.grad_passes, args.grad_chains)
    hmc = None
    if args.hmc_trajectories > 0:
        log("HMC leg")
        hmc = hmc_leg(m, x, sl, args.hmc_trajectories)
    prec = None
    if args.precision_lattices:
        prec = []
        for Lp in (int(v) for v in args.precision_lattices.split(",")):
            log(f"precision leg: {Lp}x{Lp} Lt=1000 seed 11")
            prec.append(precision_leg(Lp, 1000, 11))
    mixed = None
    if args.mixed_chains > 0:
        log(f"mixed batch leg: {args.mixed_chains} chains")
        mixed = mixed_batch_leg(args.mixed_chains, args.mixed_threads, 4, args.mixed_sweeps)
        log(f"mixed batch: {mixed['value']:.2f} chain-sweeps/s")
    cpu = None
    if not args.no_cpu:
        nc = args.cpu_cores or default_cpu_cores()
        log(f"CPU baseline: {nc} oracle chains, one full sweep each")
        cpu = cpu_baseline(L, Lt, ng, ne, seed, nc)
        log(f"CPU baseline: {cpu['value']:.4f} chain-sweeps/s per core")
    out = {
        "metric": "chain-sweeps/sec at 14x14 Lt=1000: plain FP64 vs batched mixed-prec @ G err<=1e-8",
        "value_is": "plain FP64, one chain (the batched mixed-precision figure is batched.value)",
        "value": sps,
        "unit": "chain-sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic: fields N(0,1) seed %d, trial = lowest ne eigenvectors of H0 + 1e-3 random hopping" % seed,
        "config": {"workload": f"EPSOCZ {L}x{L} Lt={Lt} dtau=0.1 ngroup={ng} lambda=0.64 omega=1 Ns={Ns} "
                               f"ne={ne} 1 chain, plain FP64, measure at Lt/2", "chains": 1},
        "roofline": roof,
        "sweep_roofline": {"flops_per_sweep": F, "achieved_tflops": F * sps / 1e12,
                           "frac_fp64": F * sps / PEAK_FP64},
        "acceptance": n_acc / max(n_prop, 1),
        "kernels": kern,
        "batched": batched,
        "mixed_batch": mixed,
        "precision": prec,
        "finite_temperature": ft,
        "gradient": grad,
        "hmc": hmc,
        "cpu_baseline": cpu,
        "speedup_vs_cpu_core": (sps / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
