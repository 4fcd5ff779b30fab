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

Prints one JSON line.  `roofline` is that of the dominant kernel class,
timed with HIP events on the chain's own stream over the timed region;
`cpu_baseline` times the oracle (a single-threaded C restatement of the
reference algorithm) on a bounded slice of the same workload.
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


def sweep_flops(Ns, ne, Nb, Lt, ng, n_acc, d=4):
    """F_sweep of the survey's accounting: rank-4 updates + wraps + propagation + stabilization."""
    return (8 * d * Ns ** 2 * n_acc + Lt * Nb * (16 * d * d * Ns + 8 * d * d * ne)
            + (Lt / ng) * (32 * Ns * ne ** 2 + (16 / 3) * ne ** 3 + 8 * Ns ** 2 * ne))


def cpu_baseline(L, Lt, ng, ne, seed, target_s):
    """The oracle on the first slices of a backward sweep, scaled to chain-sweeps/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from pqmc_allinone_amd.model import ModelParams, build_model, synthetic_inputs
    m = build_model(ModelParams(L1=L, L2=L, Lt=Lt, ngroup=ng, ne=ne))
    x, sl = synthetic_inputs(m, seed)
    orc = O.Oracle(L, L, Lt, ng, ne)
    orc.set_fields(x)
    orc.set_trial(sl)
    orc.init(1)
    rng = O.Ran2Stream(seed)
    u = rng.draws(2 * m.Nb * Lt)
    # calibrate on one stabilization group, then run a bounded sample
    t0 = time.perf_counter()
    orc.sweep(False, u, ntc=ng)
    t_group = time.perf_counter() - t0
    ngroups = max(1, min(Lt // ng - 1, int(target_s / max(t_group, 1e-9))))
    orc2 = O.Oracle(L, L, Lt, ng, ne)
    orc2.set_fields(x)
    orc2.set_trial(sl)
    orc2.init(1)
    t0 = time.perf_counter()
    orc2.sweep(False, u, ntc=ngroups * ng)
    dt = time.perf_counter() - t0
    frac = ngroups * ng / Lt
    return {"value": frac / dt, "unit": "chain-sweeps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (C, 1 thread) over the first {ngroups * ng} of {Lt} slices of a backward sweep "
                      f"({dt:.1f} s), scaled by slices"}


def batched_leg(m, x0, sl, seed, K, warmup, steps, block_bonds, helper_bonds, g_plain):
    """K chains of the workload, one batch sweep per step; chain 0 replays the
    plain chain (same fields, trial and draws), chains i > 0 take fields and
    draws from seed + i."""
    from pqmc_allinone_amd.host import Chain, Ran2, sweep_batch
    from pqmc_allinone_amd.model import synthetic_inputs
    chains, rngs = [], []
    for i in range(K):
        xi = x0 if i == 0 else synthetic_inputs(m, seed + i)[0]
        ch = Chain(m, block_bonds=block_bonds, helper_bonds=helper_bonds)
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
            "ms_per_step": 1e3 * dt / steps, "dtype": "c128 (fp64)", "max_abs_g_err_vs_plain": g_err}


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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true")
    ap.add_argument("--batch", type=int, default=9,
                    help="chains of the batched leg (0 or 1: no batched leg)")
    ap.add_argument("--batch-helper-bonds", type=int, default=4,
                    help="bonds per column-owning workgroup in the batched leg's family passes (1, 2 or 4)")
    ap.add_argument("--ft-steps", type=int, default=2,
                    help="timed sweeps of the finite-temperature leg (0: no such leg)")
    ap.add_argument("--grad-passes", type=int, default=2,
                    help="field-gradient passes on one chain (0: skip the gradient leg)")
    ap.add_argument("--grad-chains", type=int, default=16, help="chains of the concurrent gradient pass")
    ap.add_argument("--hmc-trajectories", type=int, default=1, help="HMC trajectories (0: skip the HMC leg)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read MIB, write MIB) five times: "
                         "the known-size reference for the HBM byte counters")
    args = ap.parse_args()
    # one hardware queue per chain of the batched and gradient legs (read when
    # the HIP runtime starts, so before the first device call)
    nq = max(args.batch if args.batch > 1 else 0, args.grad_chains if args.grad_passes > 0 else 0)
    if nq > 1:
        q = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        if q < nq + 2:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, nq + 3))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
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
    ch = Chain(m, block_bonds=args.block_bonds)
    ch.set_fields(x)
    ch.set_trial(sl)
    ch.init_g(1)
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
    batched = None
    if args.batch > 1:
        batched = batched_leg(m, x, sl, seed, args.batch, args.warmup, args.steps, args.block_bonds,
                              args.batch_helper_bonds, ch.g_h)
        batched["vs_plain"] = batched["value"] / sps
    ft = None
    if args.ft_steps > 0:
        ft = ft_leg(m, x, seed, args.warmup, args.ft_steps)
    grad = None
    if args.grad_passes > 0:
        grad = grad_leg(m, x, sl, args.grad_passes, args.grad_chains)
    hmc = None
    if args.hmc_trajectories > 0:
        hmc = hmc_leg(m, x, sl, args.hmc_trajectories)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(L, Lt, ng, ne, seed, args.cpu_seconds)
    out = {
        "metric": "chain-sweeps/sec at 14x14 Lt=1000 (plain FP64 single chain)",
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
        "finite_temperature": ft,
        "gradient": grad,
        "hmc": hmc,
        "cpu_baseline": cpu,
        "speedup_vs_cpu_core": (sps / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
