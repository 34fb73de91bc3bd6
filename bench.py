#!/usr/bin/env python3
"""Benchmark of the MI355X Kern-Frenkel rototranslation hot path.

Workload (one GPU): plain NVT checkerboard Monte Carlo of N=32768 hard spheres
with four Kern-Frenkel patches (delta 0.12, cosmax 0.92) at rho=0.5, T=0.12,
Disp_max 0.05, Theta_max 0.1, FP64, from a seeded random-insertion start.
A step is one rebuild on a randomly shifted grid plus 8 colour passes in which
every non-empty cell makes --trials-per-cell trials (0: one per member, so a
step is exactly N attempts).  The metric counts sweeps of N attempted moves
(MC.c:21): value = attempts / N / second.

Prints one JSON line: sweeps/s (value), the dominant kernel's achieved
algorithmic bandwidth against the HBM roofline (HIP events on the kernel's
own stream, inside the timed region), decorrelated samples/s of E/N, and the
reference's own CPU code timed on this host (cpu_baseline).  Decorrelated
samples/s = sweeps/s / (2 tau_int[E/N]); every tau_int run (plain, plain with
one trial per cell, learned) starts from the state the timed region ends in
and runs the same number of sweeps, untimed for the plain one, and is given
with its statistical error, its length in units of tau and the drift of E/N
across it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, RHO, T, SEED = 32768, 0.5, 0.12, 1
DELTA, COSMAX, DISP, THETA = 0.12, 0.92, 0.05, 0.1
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_move(rho, side, delta=DELTA):
    """Bytes one rototranslation attempt must touch with SoA FP64 cells: own
    state read and written (2 x 96), 27 cell ranges (27 x 8), positions of the
    candidates (27 rho side^3 x 24), orientations of the neighbours within the
    bond range (rho 4pi/3 ((1+delta)^3 - 1) x 72)."""
    n_cand = 27.0 * rho * side ** 3
    n_rc = rho * 4.0 * math.pi / 3.0 * ((1 + delta) ** 3 - 1)
    return 2 * 96 + 27 * 8 + n_cand * 24 + n_rc * 72


def cpu_baseline(r, O, box, seconds):
    """The reference's own code on one host core: oracle/_ref/PatchyParticles_fast
    (built from the reference sources with its -O3 -ffast-math flags) on the
    same start; two runs of different length cancel the start-up cost.  Falls
    back to the oracle's restatement (oracle/cpu_bench) without the reference."""
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    taskset = ["taskset", "-c", "0"] if shutil.which("taskset") else []
    if os.path.exists(exe):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        d = tempfile.mkdtemp(prefix="ppcpu_")
        pyoracle.rrr_save(os.path.join(d, "start.rrr"), 0, box, r, O)

        def run(steps):
            pyoracle.write_input(os.path.join(d, f"in{steps}"), "start.rrr", T=T, seed=SEED, steps=steps,
                                 print_every=10 ** 9, save_every=10 ** 9, folder=d)
            t0 = time.perf_counter()
            subprocess.run(taskset + [exe, f"in{steps}"], cwd=d, check=True, capture_output=True)
            return time.perf_counter() - t0

        t_probe = run(5) - run(0)
        per = max(t_probe / 5, 1e-3)
        k = max(10, int(seconds / per))
        t_short, t_long = run(5), run(5 + k)
        rate = k / max(t_long - t_short, 1e-9)
        return dict(value=rate, unit="sweeps/s", cores=1, kind="reference",
                    sample=f"{k} sweeps of the N={N} rho={RHO} T={T} state the GPU run ended in (difference of a {5 + k}- and a "
                           f"5-sweep run of oracle/_ref/PatchyParticles_fast, -O3 -ffast-math, 1 core)")
    exe = os.path.join(ROOT, "oracle", "cpu_bench")
    out = subprocess.run(taskset + [exe, str(N), str(RHO), str(T), str(SEED), str(seconds)], check=True,
                         capture_output=True, text=True).stdout
    j = json.loads(out)
    return dict(value=j["sweeps_per_s"], unit="sweeps/s", cores=1, kind="port",
                sample=f"{j['sweeps']} sweeps of the N={N} rho={RHO} T={T} start with oracle/cpu_bench")


def host_cpu():
    """The host the CPU baseline ran on: model name and logical CPUs."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count()}


def measured_traffic(path=os.path.join(ROOT, "profiles", "k_pass_n32768_rho0.5_pmc.txt")):
    """HBM bytes per colour-pass launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy), or None."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def series_stats(e, steps_per_sweep):
    """tau_int[E/N] of a per-step E/N series in sweeps, with its statistical
    error, the run's length in sweeps and in units of tau, and the drift of
    E/N (second half's mean minus the first's, with the block standard error
    of that difference): a drift of several standard errors means the run is
    not stationary and tau is overestimated."""
    import numpy as np
    from patchyparticles_amd.stats import block_mean, tau_int_err
    tau_s, tau_e, _ = tau_int_err(e)
    tau = tau_s / steps_per_sweep
    h = len(e) // 2
    m1, se1 = block_mean(e[:h])
    m2, se2 = block_mean(e[h:])
    run = len(e) / steps_per_sweep
    return {"tau_int_sweeps": round(tau, 2), "tau_int_err_sweeps": round(tau_e / steps_per_sweep, 2),
            "run_sweeps": round(run), "run_over_tau": round(run / tau, 1), "E_per_N": round(float(e.mean()), 5),
            "E_drift": round(m2 - m1, 5), "E_drift_se": round(float(np.hypot(se1, se2)), 5)}


def tau_leg(pp, r_eq, O_eq, box_eq, opts, trials_per_cell, sweeps, attempts_per_step):
    """tau_int[E/N] of the plain sampler from the equilibrated state over
    `sweeps` sweeps (untimed; its rate is the timed region's)."""
    steps = max(1, int(math.ceil(sweeps * N / attempts_per_step)))
    with pp.System(r_eq, O_eq, box_eq, seed=SEED + 4, trials_per_cell=trials_per_cell, **opts) as s:
        s.do_ensemble(50)
        s.reset_counters()
        s.set_trace(steps)
        s.do_ensemble(steps)
        c = s.counters()
        e = s.trace() / N
    return series_stats(e, N * steps / c["tries"][0])


def learned_leg(args, pp, r_eq, O_eq, box_eq, opts):
    """The learned proposal against the plain move on the same schedule (one
    trial per non-empty cell and pass), from the equilibrated state: a
    ProposalNet trained on --train-steps snapshots of the plain sampler (one
    Adam step per snapshot, --train-batch moves each), then --learned-steps
    timed steps with --learned-fraction of the colour passes learned and the
    rest plain; the plain run times the same steps without the network.
    Decorrelated samples/s = sweeps/s / (2 tau_int[E/N])."""
    def timed(net):
        with pp.System(r_eq, O_eq, box_eq, seed=SEED + 2, trials_per_cell=1, **opts) as s:
            if net is not None:
                s.attach(net, learned_fraction=args.learned_fraction, theta_max=THETA)
            s.do_ensemble(50)
            s.reset_counters()
            s.set_timing(args.time_every)
            s.set_trace(args.learned_steps)
            t0 = time.perf_counter()
            s.do_ensemble(args.learned_steps)
            el = time.perf_counter() - t0
            c = s.counters()
            e = s.trace() / N
            nl, ml = s.timing("learned")
            tl, al = s.learn_counters(0)
            b, ov, tot, anyo = s.energy()
            assert not anyo and tot == c["bonds"], "tracked energy disagrees with a full recompute"
        sw = c["tries"][0] / N / el
        st = series_stats(e, N * args.learned_steps / c["tries"][0])
        out = {"sweeps_per_s": round(sw, 2), **st,
               "decorrelated_samples_per_s": round(sw / (2 * st["tau_int_sweeps"]), 4),
               "acceptance": round(c["accepted"][0] / max(c["tries"][0], 1), 5)}
        if net is not None:
            out["learned_acceptance"] = round(al / max(tl, 1), 5)
            out["learned_pass_us"] = round(ml / max(nl, 1) * 1e3, 3)
        return out

    print("# learned proposal: training ...", file=sys.stderr, flush=True)
    net = pp.ProposalNet(seed=SEED)
    with pp.System(r_eq, O_eq, box_eq, seed=SEED + 3, trials_per_cell=args.trials_per_cell, **opts) as s:
        t0 = time.perf_counter()
        l0, l1 = s.train(net, steps=args.train_steps, batch=args.train_batch, lr=1e-3, pos_noise=0.03,
                         theta_max=THETA, sweeps_between=1)
        t_train = time.perf_counter() - t0
    print(f"# trained in {t_train:.1f} s (loss {l0:.3f} -> {l1:.3f}); timing learned and plain ...", file=sys.stderr,
          flush=True)
    learned = timed(net)
    plain = timed(None)
    net.close()
    learned["train"] = {"steps": args.train_steps, "batch": args.train_batch, "seconds": round(t_train, 2),
                        "loss_first": round(l0, 4), "loss_last": round(l1, 4), "pos_noise": 0.03}
    learned["learned_fraction"] = args.learned_fraction
    return learned, plain


def _jumps(s, sweeps):
    """Mean squared displacement and mean 1 - cos(angle) per attempted move of
    replica 0, from states one sweep apart."""
    import numpy as np
    tr = rot = 0.0
    att = 0
    for _ in range(sweeps):
        r0, O0, box = s.state(0)
        c0 = s.counters()["tries"][0]
        s.do_ensemble(1)
        r1, O1, _ = s.state(0)
        d = r1 - r0
        d -= box * np.rint(d / box)
        tr += float((d * d).sum())
        rot += float((1 - (np.einsum("nij,nij->n", O1, O0) - 1) / 2).sum())
        att += s.counters()["tries"][0] - c0
    return tr / att, rot / att


def bench_learned(args, pp):
    """The learned proposal at the gel state point: N=65536, rho=0.3, T=0.09.
    Plain equilibration (--warmup steps of 8 trials per cell), a network
    trained on --train-steps snapshots of the plain sampler, then --steps
    timed steps each of the learned form (--learned-fraction of the passes
    learned, the rest plain; one trial per non-empty cell and pass) and of
    the plain move on the same schedule, both from the same state.  value =
    sweeps/s of the learned form; jumps per attempt and tau_int[E/N] for
    both, with tau's error, the run's length in units of tau and the drift of
    E/N across it."""
    n, rho, temp = 65536, 0.3, 0.09
    r, O, box = pp.generate(n, rho, SEED)
    opts = dict(T=temp, disp_max=DISP, theta_max=THETA, kf_delta=DELTA, kf_cosmax=COSMAX)
    with pp.System(r, O, box, seed=SEED, trials_per_cell=8, **opts) as s:
        s.do_ensemble(args.warmup)
        st = s.state(0)
        print(f"# warmup {args.warmup} steps, E/N={-s.counters()['bonds'] / n:.4f}", file=sys.stderr, flush=True)
        net = pp.ProposalNet(seed=SEED)
        t0 = time.perf_counter()
        l0, l1 = s.train(net, steps=args.train_steps, batch=args.train_batch, lr=1e-3, pos_noise=0.03,
                         theta_max=THETA, sweeps_between=1)
        t_train = time.perf_counter() - t0
    print(f"# trained in {t_train:.1f} s (loss {l0:.3f} -> {l1:.3f})", file=sys.stderr, flush=True)
    res = {}
    for name in ("learned", "plain"):
        with pp.System(*st, seed=SEED + 2, trials_per_cell=1, **opts) as s:
            if name == "learned":
                s.attach(net, learned_fraction=args.learned_fraction, theta_max=THETA)
            s.do_ensemble(20)
            jt, jr = _jumps(s, 10)
            s.reset_counters()
            s.set_timing(args.time_every)
            s.set_trace(args.steps)
            t0 = time.perf_counter()
            s.do_ensemble(args.steps)
            el = time.perf_counter() - t0
            c = s.counters()
            e = s.trace() / n
            nl, ml = s.timing("learned")
            b, ov, tot, anyo = s.energy()
            assert not anyo and tot == c["bonds"], "tracked energy disagrees with a full recompute"
        sw = c["tries"][0] / n / el
        st_e = series_stats(e, n * args.steps / c["tries"][0])
        res[name] = {"sweeps_per_s": round(sw, 2), "acceptance": round(c["accepted"][0] / c["tries"][0], 5),
                     "jump_sq_per_attempt": round(jt, 8), "one_minus_cos_per_attempt": round(jr, 8), **st_e,
                     "decorrelated_samples_per_s": round(sw / (2 * st_e["tau_int_sweeps"]), 4)}
        if name == "learned":
            res[name]["learned_pass_us"] = round(ml / max(nl, 1) * 1e3, 3)
        print(f"# {name}: {res[name]}", file=sys.stderr, flush=True)
    net.close()
    out = {
        "metric": "particle sweeps/s + decorrelated samples/s (E/N), learned vs plain, 1 GPU",
        "value": res["learned"]["sweeps_per_s"], "unit": "sweeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 geometry, bf16/f32 network", "data": "synthetic (seeded random-insertion start)",
        "config": {"workload": f"learned-proposal NVT, N={n}, rho={rho}, T={temp}, learned fraction "
                               f"{args.learned_fraction}, one trial per non-empty cell and pass",
                   "train": {"steps": args.train_steps, "batch": args.train_batch, "seconds": round(t_train, 2),
                             "loss_first": round(l0, 4), "loss_last": round(l1, 4)}},
        "decorrelated_samples_per_s": {"learned": res["learned"]["decorrelated_samples_per_s"],
                                       "plain": res["plain"]["decorrelated_samples_per_s"]},
        "learned": res["learned"], "plain": res["plain"],
    }
    print(json.dumps(out), flush=True)


def bench_npt(args, pp):
    """16 NPT replicas of N=131072 at P=0.02, T=0.11 from a dilute start
    (rho=0.05), rescale_factor_max 0.01, one volume attempt per N particle
    attempts; with --learned-fraction > 0 a network trained on the batch
    proposes that share of the colour passes.  value = sum over replicas of
    attempts / N per second."""
    import numpy as np
    n, rho, temp, P, R = 131072, 0.05, 0.11, 0.02, 16
    rs, Os, bs = [], [], []
    for k in range(R):
        r, O, box = pp.generate(n, rho, SEED + k)
        rs.append(r)
        Os.append(O)
        bs.append(box)
    prm = [dict(T=temp, disp_max=DISP, theta_max=THETA, kf_delta=DELTA, kf_cosmax=COSMAX, ensemble=pp.NPT, P=P,
                rescale_factor_max=0.01)] * R
    s = pp.System.batch(rs, Os, bs, prm, seed=SEED, trials_per_cell=args.trials_per_cell)
    print(f"# {R} replicas of N={n} created", file=sys.stderr, flush=True)
    for k in range(0, args.warmup, 50):
        s.do_ensemble(min(50, args.warmup - k))
        print(f"# warmup {k + 50} steps, density {float((n / s.counters_all()['volume']).mean()):.4f}",
              file=sys.stderr, flush=True)
    net = None
    if args.learned_fraction > 0:
        net = pp.ProposalNet(seed=SEED)
        steps = min(args.train_steps, 1000)  # an NPT sweep of the batch between steps is the slow part
        for k in range(0, steps, 100):
            l0, l1 = s.train(net, steps=min(100, steps - k), batch=args.train_batch, lr=1e-3, pos_noise=0.03,
                             theta_max=THETA, sweeps_between=1)
            print(f"# train {k + 100} steps, loss {l1:.3f}", file=sys.stderr, flush=True)
        s.attach(net, learned_fraction=args.learned_fraction, theta_max=THETA)
    s.reset_counters()
    t0 = time.perf_counter()
    for k in range(0, args.steps, 50):
        s.do_ensemble(min(50, args.steps - k))
        print(f"# timed {min(k + 50, args.steps)} steps", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    c = s.counters_all()
    s.close()
    if net is not None:
        net.close()
    rep_sweeps = float((c["tries"][:, 0] / n).sum())
    dens = n / c["volume"]
    out = {
        "metric": "particle sweeps/s + decorrelated samples/s (E/N), learned vs plain, 1 GPU",
        "value": round(rep_sweeps / el, 2), "unit": "replica sweeps/s (sum over replicas of attempts / N per second)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random-insertion starts)",
        "config": {"workload": f"{R} NPT replicas of N={n}, P={P}, T={temp}, start rho={rho}, rescale_factor_max 0.01",
                   "trials_per_cell": args.trials_per_cell, "learned_fraction": args.learned_fraction,
                   "density_mean": round(float(dens.mean()), 5), "E_per_N_mean": round(float((-c["bonds"] / n).mean()), 5),
                   "volume_acceptance": round(float(c["accepted"][:, 1].sum() / max(c["tries"][:, 1].sum(), 1)), 4)},
    }
    print(json.dumps(out), flush=True)


def bench_batch(args, pp):
    """256 independent NVT replicas in one context: N log-uniform in [500,
    8000], T uniform in [0.08, 0.20], rho uniform in [0.05, 0.60], random
    sequential insertion starts; the launch shape is autotuned on real
    sweeps.  value = sum over replicas of attempts / N_r per second."""
    import numpy as np
    rs, Os, bs, ps = pp.uneven_replicas(256, seed=SEED)
    n = np.array([len(r) for r in rs])
    sysm = pp.System.batch(rs, Os, bs, ps, seed=SEED, trials_per_cell=args.trials_per_cell)
    print(f"# batch: {len(rs)} replicas, {n.sum()} particles", file=sys.stderr, flush=True)
    tpc, block, us = sysm.autotune(10)
    print(f"# autotuned: {tpc} lanes per cell, {block} threads per workgroup ({us:.1f} us/step)", file=sys.stderr,
          flush=True)
    sysm.do_ensemble(args.warmup)
    sysm.reset_counters()
    t0 = time.perf_counter()
    sysm.do_ensemble(args.steps)
    t1 = time.perf_counter()
    c = sysm.counters_all()
    sysm.close()
    elapsed = t1 - t0
    rep_sweeps = float((c["tries"][:, 0] / n).sum())
    out = {
        "metric": "particle sweeps/s + decorrelated samples/s (E/N), learned vs plain, 1 GPU",
        "value": round(rep_sweeps / elapsed, 2),
        "unit": "replica sweeps/s (sum over replicas of attempts / N_r per second)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded random-insertion starts)",
        "config": {"workload": "256 independent NVT replicas, N log-uniform in [500, 8000], T U[0.08, 0.20], "
                               "rho U[0.05, 0.60], KF_delta 0.12, KF_cosmax 0.92, FP64",
                   "particles": int(n.sum()), "trials_per_cell": args.trials_per_cell,
                   "autotuned": {"threads_per_cell": tpc, "block_size": block},
                   "attempts_per_s": round(float(c["tries"][:, 0].sum()) / elapsed, 1),
                   "acceptance": round(float(c["accepted"][:, 0].sum() / c["tries"][:, 0].sum()), 5)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000, help="timed steps (a step: rebuild + 8 colour passes)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (equilibration; default 40000 for the single workload, else 6000)")
    ap.add_argument("--tau-sweeps", type=int, default=24000,
                    help="sweeps of the plain tau_int run when the learned leg is off (0: no tau_int run)")
    ap.add_argument("--trials-per-cell", type=int, default=16,
                    help="trials per non-empty cell and pass (0: one per member, a step of exactly N attempts)")
    ap.add_argument("--per-member-steps", type=int, default=2000,
                    help="steps of the one-trial-per-member schedule timed after the main run (0: skip)")
    ap.add_argument("--tpc", type=int, default=0, help="threads per cell (0: default)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--occupancy", type=float, default=0.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no HIP events in the timed region")
    ap.add_argument("--time-every", type=int, default=16,
                    help="HIP events bracket the kernels of every k-th sweep of the timed region")
    ap.add_argument("--learned-steps", type=int, default=30000,
                    help="timed steps of the learned-vs-plain comparison (0: skip it)")
    ap.add_argument("--train-steps", type=int, default=10000, help="Adam steps (one sampler snapshot each)")
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--learned-fraction", type=float, default=0.8)
    ap.add_argument("--workload", choices=("single", "batch", "learned", "npt"), default="single",
                    help="single: the N=32768 system; batch: 256 independent replicas of uneven N, T and rho; "
                         "learned: the learned proposal at N=65536, rho=0.3, T=0.09; npt: 16 NPT replicas of "
                         "N=131072")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (the reference point that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profile)")
    args = ap.parse_args()
    if args.warmup is None:
        args.warmup = 40000 if args.workload == "single" else 6000
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "independent replicas would shard whole systems across GPUs.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import patchyparticles_amd as pp

    if args.workload == "batch":
        return bench_batch(args, pp)
    if args.workload == "learned":
        return bench_learned(args, pp)
    if args.workload == "npt":
        return bench_npt(args, pp)

    r, O, box = pp.generate(N, RHO, SEED)
    opts = dict(T=T, disp_max=DISP, theta_max=THETA, kf_delta=DELTA, kf_cosmax=COSMAX, threads_per_cell=args.tpc,
                block_size=args.block, cell_occupancy=args.occupancy)
    sysm = pp.System(r, O, box, seed=SEED, trials_per_cell=args.trials_per_cell, **opts)
    print(f"# start: N={N} box={box[0]:.4f} grid={sysm.grid()}", file=sys.stderr, flush=True)
    sysm.do_ensemble(args.warmup)
    print(f"# warmup {args.warmup} steps done, E/N={-sysm.counters()['bonds'] / N:.4f}", file=sys.stderr, flush=True)

    # timed region: K steps; HIP events bracket the kernels of every
    # --time-every-th step (the per-kernel averages of the roofline line)
    sysm.reset_counters()
    sysm.set_timing(0 if args.no_timing else args.time_every)
    t0 = time.perf_counter()
    sysm.do_ensemble(args.steps)  # synchronous: returns when the device is done
    t1 = time.perf_counter()
    elapsed = t1 - t0
    n_pass, ms_pass = sysm.timing("pass")
    n_reb, ms_reb = sysm.timing("rebuild")
    c = sysm.counters()
    sysm.set_timing(0)
    print(f"# timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    attempts = c["tries"][0]
    sweeps_per_s = attempts / N / elapsed  # a sweep is N attempted moves (MC.c:21)
    nside, side = sysm.grid()
    bpm = algorithmic_bytes_per_move(RHO, side[0])
    moves_per_launch = attempts / args.steps / 8  # attempted moves one colour pass makes
    bytes_per_launch = bpm * moves_per_launch
    avg_pass_s = ms_pass / max(n_pass, 1) / 1e3
    achieved = bytes_per_launch / avg_pass_s / 1e9 if avg_pass_s > 0 else 0.0

    # parity spot check of the final snapshot (energy recomputed vs tracked)
    b, ov, tot, anyo = sysm.energy()
    assert not anyo and tot == c["bonds"], "tracked energy disagrees with a full recompute"
    r_eq, O_eq, box_eq = sysm.state()
    sysm.close()

    per_member = None
    if args.per_member_steps > 0 and args.trials_per_cell != 0:
        # the same state advanced with one trial per member of each cell
        with pp.System(r_eq, O_eq, box_eq, seed=SEED + 1, trials_per_cell=0, **opts) as s0:
            s0.do_ensemble(50)
            s0.reset_counters()
            ta = time.perf_counter()
            s0.do_ensemble(args.per_member_steps)
            tb = time.perf_counter()
            c0 = s0.counters()
        per_member = {"trials_per_cell": 0, "sweeps_per_s": round(c0["tries"][0] / N / (tb - ta), 2),
                      "steps": args.per_member_steps}

    # decorrelated samples: every tau run starts from this state and runs as
    # many sweeps as the one-trial-per-cell plain run (or --tau-sweeps)
    learned = plain_same = None
    tau_sweeps = args.tau_sweeps
    if args.learned_steps > 0:
        learned, plain_same = learned_leg(args, pp, r_eq, O_eq, box_eq, opts)
        tau_sweeps = plain_same["run_sweeps"]
    plain_tau = decorrelated = None
    if tau_sweeps > 0:
        print(f"# tau_int of the plain run over {tau_sweeps} sweeps ...", file=sys.stderr, flush=True)
        plain_tau = tau_leg(pp, r_eq, O_eq, box_eq, opts, args.trials_per_cell, tau_sweeps, attempts / args.steps)
        decorrelated = sweeps_per_s / (2 * plain_tau["tau_int_sweeps"])

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(r_eq, O_eq, box_eq, args.cpu_seconds)

    out = {
        "metric": "particle sweeps/s + decorrelated samples/s (E/N), learned vs plain, 1 GPU",
        "value": round(sweeps_per_s, 2),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-insertion start, uniform random orientations)",
        "config": {
            "workload": f"plain NVT checkerboard rototranslation, N={N}, rho={RHO}, T={T}, KF_delta={DELTA}, "
                        f"KF_cosmax={COSMAX}, Disp_max={DISP}, Theta_max={THETA}, FP64",
            "grid": nside, "cell_side": round(side[0], 5),
            "threads_per_cell": args.tpc or 32, "trials_per_cell": args.trials_per_cell,
            "step": "rebuild on a randomly shifted grid + 8 colour passes",
            "attempts": attempts, "attempts_per_step": round(attempts / args.steps, 1),
            "acceptance": round(c["accepted"][0] / max(attempts, 1), 5),
            "E_per_N": round(float(-c["bonds"] / N), 6),
        },
        "decorrelated_samples_per_s": {
            "plain": round(decorrelated, 3) if decorrelated else None, "plain_run": plain_tau,
            "learned": learned["decorrelated_samples_per_s"] if learned else None,
            "learned_over_plain": round(learned["decorrelated_samples_per_s"] / decorrelated, 4)
            if learned and decorrelated else None,
            "learned_run": learned,
            "plain_one_trial_per_cell": plain_same,
        },
        "one_trial_per_member": per_member,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(),
            "kernel": "k_pass (colour pass)", "bytes_per_move": round(bpm, 1),
            "moves_per_launch": round(moves_per_launch, 1),
            "algorithmic_bytes_per_launch": round(bytes_per_launch), "avg_launch_us": round(avg_pass_s * 1e6, 3),
            "rebuild_us_per_sweep": round(ms_reb / max(n_reb, 1) * 1e3, 3),
            "timed_sweeps": n_reb,
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        cpu.update(host_cpu())
        out["speedup_vs_cpu"] = round(sweeps_per_s / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
