#!/usr/bin/env python3
"""Benchmark of the MI355X Kern-Frenkel rototranslation hot path.

Workload (one GPU): plain NVT checkerboard Monte Carlo of N=32768 hard spheres
with four Kern-Frenkel patches (delta 0.12, cosmax 0.92) at rho=0.5, T=0.12,
Disp_max 0.05, Theta_max 0.1, FP64, from a seeded random-insertion start.
A step is one rebuild on a randomly shifted grid plus 8 colour passes in which
every non-empty cell makes --trials-per-cell trials (0: one per member, so a
step is exactly N attempts).  The metric counts sweeps of N attempted moves
(MC.c:21): value = attempts / N / second.

The system is equilibrated (--equil steps, 40000 by default, about 5e5
sweeps) before the warmup and the timed steps, whatever --warmup is, so the
headline is timed at the state point.  HIP events on the pass's own stream
bracket the kernels of every step of a short timed run (256 evenly spaced
steps of a long one) for the roofline line.

Prints one JSON line: sweeps/s (value), the dominant kernel's launch time and
bytes against the HBM roofline (bound "latency": each cell's trials are one
dependent chain), decorrelated samples/s of E/N, and the reference's own CPU
code timed on this host (cpu_baseline).  Decorrelated samples/s = sweeps/s /
(2 tau_int[E/N]), with tau_int pooled over 8 independent copies of an
equilibrated state in one batched context; a run whose tau has a relative
error above 0.25 or whose E/N drifts by 2 standard errors or more across it
reports null and "non-stationary" instead.  Three such measurements: the
headline form at its own state point; the learned proposal against the plain
move at a stationary state point (rho=0.5, --stat-T, --stat-n particles per
copy); and the same at the gel state point (N=65536, rho=0.3, T=0.09).
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
CLOCK_GHZ = 2.4  # MI355X peak engine clock


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


def decorrelation(pp, start, prm, *, seed, trials_per_cell, sweeps, replicas=8, net=None, learned_fraction=0.0,
                  label=""):
    """tau_int[E/N] of one sampler form from `replicas` independent copies of
    the state `start` = (r, O, box) in one batched context, each copy with its
    own random streams: the E/N series of every copy after every step
    (pp_set_trace), pooled by stats.tau_int_replicas; the drift of E/N across
    the run from the spread over the copies (stats.replica_drift).  The rate
    is the batch's own: sweeps of one copy per second (attempts / N / copies
    / s).  decorrelated_samples_per_s = that rate / (2 tau) is given only when
    the run passes the stationarity rule (tau's relative error below 0.25 and
    |drift| below 2 standard errors); otherwise it is null and `status` says
    "non-stationary" (the drift) or "tau error above 0.25" (a run too short
    for its tau)."""
    import numpy as np
    from patchyparticles_amd.stats import replica_drift, stationary, tau_int_replicas
    r, O, box = start
    n = len(r)
    with pp.System.batch([r] * replicas, [O] * replicas, [box] * replicas, [prm] * replicas, seed=seed,
                         trials_per_cell=trials_per_cell) as s:
        if net is not None:
            s.attach(net, learned_fraction=learned_fraction, theta_max=prm.get("theta_max", THETA))
        s.do_ensemble(20)
        c0 = s.counters_all()
        s.do_ensemble(20)
        per_step = float((s.counters_all()["tries"][:, 0] - c0["tries"][:, 0]).mean()) / n / 20
        steps = max(64, int(math.ceil(sweeps / per_step)))
        s.reset_counters()
        s.set_trace(steps)
        t0 = time.perf_counter()
        s.do_ensemble(steps)
        el = time.perf_counter() - t0
        c = s.counters_all()
        X = np.stack([s.trace(k) for k in range(replicas)], 1) / n
        for k in range(replicas):
            b, ov, tot, anyo = s.energy(k)
            assert not anyo and tot == c["bonds"][k], "tracked energy disagrees with a full recompute"
        if net is not None:
            lt, la = zip(*[s.learn_counters(k) for k in range(replicas)])
    sweeps_per_step = float(c["tries"][:, 0].mean()) / n / steps
    tau_s, err_s, _ = tau_int_replicas(X)
    tau, err = tau_s * sweeps_per_step, err_s * sweeps_per_step
    drift, drift_se = replica_drift(X)
    ok = stationary(tau, err, drift, drift_se)
    still = abs(drift) < 2.0 * drift_se
    rate = float(c["tries"][:, 0].sum()) / n / replicas / el
    out = {"form": label, "replicas": replicas, "sweeps_per_s": round(rate, 2), "run_sweeps": round(steps * sweeps_per_step),
           "tau_int_sweeps": round(tau, 2), "tau_int_err_sweeps": round(err, 2),
           "run_over_tau": round(steps * sweeps_per_step / tau, 1), "E_per_N": round(float(X.mean()), 5),
           "E_drift": round(drift, 6), "E_drift_se": round(drift_se, 6),
           "acceptance": round(float(c["accepted"][:, 0].sum() / max(c["tries"][:, 0].sum(), 1)), 5),
           "status": "stationary" if ok else ("tau error above 0.25" if still else "non-stationary"),
           "decorrelated_samples_per_s": round(rate / (2 * tau), 4) if ok else None}
    if net is not None:
        out["learned_acceptance"] = round(sum(la) / max(sum(lt), 1), 5)
    print(f"# tau {label}: {out}", file=sys.stderr, flush=True)
    return out


def ratio(a, b):
    """a / b of two decorrelation results with its standard error (from the
    errors of both tau estimates; the rates' own noise is far smaller), or
    None when either is not stationary."""
    if not a or not b or a["decorrelated_samples_per_s"] is None or b["decorrelated_samples_per_s"] is None:
        return None
    q = a["decorrelated_samples_per_s"] / b["decorrelated_samples_per_s"]
    rel = math.hypot(a["tau_int_err_sweeps"] / a["tau_int_sweeps"], b["tau_int_err_sweeps"] / b["tau_int_sweeps"])
    return {"value": round(q, 4), "se": round(q * rel, 4)}


def train_net(pp, start, prm, args, seed):
    """A ProposalNet trained on --train-steps snapshots of the plain sampler
    at the state point of `start` (one Adam step per snapshot)."""
    net = pp.ProposalNet(seed=seed)
    with pp.System(*start, seed=seed + 3, trials_per_cell=8, **prm) as s:
        t0 = time.perf_counter()
        l0, l1 = s.train(net, steps=args.train_steps, batch=args.train_batch, lr=1e-3, pos_noise=0.03,
                         theta_max=prm.get("theta_max", THETA), sweeps_between=1)
        t_train = time.perf_counter() - t0
    print(f"# trained in {t_train:.1f} s (loss {l0:.3f} -> {l1:.3f})", file=sys.stderr, flush=True)
    return net, {"steps": args.train_steps, "batch": args.train_batch, "seconds": round(t_train, 2),
                 "loss_first": round(l0, 4), "loss_last": round(l1, 4)}


def learned_vs_plain(pp, args, n, rho, temp, equil_steps, sweeps, label):
    """The learned proposal against the plain move at one state point: a
    seeded start equilibrated by --equil steps of the plain sampler (8 trials
    per cell), a network trained there, then decorrelation runs of the
    learned form (--learned-fraction of the passes learned, one move per
    non-empty cell in each), of the plain move on the same schedule (one trial
    per non-empty cell and pass) and of the plain move with 16 trials per
    cell (the fastest plain form), all from the equilibrated state."""
    r, O, box = pp.generate(n, rho, SEED + 7)
    prm = dict(T=temp, disp_max=DISP, theta_max=THETA, kf_delta=DELTA, kf_cosmax=COSMAX)
    with pp.System(r, O, box, seed=SEED + 7, trials_per_cell=8, **prm) as s:
        for k in range(0, equil_steps, 5000):
            s.do_ensemble(min(5000, equil_steps - k))
        start = s.state()
        e0 = -s.counters()["bonds"] / n
    print(f"# {label}: N={n} rho={rho} T={temp} equilibrated {equil_steps} steps, E/N={e0:.4f}", file=sys.stderr,
          flush=True)
    net, train = train_net(pp, start, prm, args, SEED)
    res = {"workload": f"NVT N={n}, rho={rho}, T={temp}", "equilibration_steps": equil_steps, "train": train}
    res["learned"] = decorrelation(pp, start, prm, seed=SEED + 11, trials_per_cell=1, sweeps=sweeps, net=net,
                                   learned_fraction=args.learned_fraction, label=f"{label} learned")
    res["plain_same_schedule"] = decorrelation(pp, start, prm, seed=SEED + 12, trials_per_cell=1, sweeps=sweeps,
                                               label=f"{label} plain 1 trial/cell")
    res["plain_16"] = decorrelation(pp, start, prm, seed=SEED + 13, trials_per_cell=16, sweeps=sweeps,
                                    label=f"{label} plain 16 trials/cell")
    net.close()
    res["learned_over_plain"] = ratio(res["learned"], res["plain_16"])
    res["learned_over_plain_same_schedule"] = ratio(res["learned"], res["plain_same_schedule"])
    res["learned_fraction"] = args.learned_fraction
    return res


def bench_learned(args, pp):
    """The learned proposal at the gel state point alone: N=65536, rho=0.3,
    T=0.09 (learned_vs_plain with --c4-equil steps of equilibration and
    --c4-sweeps sweeps per decorrelation run).  value = sweeps/s of the
    learned form."""
    res = learned_vs_plain(pp, args, 65536, 0.3, 0.09, args.c4_equil, args.c4_sweeps, "gel")
    out = {
        "metric": "particle sweeps/s + decorrelated samples/s (E/N), learned vs plain, 1 GPU",
        "value": res["learned"]["sweeps_per_s"], "unit": "sweeps/s (per replica of an 8-replica batch)", "n_gpus": 1,
        "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 geometry, bf16/f32 network", "data": "synthetic (seeded random-insertion start)",
        "config": {"workload": f"learned-proposal NVT, N=65536, rho=0.3, T=0.09, learned fraction "
                               f"{args.learned_fraction}, one move per non-empty cell and learned pass"},
        "learned_vs_plain": res,
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
    s = pp.System.batch(rs, Os, bs, prm, seed=SEED, trials_per_cell=args.trials_per_cell,
                        threads_per_cell=args.tpc, block_size=args.block)
    print(f"# {R} replicas of N={n} created", file=sys.stderr, flush=True)
    for k in range(0, args.warmup, 50):
        s.do_ensemble(min(50, args.warmup - k))
        print(f"# warmup {k + 50} steps, density {float((n / s.counters_all()['volume']).mean()):.4f}",
              file=sys.stderr, flush=True)
    tuned = None
    if not args.tpc and not args.block:
        # the launch shape of the warm state (the trajectory does not depend on it)
        tpc, block, us = s.autotune(5)
        tuned = {"threads_per_cell": tpc, "block_size": block}
        print(f"# autotuned: {tpc} lanes per cell, {block} threads per workgroup ({us:.1f} us/step)", file=sys.stderr,
              flush=True)
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
                   "autotuned": tuned,
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
                    help="untimed steps just before the timed ones (default 100; batch and npt workloads: 6000)")
    ap.add_argument("--equil", type=int, default=40000,
                    help="equilibration steps of the N=32768 system before the warmup, whatever --warmup is")
    ap.add_argument("--tau-sweeps", type=int, default=60000,
                    help="sweeps of each of the 8 copies in the headline form's tau_int run (0: none)")
    ap.add_argument("--stat-T", type=float, default=0.2,
                    help="temperature of the stationary state point (rho=0.5) of the learned comparison")
    ap.add_argument("--stat-n", type=int, default=4096,
                    help="particles per copy there (tau_int of E/N is intensive; small copies reach many tau)")
    ap.add_argument("--stat-equil", type=int, default=40000, help="equilibration steps there (8 trials per cell)")
    ap.add_argument("--stat-sweeps", type=int, default=100000,
                    help="sweeps of each copy in each decorrelation run there (0: skip the stationary point)")
    ap.add_argument("--c4-equil", type=int, default=6000,
                    help="equilibration steps of the gel (N=65536, rho=0.3, T=0.09; 8 trials per cell)")
    ap.add_argument("--c4-sweeps", type=int, default=10000,
                    help="sweeps of each copy in each decorrelation run of the gel (0: skip the gel)")
    ap.add_argument("--trials-per-cell", type=int, default=16,
                    help="trials per non-empty cell and pass (0: one per member, a step of exactly N attempts)")
    ap.add_argument("--per-member-steps", type=int, default=2000,
                    help="steps of the one-trial-per-member schedule timed after the main run (0: skip)")
    ap.add_argument("--tpc", type=int, default=0, help="threads per cell (0: default)")
    ap.add_argument("--launch", type=int, default=0,
                    help="launch form of the colour passes: 0/1 one launch per pass, 3 dependency-driven")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--occupancy", type=float, default=0.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no HIP events in the timed region")
    ap.add_argument("--time-every", type=int, default=0,
                    help="HIP events bracket the kernels of every k-th step of the timed region "
                         "(0: every step of up to 256, else 256 evenly spaced steps)")
    ap.add_argument("--train-steps", type=int, default=10000, help="Adam steps (one sampler snapshot each)")
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--learned-fraction", type=float, default=0.8)
    ap.add_argument("--workload", choices=("single", "batch", "learned", "npt"), default="single",
                    help="single: the N=32768 system (and the decorrelation legs); batch: 256 independent replicas "
                         "of uneven N, T and rho; learned: the learned proposal at N=65536, rho=0.3, T=0.09 "
                         "alone; npt: 16 NPT replicas of N=131072")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (the reference point that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profile)")
    args = ap.parse_args()
    if args.warmup is None:
        args.warmup = 100 if args.workload == "single" else 6000
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
    opts = dict(T=T, disp_max=DISP, theta_max=THETA, kf_delta=DELTA, kf_cosmax=COSMAX)
    sysm = pp.System(r, O, box, seed=SEED, trials_per_cell=args.trials_per_cell, threads_per_cell=args.tpc,
                     block_size=args.block, cell_occupancy=args.occupancy, launch=args.launch, **opts)
    print(f"# start: N={N} box={box[0]:.4f} grid={sysm.grid()}", file=sys.stderr, flush=True)
    # equilibration (whatever --warmup says): the timed steps run at the
    # state point, not on the relaxing start
    for k in range(0, args.equil, 10000):
        sysm.do_ensemble(min(10000, args.equil - k))
        print(f"# equilibration {min(k + 10000, args.equil)} steps, E/N={-sysm.counters()['bonds'] / N:.4f}",
              file=sys.stderr, flush=True)
    sysm.do_ensemble(args.warmup)
    print(f"# warmup {args.warmup} steps done, E/N={-sysm.counters()['bonds'] / N:.4f}", file=sys.stderr, flush=True)

    # timed region: K steps; HIP events bracket the kernels of every
    # time_every-th step (every step of a short run)
    time_every = args.time_every or max(1, args.steps // 256)
    sysm.reset_counters()
    sysm.set_timing(0 if args.no_timing else time_every)
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
    # one launch makes one colour pass, or all 8 in the dependency-driven form
    passes_per_launch = 8 if args.launch == 3 else 1
    moves_per_launch = attempts / args.steps / 8 * passes_per_launch  # attempted moves of one launch
    bytes_per_launch = bpm * moves_per_launch
    avg_pass_s = ms_pass / max(n_pass, 1) / 1e3
    achieved = bytes_per_launch / avg_pass_s / 1e9 if avg_pass_s > 0 else 0.0
    traffic = measured_traffic() if passes_per_launch == 1 else None  # measured for the per-pass kernel
    chain = args.trials_per_cell if args.trials_per_cell > 0 else None

    # parity spot check of the final snapshot (energy recomputed vs tracked)
    b, ov, tot, anyo = sysm.energy()
    assert not anyo and tot == c["bonds"], "tracked energy disagrees with a full recompute"
    start = sysm.state()
    sysm.close()

    per_member = None
    if args.per_member_steps > 0 and args.trials_per_cell != 0:
        # the same state advanced with one trial per member of each cell
        with pp.System(*start, seed=SEED + 1, trials_per_cell=0, **opts) as s0:
            s0.do_ensemble(50)
            s0.reset_counters()
            ta = time.perf_counter()
            s0.do_ensemble(args.per_member_steps)
            tb = time.perf_counter()
            c0 = s0.counters()
        per_member = {"trials_per_cell": 0, "sweeps_per_s": round(c0["tries"][0] / N / (tb - ta), 2),
                      "steps": args.per_member_steps}

    # decorrelated samples of the headline form at its own state point, from
    # 8 independent copies of the equilibrated state
    plain_tau = None
    if args.tau_sweeps > 0:
        plain_tau = decorrelation(pp, start, opts, seed=SEED + 4, trials_per_cell=args.trials_per_cell,
                                  sweeps=args.tau_sweeps, label="N=32768 rho=0.5 T=0.12 plain")
        # the headline rate is the single system's timed one
        if plain_tau["decorrelated_samples_per_s"] is not None:
            plain_tau["decorrelated_samples_per_s_at_value"] = round(sweeps_per_s / (2 * plain_tau["tau_int_sweeps"]), 4)
    stat_point = None
    if args.stat_sweeps > 0:
        stat_point = learned_vs_plain(pp, args, args.stat_n, RHO, args.stat_T, args.stat_equil, args.stat_sweeps,
                                      "stationary")
    gel = None
    if args.c4_sweeps > 0:
        gel = learned_vs_plain(pp, args, 65536, 0.3, 0.09, args.c4_equil, args.c4_sweeps, "gel")

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(*start, args.cpu_seconds)

    avg_launch_us = avg_pass_s * 1e6
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
            "equilibration_steps": args.equil,
            "attempts": attempts, "attempts_per_step": round(attempts / args.steps, 1),
            "acceptance": round(c["accepted"][0] / max(attempts, 1), 5),
            "E_per_N": round(float(-c["bonds"] / N), 6),
        },
        "decorrelated_samples_per_s": {
            "plain": plain_tau["decorrelated_samples_per_s_at_value"]
            if plain_tau and plain_tau["decorrelated_samples_per_s"] is not None else None,
            "status": plain_tau["status"] if plain_tau else None,
            "plain_run": plain_tau,
            "stationary_point": stat_point,
            "gel": gel,
        },
        "one_trial_per_member": per_member,
        "roofline": {
            # each cell's team runs its trials as one dependent chain: the
            # pass is bound by that chain's latency, not by bytes
            "bound": "latency",
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "achieved_is": "algorithmic bytes per launch (attempted moves x bytes per move) / average launch",
            "hbm_measured_gbs": round(traffic / avg_pass_s / 1e9, 2) if traffic and avg_pass_s > 0 else None,
            "hbm_measured_frac": round(traffic / avg_pass_s / 1e9 / HBM_PEAK_GBS, 5) if traffic and avg_pass_s > 0 else None,
            "kernel": "k_pass_dd (the 8 colour passes of a sweep)" if passes_per_launch == 8 else "k_pass (colour pass)",
            "bytes_per_move": round(bpm, 1),
            "moves_per_launch": round(moves_per_launch, 1),
            "algorithmic_bytes_per_launch": round(bytes_per_launch), "avg_launch_us": round(avg_launch_us, 3),
            "trials_per_chain": chain,
            "us_per_trial_of_chain": round(avg_launch_us / passes_per_launch / chain, 3) if chain else None,
            "cycles_per_trial_of_chain": round(avg_launch_us / passes_per_launch / chain * CLOCK_GHZ * 1e3)
            if chain else None,
            "rebuild_us_per_sweep": round(ms_reb / max(n_reb, 1) * 1e3, 3),
            "timed_steps": n_reb, "pass_launches_timed": n_pass,
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        cpu.update(host_cpu())
        out["speedup_vs_cpu"] = round(sweeps_per_s / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
