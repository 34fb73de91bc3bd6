"""Write the golden vectors of tests/golden/ from the reference itself.

Runs only where the reference sources exist (/root/reference): builds
oracle/_ref (oracle/Makefile, straight from the reference sources) and calls
the reference's own functions through oracle/_ref/libppref.so and its
executables.  Outputs are data only (inputs and expected outputs):

  pair_kat.npz      MC_would_interact on hand-built pairs at the strict
                    thresholds and across periodic faces/corners
  snap_*.npz        per-particle MC_energy (bonds, overlap) on snapshots
  traj_*.npz        positions/orientations/box after k steps of the
                    reference main loop (NVT and NPT), from a given start
  generator.json    sha256 of the reference generator's generated.rrr
  stats_ref.json    equilibrium <E/N> (and NPT <N/V>) of the reference
                    program at small state points: 8 independent runs each,
                    block statistics over the runs
  host_runs.json    the lines the reference program prints at step 0
                    (stdout, energy.dat, acceptance.dat, density.dat) on the
                    reference generator's N=1000 rho=0.1 start, NVT, NPT and
                    grand canonical
  bonds_start.rrr,  the reference program's Print_bonds file at step 0 of a
  bonds_0_ref.dat,  run from the snap_rho05_T01 liquid (N=500) and the keys
  bonds_keys.json   it ran with
  batch_ref.json    per-replica <E/N> of the reference program for 8 uneven
                    state points (N, T, rho), 8 runs each

    python tests/golden/make_golden.py            (everything)
    python tests/golden/make_golden.py host       (host_runs.json only)
    python tests/golden/make_golden.py stats      (stats_ref.json only)
    python tests/golden/make_golden.py bonds      (the bond-list files only)
    python tests/golden/make_golden.py batch      (batch_ref.json only)
    python tests/golden/make_golden.py c1         (c1_ref.json only)
    python tests/golden/make_golden.py cold       (cold_ref.json only)
    python tests/golden/make_golden.py lx         (lx_ref.json only)
    python tests/golden/make_golden.py gc         (gc_ref.json only)

  c1_ref.json       E/N series of 8 reference runs at the reference-runnable
                    configuration (N=1000, rho=0.1, T=0.1, 10^5 sweeps)
  cold_ref.json     the same for 8 state points of the batched workload's
                    distribution, down to T=0.08
  lx_ref.json       <E/N> and <Lx> of 8 reference runs with box-shape moves
                    (Lx_move = 1) under the reference's acceptance test
  gc_ref.json       <N/V> and <E/N> of 8 grand-canonical reference runs
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

REF_SRC = "/root/reference"


def build_ref():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all", "ref", f"REF={REF_SRC}"], check=True,
                   stdout=subprocess.DEVNULL)


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def align_patch(rng, target, patch_index, base):
    """A rotation whose patch `patch_index` points along unit vector target."""
    b = base[patch_index]
    v = np.cross(b, target)
    s = np.linalg.norm(v)
    c = float(np.dot(b, target))
    if s < 1e-12:
        R = np.eye(3) if c > 0 else -np.eye(3)
    else:
        k = v / s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * K @ K
    # a random spin about the target keeps the patch where it is
    t = rng.uniform(0, 2 * np.pi)
    K = np.array([[0, -target[2], target[1]], [target[2], 0, -target[0]], [-target[1], target[0], 0]])
    S = np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K
    return S @ R


def pair_cases(rng, base, delta=0.12, cosmax=0.92):
    rc = 1 + delta
    box = np.array([7.3, 8.1, 9.7])
    cases = []

    def add(rp, Op, rq, Oq, bx=box):
        cases.append((np.array(rp, float), np.array(Op, float), np.array(rq, float), np.array(Oq, float),
                      np.array(bx, float)))

    # radial thresholds along random directions, with aligned patches
    for target_d2 in (1.0, rc * rc):
        for _ in range(40):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            dist = np.sqrt(target_d2)
            dist = dist * (1 + rng.choice([-1, 0, 1]) * rng.choice([1e-16, 2e-16, 1e-15, 1e-12]))
            rp = rng.uniform(0, box)
            rq = rp + u * dist
            Op = align_patch(rng, u, rng.integers(4), base)
            Oq = align_patch(rng, -u, rng.integers(4), base)
            add(rp, Op, rq, Oq)
    # angular threshold: patch at an angle whose cosine is ~cosmax
    for _ in range(60):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        dist = rng.uniform(1.0, rc)
        rp = rng.uniform(0, box)
        rq = rp + u * dist
        th = np.arccos(cosmax) * (1 + rng.choice([-1, 1]) * rng.choice([1e-15, 1e-13, 1e-10, 1e-6]))
        w = np.cross(u, rng.normal(size=3))
        w /= np.linalg.norm(w)
        tp = np.cos(th) * u + np.sin(th) * w
        Op = align_patch(rng, tp, rng.integers(4), base)
        Oq = align_patch(rng, -u, rng.integers(4), base) if rng.random() < 0.5 else rand_rot(rng)
        add(rp, Op, rq, Oq)
    # minimum image across faces, edges and corners, coordinates far outside the box
    for _ in range(120):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        dist = rng.uniform(0.9, 1.3)
        corner = rng.integers(0, 2, size=3) * box
        rp = corner + rng.uniform(-0.3, 0.3, size=3) + rng.integers(-3, 4, size=3) * box
        rq = rp + u * dist + rng.integers(-3, 4, size=3) * box
        Op = align_patch(rng, u, rng.integers(4), base) if rng.random() < 0.7 else rand_rot(rng)
        Oq = align_patch(rng, -u, rng.integers(4), base) if rng.random() < 0.7 else rand_rot(rng)
        add(rp, Op, rq, Oq)
    # random pairs
    for _ in range(80):
        rp = rng.uniform(0, box)
        rq = rp + rng.uniform(-1.4, 1.4, size=3)
        add(rp, rand_rot(rng), rq, rand_rot(rng))
    return cases


def write_conf(path, r, O, box):
    po.rrr_save(path, 0, box, r, O)


def main():
    build_ref()
    tmp = tempfile.mkdtemp(prefix="golden_")
    rng = np.random.default_rng(20261202)
    base = po.base_patches()

    # --- reference run-time state: load a small configuration -------------
    r0, O0, box0 = po.generate(64, 0.1, seed=3)
    conf0 = os.path.join(tmp, "init0.rrr")
    write_conf(conf0, r0, O0, box0)
    ref = po.Ref(conf0, T=0.1, seed=1)

    # --- pair KATs ----------------------------------------------------------
    cases = pair_cases(rng, base)
    res = np.array([ref.pair(*c) for c in cases], np.int32)
    np.savez_compressed(os.path.join(HERE, "pair_kat.npz"),
                        rp=np.array([c[0] for c in cases]), Op=np.array([c[1] for c in cases]),
                        rq=np.array([c[2] for c in cases]), Oq=np.array([c[3] for c in cases]),
                        box=np.array([c[4] for c in cases]), expected=res, delta=0.12, cosmax=0.92)
    print("pair_kat:", np.bincount(res + 1))

    # --- snapshot energies ----------------------------------------------------
    def snapshot(name, conf, T=0.1, seed=1, sweeps=0, print_every=0, store_inputs=True, note=""):
        ref = po.Ref(conf, T=T, seed=seed)
        if sweeps:
            ref.run(sweeps, print_every)
        r, O, box = ref.state()
        ref.set_state(r, O, box)
        b, ov = ref.energies()
        out = dict(bonds=b, overlap=ov, box=box, delta=0.12, cosmax=0.92, note=note)
        out["r"], out["O"] = r, O
        np.savez_compressed(os.path.join(HERE, f"snap_{name}.npz"), **out)
        print(f"snap_{name}: N={len(b)} bonds={b.sum() // 2} overlaps={int(ov.sum())}")
        return r, O, box

    # C1 start: N=1000 rho=0.1 random sequential insertion, seed 1
    r1, O1, b1 = po.generate(1000, 0.1, seed=1)
    c1 = os.path.join(tmp, "c1.rrr")
    write_conf(c1, r1, O1, b1)
    snapshot("c1_start", c1, note="N=1000 rho=0.1 RSA seed 1 (as loaded from .rrr)")
    # bonded liquid: N=500 rho=0.5, 300 reference sweeps at T=0.1
    r2, O2, b2 = po.generate(500, 0.5, seed=2)
    c2 = os.path.join(tmp, "c2.rrr")
    write_conf(c2, r2, O2, b2)
    snapshot("rho05_T01", c2, T=0.1, seed=2, sweeps=300, print_every=100,
             note="N=500 rho=0.5 after 300 reference sweeps at T=0.1")
    # low temperature, strongly bonded
    r3, O3, b3 = po.generate(400, 0.3, seed=4)
    c3 = os.path.join(tmp, "c3.rrr")
    write_conf(c3, r3, O3, b3)
    snapshot("rho03_T005", c3, T=0.05, seed=4, sweeps=400, print_every=100,
             note="N=400 rho=0.3 after 400 reference sweeps at T=0.05")
    # the reference's own crystal seeds
    for name in ("fcc", "bcc"):
        snapshot(name, os.path.join(REF_SRC, f"{name}.rrr"), note=f"{name}.rrr of the reference, as loaded")
    # unwrapped coordinates far outside the box (positions are never wrapped)
    ref = po.Ref(c2, T=0.1, seed=2)
    ref.run(200, 100)
    r, O, box = ref.state()
    shift = rng.integers(-5, 6, size=r.shape) * box
    ref.set_state(r + shift, O, box)
    b, ov = ref.energies()
    np.savez_compressed(os.path.join(HERE, "snap_unwrapped.npz"), r=r + shift, O=O, box=box, bonds=b, overlap=ov,
                        delta=0.12, cosmax=0.92, note="N=500 rho=0.5, positions moved by random box multiples")
    print("snap_unwrapped:", b.sum() // 2)
    # a snapshot with overlaps (MC_energy returns 0 and flags them)
    r_ov = r.copy()
    r_ov[7] = r_ov[8] + np.array([0.5, 0.1, 0.0])
    r_ov[100] = r_ov[101] + np.array([0.0, 0.99, 0.0])
    ref.set_state(r_ov, O, box)
    b, ov = ref.energies()
    np.savez_compressed(os.path.join(HERE, "snap_overlap.npz"), r=r_ov, O=O, box=box, bonds=b, overlap=ov,
                        delta=0.12, cosmax=0.92, note="two planted overlaps")
    print("snap_overlap: overlaps", int(ov.sum()))

    # --- trajectories ----------------------------------------------------------
    def traj(name, conf, steps, print_every, **kw):
        ref = po.Ref(conf, **kw)
        step, box, r_in, O_in = po.rrr_load(conf)
        ref.run(steps, print_every)
        r, O, box_out = ref.state()
        tries, acc = ref.counters()
        np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), r_in=r_in, O_in=O_in, box_in=box,
                            r=r, O=O, box=box_out, energy=ref.energy, steps=steps, print_every=print_every,
                            params=json.dumps(kw))
        print(f"traj_{name}: E={ref.energy} tries={tries} acc={acc} box={box_out}")

    r4, O4, b4 = po.generate(150, 0.3, seed=6)
    c4 = os.path.join(tmp, "c4.rrr")
    write_conf(c4, r4, O4, b4)
    traj("nvt", c4, 120, 40, T=0.1, seed=9)
    r5, O5, b5 = po.generate(120, 0.15, seed=8)
    c5 = os.path.join(tmp, "c5.rrr")
    write_conf(c5, r5, O5, b5)
    traj("npt", c5, 150, 50, T=0.12, seed=13, ensemble=5, pressure=0.3, rescale_factor_max=0.1)

    # --- the generator --------------------------------------------------------
    gens = []
    for n, rho in ((300, 0.2), (1000, 0.1)):
        d = tempfile.mkdtemp(prefix="gen_")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(n), str(rho)], cwd=d, check=True,
                       capture_output=True)
        with open(os.path.join(d, "generated.rrr"), "rb") as f:
            gens.append(dict(n=n, rho=rho, sha256=hashlib.sha256(f.read()).hexdigest()))
    with open(os.path.join(HERE, "generator.json"), "w") as f:
        json.dump(gens, f, indent=1)
    print("generator:", gens)


def host_runs():
    """Step-0 outputs of the reference program: same input keys and start as
    the host-program test uses (the start is the reference generator's
    generated.rrr for N=1000, rho=0.1, which generator_hip reproduces)."""
    build_ref()
    d = tempfile.mkdtemp(prefix="host_")
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), "1000", "0.1"], cwd=d, check=True,
                   capture_output=True)
    out = {}
    for name, extra in (("nvt", {}), ("npt", {"Ensemble": 5, "P": 0.05, "rescale_factor_max": 0.05}),
                        ("gc", {"Ensemble": 1, "Activity": 0.1, "GC_N_max": 3000})):
        keys = {"Dynamics": 0, "Ensemble": 0, "Seed": 17, "Temperature": 0.1, "Disp_max": 0.05, "Theta_max": 0.1,
                "KF_delta": 0.12, "KF_cosmax": 0.92, "Initial_conditions_file": "generated.rrr", "Steps": 1,
                "Print_every": 1, "Save_every": 1000, "Configuration_folder": "."}
        keys.update(extra)
        with open(os.path.join(d, "input"), "w") as f:
            f.write("".join(f"{k} = {v}\n" for k, v in keys.items()))
        for fn in ("energy.dat", "acceptance.dat", "density.dat"):
            if os.path.exists(os.path.join(d, fn)):
                os.remove(os.path.join(d, fn))
        res = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "PatchyParticles"), "input"], cwd=d, check=True,
                             capture_output=True, text=True)

        def first(fn):
            p = os.path.join(d, fn)
            return open(p).readline() if os.path.exists(p) else None

        init = [x for x in res.stderr.splitlines() if x.startswith("Initial energy")]
        out[name] = dict(keys=keys, stdout0=res.stdout.splitlines()[0] + "\n", energy0=first("energy.dat"),
                         acceptance0=first("acceptance.dat"), density0=first("density.dat"), initial=init[0])
    with open(os.path.join(HERE, "host_runs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("host_runs:", json.dumps(out)[:300])


def bonds_ref():
    """The reference program's bond list (Print_bonds = 1) at step 0 of a run
    from a bonded liquid (snap_rho05_T01: N=500, rho=0.5), written by
    output_print_bonds right after the cells were filled: bonds_start.rrr
    (the input) and bonds_0_ref.dat (its output)."""
    build_ref()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    z = np.load(os.path.join(HERE, "snap_rho05_T01.npz"))
    d = tempfile.mkdtemp(prefix="bonds_")
    start = os.path.join(HERE, "bonds_start.rrr")
    pyoracle.rrr_save(start, 0, z["box"], z["r"], z["O"])
    shutil.copy(start, os.path.join(d, "start.rrr"))
    os.mkdir(os.path.join(d, "bonds"))
    keys = {"Dynamics": 0, "Ensemble": 0, "Seed": 3, "Temperature": 0.1, "Disp_max": 0.05, "Theta_max": 0.1,
            "KF_delta": 0.12, "KF_cosmax": 0.92, "Initial_conditions_file": "start.rrr", "Steps": 1,
            "Print_every": 1, "Save_every": 1000, "Configuration_folder": ".", "Print_bonds": 1,
            "Bonds_folder": "bonds"}
    with open(os.path.join(d, "input"), "w") as f:
        f.write("".join(f"{k} = {v}\n" for k, v in keys.items()))
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", "PatchyParticles"), "input"], cwd=d, check=True,
                   capture_output=True)
    shutil.copy(os.path.join(d, "bonds", "bonds_0.dat"), os.path.join(HERE, "bonds_0_ref.dat"))
    with open(os.path.join(HERE, "bonds_keys.json"), "w") as f:
        json.dump(keys, f, indent=1)
    print("bonds: reference bond list of", len(z["r"]), "particles")


BATCH_REPLICAS, BATCH_RUNS, BATCH_SWEEPS = 8, 8, 40000


def batch_points():
    """8 uneven replicas in the manner of the batched workload, restricted to
    state points that equilibrate well within the runs: N log-uniform in
    [500, 1500], T uniform in [0.19, 0.23], rho uniform in [0.2, 0.5]."""
    rng = np.random.default_rng(20261203)
    n = np.exp(rng.uniform(np.log(500), np.log(1500), BATCH_REPLICAS)).astype(int)
    T = rng.uniform(0.19, 0.23, BATCH_REPLICAS)
    rho = rng.uniform(0.2, 0.5, BATCH_REPLICAS)
    return [dict(n=int(n[k]), T=round(float(T[k]), 4), rho=round(float(rho[k]), 4)) for k in range(BATCH_REPLICAS)]


def batch_ref():
    """Per-replica <E/N> of the reference program for batch_points(): 8
    independent runs each from the reference generator's start, mean over
    the second half of each run, then mean and standard error over runs."""
    build_ref()
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    pts = batch_points()
    jobs = []
    for k, sp in enumerate(pts):
        d0 = tempfile.mkdtemp(prefix=f"batch{k}_")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(sp["n"]), str(sp["rho"])], cwd=d0,
                       check=True, capture_output=True)
        for j in range(BATCH_RUNS):
            d = os.path.join(d0, f"run{j}")
            os.makedirs(d)
            keys = {"Dynamics": 0, "Ensemble": 0, "Seed": 2000 + j, "Temperature": sp["T"], "Disp_max": 0.05,
                    "Theta_max": 0.1, "KF_delta": 0.12, "KF_cosmax": 0.92,
                    "Initial_conditions_file": os.path.join(d0, "generated.rrr"), "Steps": BATCH_SWEEPS,
                    "Print_every": STAT_PRINT, "Save_every": BATCH_SWEEPS, "Configuration_folder": "."}
            with open(os.path.join(d, "input"), "w") as f:
                f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
            jobs.append((k, d))
    running = []
    for k, d in jobs:  # at most os.cpu_count() reference runs at once
        while len(running) >= (os.cpu_count() or 4):
            running.pop(0).wait()
        running.append(subprocess.Popen([exe, "input"], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for p in running:
        assert p.wait() == 0
    out = []
    for k, sp in enumerate(pts):
        e_runs = []
        for kk, d in jobs:
            if kk != k:
                continue
            e = np.loadtxt(os.path.join(d, "energy.dat"))[:, 1]
            e_runs.append(e[int(STAT_DISCARD * len(e)):].mean())
        e_runs = np.array(e_runs)
        out.append(dict(sp, E_per_N=float(e_runs.mean()), E_per_N_se=float(e_runs.std(ddof=1) / np.sqrt(len(e_runs))),
                        E_runs=e_runs.tolist()))
        print("batch", k, out[-1]["n"], out[-1]["T"], out[-1]["rho"], out[-1]["E_per_N"], out[-1]["E_per_N_se"])
    with open(os.path.join(HERE, "batch_ref.json"), "w") as f:
        json.dump(dict(sweeps=BATCH_SWEEPS, print_every=STAT_PRINT, discard=STAT_DISCARD, runs=BATCH_RUNS,
                       replicas=out), f, indent=1)


# state points of the statistical comparison: the reference program's own
# output series (energy.dat, density.dat) from its generator's start
STAT_POINTS = {
    "nvt_rho03_T020": dict(n=300, rho=0.3, keys={"Ensemble": 0, "Temperature": 0.2}),
    "nvt_rho05_T020": dict(n=300, rho=0.5, keys={"Ensemble": 0, "Temperature": 0.2}),
    # the NPT density relaxes for ~2.5e5 sweeps from the rho = 0.3 start: the
    # runs start from a reference-equilibrated configuration (a committed
    # input file, stat_npt_T020_P02_start.rrr) written by a first long run
    "npt_T020_P02": dict(n=300, rho=0.3, equilibrate=400000,
                         keys={"Ensemble": 5, "Temperature": 0.2, "P": 0.2, "rescale_factor_max": 0.05}),
}
# (lower temperatures keep relaxing for 10^5+ sweeps at these sizes: not an
# equilibrium comparison)
STAT_RUNS, STAT_SWEEPS, STAT_PRINT, STAT_DISCARD = 8, 60000, 100, 0.5


def stats_ref():
    """8 independent runs of the reference program (production flags) per
    state point; the mean of each run over its last 60 % of print steps,
    then mean and standard error over the runs."""
    build_ref()
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    out = {}
    for name, sp in STAT_POINTS.items():
        d0 = tempfile.mkdtemp(prefix=f"stat_{name}_")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(sp["n"]), str(sp["rho"])], cwd=d0,
                       check=True, capture_output=True)
        start = os.path.join(d0, "generated.rrr")
        if sp.get("equilibrate"):
            d = os.path.join(d0, "equil")
            os.makedirs(d)
            keys = {"Dynamics": 0, "Seed": 999, "Disp_max": 0.05, "Theta_max": 0.1, "KF_delta": 0.12,
                    "KF_cosmax": 0.92, "Initial_conditions_file": start, "Steps": sp["equilibrate"],
                    "Print_every": 10000, "Save_every": sp["equilibrate"], "Configuration_folder": "."}
            keys.update(sp["keys"])
            with open(os.path.join(d, "input"), "w") as f:
                f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
            subprocess.run([exe, "input"], cwd=d, check=True, capture_output=True)
            start = os.path.join(HERE, f"stat_{name}_start.rrr")
            with open(os.path.join(d, "last.rrr")) as f, open(start, "w") as g:
                g.write(f.read())
        procs = []
        for k in range(STAT_RUNS):
            d = os.path.join(d0, f"run{k}")
            os.makedirs(d)
            keys = {"Dynamics": 0, "Seed": 1000 + k, "Disp_max": 0.05, "Theta_max": 0.1, "KF_delta": 0.12,
                    "KF_cosmax": 0.92, "Initial_conditions_file": start,
                    "Steps": STAT_SWEEPS, "Print_every": STAT_PRINT, "Save_every": STAT_SWEEPS,
                    "Configuration_folder": "."}
            keys.update(sp["keys"])
            with open(os.path.join(d, "input"), "w") as f:
                f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
            procs.append(subprocess.Popen([exe, "input"], cwd=d, stdout=subprocess.DEVNULL,
                                          stderr=subprocess.DEVNULL))
        for p in procs:
            assert p.wait() == 0
        e_runs, r_runs = [], []
        for k in range(STAT_RUNS):
            d = os.path.join(d0, f"run{k}")
            e = np.loadtxt(os.path.join(d, "energy.dat"))[:, 1]
            e_runs.append(e[int(STAT_DISCARD * len(e)):].mean())
            if sp["keys"]["Ensemble"] == 5:
                rho = np.loadtxt(os.path.join(d, "density.dat"))[:, 1]
                r_runs.append(rho[int(STAT_DISCARD * len(rho)):].mean())
        e_runs = np.array(e_runs)
        res = dict(n=sp["n"], rho0=sp["rho"], keys=sp["keys"], runs=STAT_RUNS, sweeps=STAT_SWEEPS,
                   start=os.path.basename(start) if sp.get("equilibrate") else None,
                   print_every=STAT_PRINT, discard=STAT_DISCARD, E_per_N=float(e_runs.mean()),
                   E_per_N_se=float(e_runs.std(ddof=1) / np.sqrt(STAT_RUNS)), E_runs=e_runs.tolist())
        if r_runs:
            r_runs = np.array(r_runs)
            res.update(density=float(r_runs.mean()), density_se=float(r_runs.std(ddof=1) / np.sqrt(STAT_RUNS)),
                       density_runs=r_runs.tolist())
        out[name] = res
        print(name, {k: v for k, v in res.items() if not k.endswith("_runs")})
    with open(os.path.join(HERE, "stats_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


# The state points of the configurations themselves, where the system keeps
# relaxing over a whole run: the reference-runnable configuration (N=1000,
# rho=0.1, T=0.1, 10^5 sweeps from the generator's start) and 8 points drawn
# from the batched workload's distribution (N log-uniform in [500, 8000], T
# uniform in [0.08, 0.20], rho uniform in [0.05, 0.60]; this draw holds
# three points below T=0.1).  The whole E/N series of every run is kept, so
# the GPU can be compared with it at matched attempt counts.
SERIES_RUNS, SERIES_SWEEPS, SERIES_PRINT = 8, 100000, 1000
C1_POINT = dict(n=1000, rho=0.1, T=0.1)


def cold_points():
    rng = np.random.default_rng(20268)
    n = np.exp(rng.uniform(np.log(500), np.log(8000), 8)).astype(int)
    T = rng.uniform(0.08, 0.20, 8)
    rho = rng.uniform(0.05, 0.60, 8)
    return [dict(n=int(n[k]), T=round(float(T[k]), 4), rho=round(float(rho[k]), 4)) for k in range(8)]


def series_ref(points, out_name, seed0):
    """SERIES_RUNS runs of the reference program (production flags) per
    point from its generator's start; energy.dat of every run."""
    build_ref()
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    jobs = []
    for k, sp in enumerate(points):
        d0 = tempfile.mkdtemp(prefix=f"series{k}_")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(sp["n"]), str(sp["rho"])], cwd=d0,
                       check=True, capture_output=True)
        for j in range(SERIES_RUNS):
            d = os.path.join(d0, f"run{j}")
            os.makedirs(d)
            keys = {"Dynamics": 0, "Ensemble": 0, "Seed": seed0 + 100 * k + j, "Temperature": sp["T"],
                    "Disp_max": 0.05, "Theta_max": 0.1, "KF_delta": 0.12, "KF_cosmax": 0.92,
                    "Initial_conditions_file": os.path.join(d0, "generated.rrr"), "Steps": SERIES_SWEEPS,
                    "Print_every": SERIES_PRINT, "Save_every": SERIES_SWEEPS, "Configuration_folder": "."}
            with open(os.path.join(d, "input"), "w") as f:
                f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
            jobs.append((sp["n"], k, d))
    running = []
    for _, k, d in sorted(jobs, key=lambda j: -j[0]):  # longest first, one run per core
        while len(running) >= (os.cpu_count() or 4):
            for p in list(running):
                if p.poll() is not None:
                    assert p.returncode == 0
                    running.remove(p)
            if len(running) >= (os.cpu_count() or 4):
                running[0].wait()
        running.append(subprocess.Popen([exe, "input"], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for p in running:
        assert p.wait() == 0
    out = []
    for k, sp in enumerate(points):
        series = []
        for _, kk, d in jobs:
            if kk == k:
                e = np.loadtxt(os.path.join(d, "energy.dat"))
                assert len(e) == SERIES_SWEEPS // SERIES_PRINT  # steps 0, P, ..., Steps - P
                series.append(e[:, 1].tolist())
        half = np.array(series)[:, len(series[0]) // 2:].mean(axis=1)
        out.append(dict(sp, E_series=series, E_late=float(half.mean()),
                        E_late_se=float(half.std(ddof=1) / np.sqrt(len(half)))))
        print(out_name, k, sp, "late E/N", out[-1]["E_late"], "+-", out[-1]["E_late_se"])
    with open(os.path.join(HERE, out_name), "w") as f:
        json.dump(dict(sweeps=SERIES_SWEEPS, print_every=SERIES_PRINT, runs=SERIES_RUNS, disp_max=0.05,
                       theta_max=0.1, points=out), f)


# Box-shape moves (Lx_move = 1) with the reference's own acceptance test
# (MC.c:562): N=300 from the generator's start at rho=0.3, T=0.2, Ly = Lz
# allowed within 10 % of the start, Lx_change_max 0.5; energy.dat and
# boxshape.dat of 8 runs.
LX_POINT = dict(n=300, rho=0.3, T=0.2, lx_change_max=0.5, lyz_rel=0.1)
LX_RUNS, LX_SWEEPS, LX_PRINT = 8, 60000, 100


def lx_ref():
    build_ref()
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    sp = LX_POINT
    d0 = tempfile.mkdtemp(prefix="lx_")
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(sp["n"]), str(sp["rho"])], cwd=d0,
                   check=True, capture_output=True)
    with open(os.path.join(d0, "generated.rrr")) as f:
        L = float(f.readline().split()[2])
    lyz = (L * (1 - sp["lyz_rel"]), L * (1 + sp["lyz_rel"]))
    procs = []
    for j in range(LX_RUNS):
        d = os.path.join(d0, f"run{j}")
        os.makedirs(d)
        keys = {"Dynamics": 0, "Ensemble": 0, "Seed": 7000 + j, "Temperature": sp["T"], "Disp_max": 0.05,
                "Theta_max": 0.1, "KF_delta": 0.12, "KF_cosmax": 0.92, "Lx_move": 1,
                "Lx_change_max": sp["lx_change_max"], "Lyz_min": lyz[0], "Lyz_max": lyz[1],
                "Initial_conditions_file": os.path.join(d0, "generated.rrr"), "Steps": LX_SWEEPS,
                "Print_every": LX_PRINT, "Save_every": LX_SWEEPS, "Configuration_folder": "."}
        with open(os.path.join(d, "input"), "w") as f:
            f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
        procs.append(subprocess.Popen([exe, "input"], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for p in procs:
        assert p.wait() == 0
    e_runs, lx_runs, acc = [], [], []
    for j in range(LX_RUNS):
        d = os.path.join(d0, f"run{j}")
        e = np.loadtxt(os.path.join(d, "energy.dat"))[:, 1]
        b = np.loadtxt(os.path.join(d, "boxshape.dat"))[:, 1]
        e_runs.append(e[int(STAT_DISCARD * len(e)):].mean())
        lx_runs.append(b[int(STAT_DISCARD * len(b)):].mean())
    e_runs, lx_runs = np.array(e_runs), np.array(lx_runs)
    out = dict(sp, box0=L, lyz_min=lyz[0], lyz_max=lyz[1], runs=LX_RUNS, sweeps=LX_SWEEPS, print_every=LX_PRINT,
               discard=STAT_DISCARD, E_per_N=float(e_runs.mean()),
               E_per_N_se=float(e_runs.std(ddof=1) / np.sqrt(LX_RUNS)), Lx=float(lx_runs.mean()),
               Lx_se=float(lx_runs.std(ddof=1) / np.sqrt(LX_RUNS)), E_runs=e_runs.tolist(), Lx_runs=lx_runs.tolist())
    print("lx", {k: v for k, v in out.items() if not k.endswith("_runs")})
    with open(os.path.join(HERE, "lx_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


# Grand-canonical runs (Ensemble = 1, MC_add_remove, MC.c:342-448): T=0.2,
# Activity 0.2 in a box of side 10 from the generator's N=100 start, GC_N_max
# 1000; the mean density and E/N of 8 runs over their second halves.
GC_POINT = dict(n0=100, rho0=0.1, T=0.2, activity=0.2, n_max=1000)
GC_RUNS, GC_SWEEPS, GC_PRINT = 8, 60000, 100


def gc_ref():
    build_ref()
    exe = os.path.join(ROOT, "oracle", "_ref", "PatchyParticles_fast")
    sp = GC_POINT
    d0 = tempfile.mkdtemp(prefix="gc_")
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", "generator"), str(sp["n0"]), str(sp["rho0"])], cwd=d0,
                   check=True, capture_output=True)
    procs = []
    for j in range(GC_RUNS):
        d = os.path.join(d0, f"run{j}")
        os.makedirs(d)
        keys = {"Dynamics": 0, "Ensemble": 1, "Seed": 9000 + j, "Temperature": sp["T"], "Disp_max": 0.05,
                "Theta_max": 0.1, "KF_delta": 0.12, "KF_cosmax": 0.92, "Activity": sp["activity"],
                "GC_N_max": sp["n_max"], "Initial_conditions_file": os.path.join(d0, "generated.rrr"),
                "Steps": GC_SWEEPS, "Print_every": GC_PRINT, "Save_every": GC_SWEEPS, "Configuration_folder": "."}
        with open(os.path.join(d, "input"), "w") as f:
            f.write("".join(f"{a} = {b}\n" for a, b in keys.items()))
        procs.append(subprocess.Popen([exe, "input"], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for p in procs:
        assert p.wait() == 0
    e_runs, rho_runs = [], []
    for j in range(GC_RUNS):
        d = os.path.join(d0, f"run{j}")
        e = np.loadtxt(os.path.join(d, "energy.dat"))[:, 1]
        rho = np.loadtxt(os.path.join(d, "density.dat"))[:, 1]
        e_runs.append(e[int(STAT_DISCARD * len(e)):].mean())
        rho_runs.append(rho[int(STAT_DISCARD * len(rho)):].mean())
    e_runs, rho_runs = np.array(e_runs), np.array(rho_runs)
    with open(os.path.join(d0, "generated.rrr")) as f:
        L = float(f.readline().split()[2])
    out = dict(sp, box=L, runs=GC_RUNS, sweeps=GC_SWEEPS, print_every=GC_PRINT, discard=STAT_DISCARD,
               E_per_N=float(e_runs.mean()), E_per_N_se=float(e_runs.std(ddof=1) / np.sqrt(GC_RUNS)),
               density=float(rho_runs.mean()), density_se=float(rho_runs.std(ddof=1) / np.sqrt(GC_RUNS)),
               E_runs=e_runs.tolist(), density_runs=rho_runs.tolist())
    print("gc", {k: v for k, v in out.items() if not k.endswith("_runs")})
    with open(os.path.join(HERE, "gc_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if sys.argv[1:] == ["gc"]:
        gc_ref()
        sys.exit(0)
    if sys.argv[1:] == ["lx"]:
        lx_ref()
        sys.exit(0)
    if sys.argv[1:] == ["c1"]:
        series_ref([C1_POINT], "c1_ref.json", 3000)
        sys.exit(0)
    if sys.argv[1:] == ["cold"]:
        series_ref(cold_points(), "cold_ref.json", 5000)
        sys.exit(0)
    if sys.argv[1:] == ["host"]:
        host_runs()
    elif sys.argv[1:] == ["stats"]:
        stats_ref()
    elif sys.argv[1:] == ["bonds"]:
        bonds_ref()
    elif sys.argv[1:] == ["batch"]:
        batch_ref()
    else:
        main()
        host_runs()
        bonds_ref()
        stats_ref()
        batch_ref()
