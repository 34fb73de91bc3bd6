"""GPU parity: the HIP kernels against the CPU oracle and the reference's
golden vectors.  Integer results (bonds, overlap flags) must be bit-exact."""
import math

import numpy as np
import pytest

from conftest import golden, golden_files

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


@pytest.mark.parametrize("name", golden_files("snap_"))
def test_energy_snapshot_matches_reference(pp, name):
    g = golden(name)
    b, ov, tot, anyo = pp.energy_snapshot(g["r"], g["O"], g["box"], float(g["delta"]), float(g["cosmax"]))
    assert np.array_equal(b, g["bonds"]), name
    assert np.array_equal(ov, g["overlap"].astype(bool)), name
    assert anyo == bool(g["overlap"].any())


def test_pair_known_answers_on_gpu(pp):
    """Each reference pair KAT as a two-particle snapshot."""
    g = golden("pair_kat.npz")
    for i in range(len(g["expected"])):
        r = np.stack([g["rp"][i], g["rq"][i]])
        O = np.stack([g["Op"][i], g["Oq"][i]])
        b, ov, _, _ = pp.energy_snapshot(r, O, g["box"][i])
        want = int(g["expected"][i])
        if want == -1:
            assert ov.all(), i
        else:
            assert not ov.any() and list(b) == [want, want], i


def _start(oracle, n, rho, seed):
    r, O, box = oracle.generate(n, rho, seed=seed)
    return r, O, box


def test_sweeps_track_energy_and_match_oracle(pp, oracle):
    """N=4000, rho=0.5, T=0.12: the incrementally tracked energy equals a full
    recompute, and the GPU per-particle energies of the evolved snapshot equal
    the oracle's (the reference energy routine) bit for bit."""
    r, O, box = _start(oracle, 4000, 0.5, 3)
    with pp.System(r, O, box, T=0.12, seed=42) as s:
        for _ in range(3):
            s.do_ensemble(40)
            c = s.counters()
            b, ov, tot, anyo = s.energy()
            assert not anyo
            assert c["bonds"] == tot
            rr, OO, bb = s.state()
            ob, oov, otot = oracle.energy(rr, OO, bb)
            assert np.array_equal(b, ob)
            assert not oov.any()
        c = s.counters()
        assert c["tries"][0] == 120 * 4000
        assert 0.3 < c["accepted"][0] / c["tries"][0] < 0.99
        assert tot > 0


@pytest.mark.parametrize("trials_per_cell", [0, 16])
def test_config2_snapshot_parity(pp, oracle, trials_per_cell):
    """N=32768, rho=0.5, T=0.12: snapshot bonds/overlaps bit-exact against the
    oracle after 1000 sweeps, with one trial per member and with the
    benchmarked form (16 trials per non-empty cell, kept bonds)."""
    r, O, box = _start(oracle, 32768, 0.5, 1)
    with pp.System(r, O, box, T=0.12, seed=7, trials_per_cell=trials_per_cell) as s:
        s.do_ensemble(1000)
        rr, OO, bb = s.state()
        b, ov, tot, anyo = s.energy()
        ob, oov, otot = oracle.energy(rr, OO, bb)
        assert np.array_equal(b, ob) and not anyo and not oov.any()
        assert s.counters()["bonds"] == tot == otot
        s.check_energy()


def test_check_energy_gram_schmidt_bit_exact(pp, oracle):
    r, O, box = _start(oracle, 2000, 0.3, 5)
    with pp.System(r, O, box, T=0.1, seed=3) as s:
        s.do_ensemble(30)
        r1, O1, _ = s.state()
        s.check_energy()
        r2, O2, _ = s.state()
    assert np.array_equal(r1, r2)
    want = np.array([oracle.gram_schmidt(o) for o in O1])
    assert np.array_equal(O2, want)


@pytest.mark.parametrize("tpc,block,launch,trials", [(32, 64, 1, 0), (32, 128, 2, 0), (32, 64, 2, 0), (64, 64, 2, 0),
                                                     (64, 128, 1, 0), (64, 256, 2, 0), (16, 64, 1, 0), (16, 64, 1, 16),
                                                     (16, 64, 1, 4), (32, 128, 1, 16)])
def test_trajectory_independent_of_launch_shape(pp, oracle, tpc, block, launch, trials):
    """Lanes per cell (16, 32 or 64: 4, 2 or 1 cells per wavefront),
    workgroup size and launch form (one launch per colour pass, or one
    cooperative launch per sweep with grid barriers) leave the trajectory
    unchanged bit for bit, with one trial per member and with 16 trials per
    cell (kept bonds)."""
    r, O, box = _start(oracle, 3000, 0.4, 9)
    ref = None
    kw = dict(T=0.11, seed=99, trials_per_cell=trials)
    with pp.System(r, O, box, threads_per_cell=64, block_size=256, launch=1, **kw) as s:
        s.do_ensemble(25)
        ref = s.state()
        cref = s.counters()
    with pp.System(r, O, box, threads_per_cell=tpc, block_size=block, launch=launch, **kw) as s:
        s.do_ensemble(25)
        got = s.state()
        c = s.counters()
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert c == cref


def test_upload_download_roundtrip(pp, oracle):
    r, O, box = _start(oracle, 1500, 0.3, 2)
    with pp.System(r, O, box, T=0.1, seed=1) as s:
        r1, O1, b1 = s.state()
        assert np.array_equal(r1, r) and np.array_equal(O1, O) and np.array_equal(b1, box)
        s.do_ensemble(5)
        r2, O2, b2 = s.state()
        s.upload(0, r, O, box)
        r3, O3, b3 = s.state()
        assert np.array_equal(r3, r) and np.array_equal(O3, O)
        _, _, tot, _ = s.energy()
        assert s.counters()["bonds"] == tot


def test_overlapping_start_is_refused(pp, oracle):
    r, O, box = _start(oracle, 200, 0.3, 2)
    r[5] = r[6] + np.array([0.3, 0.0, 0.0])
    with pytest.raises(pp.PPError) as e:
        pp.System(r, O, box, T=0.1)
    assert e.value.status == 3


@pytest.mark.parametrize("trials_per_cell,tpc", [(0, 0), (16, 0), (16, 64)])
def test_batched_replicas_match_single(pp, oracle, trials_per_cell, tpc):
    """Replica 0 of a batch follows the same trajectory as the same system
    alone (Philox streams are keyed by replica index).  The lone system runs
    the single-replica form of the colour pass (scalar replica constants),
    the batch the general one; with 16 trials per cell both keep static and
    member bonds across trials."""
    ra, Oa, ba = _start(oracle, 1000, 0.2, 1)
    rb, Ob, bb = _start(oracle, 700, 0.5, 2)
    kw = dict(seed=5, trials_per_cell=trials_per_cell, threads_per_cell=tpc)
    with pp.System(ra, Oa, ba, T=0.1, **kw) as s:
        s.do_ensemble(20)
        single = s.state()
    prm = [dict(T=0.1), dict(T=0.15)]
    with pp.System.batch([ra, rb], [Oa, Ob], [ba, bb], prm, **kw) as s:
        s.do_ensemble(20)
        both = s.state(0)
        for k in range(2):
            b, ov, tot, anyo = s.energy(k)
            assert s.counters(k)["bonds"] == tot and not anyo
    for a, b in zip(single, both):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("trials_per_cell", [0, 2])
def test_two_particle_bonding_probability_gpu(pp, trials_per_cell):
    """256 replicas of two particles in a periodic box L = 5 (4^3 sweep cells):
    P(bonded) = Vb e^{1/T} / (Vb e^{1/T} + V - 4pi/3 - Vb), with one trial per
    member of a cell per pass or a fixed number per non-empty cell."""
    delta, cosmax, T, L = 0.12, 0.92, 0.125, 5.0
    Vb = 16 * math.pi * (delta ** 3 + 3 * delta ** 2 + 3 * delta) * (1 - cosmax) ** 2 / 3
    w = Vb * math.exp(1 / T)
    p_exact = w / (w + L ** 3 - 4 * math.pi / 3 - Vb)
    R = 256
    rng = np.random.default_rng(0)
    rs, Os, bs = [], [], []
    for _ in range(R):
        rs.append(np.array([[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]]) + rng.uniform(0, 0.3, (2, 3)))
        Os.append(np.array([np.eye(3), np.eye(3)]))
        bs.append(np.array([L, L, L]))
    prm = [dict(T=T, disp_max=0.8, theta_max=1.5)] * R
    with pp.System.batch(rs, Os, bs, prm, seed=17, trials_per_cell=trials_per_cell) as s:
        s.do_ensemble(5000)
        samples = []
        for _ in range(100):
            s.do_ensemble(200)
            samples.append(s.counters_all()["bonds"])
    per_rep = np.array(samples).mean(axis=0)  # independent replicas
    m, se = per_rep.mean(), per_rep.std(ddof=1) / np.sqrt(R)
    assert abs(m - p_exact) < 4 * se + 2e-3, (m, se, p_exact)


def _fcc(cells, a, seed):
    """FCC lattice of cells^3 unit cells (lattice constant a), random rotations."""
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(cells)] * 3, indexing="ij"), -1).reshape(-1, 1, 3)
    r = ((g + base[None]) * a).reshape(-1, 3) + 0.25 * a
    rng = np.random.default_rng(seed)
    q, rr = np.linalg.qr(rng.normal(size=(len(r), 3, 3)))
    q = q * np.sign(np.diagonal(rr, axis1=1, axis2=2))[:, None, :]
    q[np.linalg.det(q) < 0, :, 0] *= -1
    return r, q, np.full(3, cells * a)


@pytest.mark.parametrize("trials", [0, 16])
@pytest.mark.parametrize("lattice", ["rho0.89", "rho0.70"])
@pytest.mark.parametrize("tpc", [16, 32, 64])
def test_dense_neighbourhood_beyond_registers(pp, oracle, tpc, lattice, trials):
    """FCC at rho=0.89 in a box of 4^3 sweep cells of side 1.65: the 26 cells
    around a sweep cell hold ~104 particles, more than one team keeps in
    registers (32 or 64), so the colour pass reads the rest from memory.  FCC
    at rho=0.70 in 12^3 cells of side 1.19: the 27 cells hold ~32 particles,
    so with 32 register slots per team (16 lanes x 2, or 32 lanes) the cell's
    members and its first static neighbours fill them and about half the
    cells read their last few static neighbours from memory, while 64-lane
    teams keep all of them.  Overlaps and bonds must still be seen: the run
    stays overlap-free, the tracked energy equals a full recompute, the bonds
    equal the oracle's, and another team width gives the same trajectory.
    With 16 trials per cell, 16-lane teams keep only the static neighbours
    a member can reach during the pass and compact them into their register
    slots: many of them here, against the few of a liquid."""
    cells, a, nside = (4, 1.65, 4) if lattice == "rho0.89" else (8, (4 / 0.7) ** (1 / 3), 12)
    r, O, box = _fcc(cells, a, 4)
    with pp.System(r, O, box, T=0.12, seed=21, threads_per_cell=tpc, trials_per_cell=trials) as s:
        assert s.grid()[0] == [nside] * 3
        s.do_ensemble(40)
        rr, OO, bb = s.state()
        b, ov, tot, anyo = s.energy()
        c = s.counters()
    ob, oov, otot = oracle.energy(rr, OO, bb)
    assert not anyo and not oov.any()
    assert np.array_equal(b, ob) and c["bonds"] == tot == otot
    assert c["accepted"][0] > 0.2 * c["tries"][0]
    with pp.System(r, O, box, T=0.12, seed=21, threads_per_cell=64 if tpc != 64 else 32, trials_per_cell=trials) as s:
        s.do_ensemble(40)
        other = s.state()
    for x, y in zip((rr, OO, bb), other):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("trials", [0, 16])
def test_small_cell_pass_defers_full_cells(pp, oracle, monkeypatch, trials):
    """Dilute systems (fewer than 16 candidates in a cell's 27 cells) take the
    small-cell pass with 16-lane teams: half the LDS of a general team, and a
    cell of more than 8 members is deferred to the general 32-lane pass that
    follows.  With the threshold forced down to one member
    (PP_SMALL_CELL_CAP=1), every cell of two or more members takes the
    deferred path.  N=3000 at rho=0.15: the trajectory is the same bit for bit
    with the threshold at 8, at 1 and with 32-lane teams, and the final bonds
    equal the oracle's."""
    r, O, box = _start(oracle, 3000, 0.15, 12)

    def run(tpc, cap):
        if cap:
            monkeypatch.setenv("PP_SMALL_CELL_CAP", str(cap))
        else:
            monkeypatch.delenv("PP_SMALL_CELL_CAP", raising=False)
        with pp.System(r, O, box, T=0.12, seed=31, threads_per_cell=tpc, trials_per_cell=trials) as s:
            s.do_ensemble(30)
            return s.state(), s.counters(), s.energy()

    ref = run(32, 0)
    for tpc, cap in ((16, 0), (16, 1)):
        got = run(tpc, cap)
        for x, y in zip(ref[0], got[0]):
            assert np.array_equal(x, y), (tpc, cap)
        assert got[1] == ref[1], (tpc, cap)
    rr, OO, bb = ref[0]
    b, ov, tot, anyo = ref[2]
    ob, oov, otot = oracle.energy(rr, OO, bb)
    assert not anyo and np.array_equal(b, ob) and tot == otot == ref[1]["bonds"]


def test_small_cell_pass_in_a_batch(pp, oracle, monkeypatch):
    """The small-cell pass of several replicas (replica constants per team,
    orientations in LDS): four dilute replicas of uneven N and T, 16 trials
    per cell, with the deferral threshold at 8 and at 1 member, give the
    trajectories of 32-lane teams bit for bit."""
    starts = [oracle.generate(n, rho, seed=60 + k) for k, (n, rho) in enumerate([(800, 0.1), (1500, 0.15), (600, 0.05),
                                                                                  (1200, 0.2)])]
    prm = [dict(T=0.12), dict(T=0.15), dict(T=0.1), dict(T=0.2)]

    def run(tpc, cap):
        if cap:
            monkeypatch.setenv("PP_SMALL_CELL_CAP", str(cap))
        else:
            monkeypatch.delenv("PP_SMALL_CELL_CAP", raising=False)
        with pp.System.batch([x[0] for x in starts], [x[1] for x in starts], [x[2] for x in starts], prm, seed=12,
                             trials_per_cell=16, threads_per_cell=tpc) as s:
            s.do_ensemble(25)
            return [s.state(k) for k in range(s.R)], s.counters_all()

    ref = run(32, 0)
    for tpc, cap in ((16, 0), (16, 1)):
        got = run(tpc, cap)
        for a, b in zip(ref[0], got[0]):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (tpc, cap)
        for key in ("tries", "accepted", "bonds"):
            assert np.array_equal(ref[1][key], got[1][key]), (key, tpc, cap)


@pytest.mark.parametrize("trials_per_cell", [1, 3])
def test_fixed_trials_per_cell(pp, oracle, trials_per_cell):
    """s trials per non-empty cell and pass: the tracked energy equals a full
    recompute, the snapshot's bonds equal the oracle's, and every pass of
    every sweep makes a multiple of s attempts."""
    r, O, box = _start(oracle, 4000, 0.5, 3)
    with pp.System(r, O, box, T=0.12, seed=42, trials_per_cell=trials_per_cell) as s:
        s.do_ensemble(60)
        c = s.counters()
        b, ov, tot, anyo = s.energy()
        rr, OO, bb = s.state()
    ob, oov, otot = oracle.energy(rr, OO, bb)
    assert not anyo and not oov.any()
    assert np.array_equal(b, ob) and c["bonds"] == tot == otot
    assert c["tries"][0] % trials_per_cell == 0
    assert 0.3 < c["accepted"][0] / c["tries"][0] < 0.99


def test_autotune_keeps_the_trajectory(pp, oracle):
    """The tuner's sweeps are real sweeps: a tuned context ends where an
    untuned one with the same seed and sweep count ends, bit for bit."""
    r, O, box = _start(oracle, 3000, 0.4, 9)
    with pp.System(r, O, box, T=0.11, seed=5) as a:
        t, b, us = a.autotune(2)
        done = a.counters()["sweeps"]
        a.do_ensemble(5)
        sa = a.state()
    assert t in (32, 64) and b in (64, 128, 256) and us > 0 and done >= 6
    with pp.System(r, O, box, T=0.11, seed=5) as c:
        c.do_ensemble(done + 5)
        sc = c.state()
    for x, y in zip(sa, sc):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("trials_per_cell", [0, 4])
def test_uneven_batch(pp, oracle, trials_per_cell):
    """24 replicas of uneven N, T and density in one context: every replica's
    tracked energy equals a full recompute, without overlaps, and its bonds
    equal the oracle's."""
    rs, Os, bs, ps = pp.uneven_replicas(24, seed=3, n_range=(200, 1500))
    with pp.System.batch(rs, Os, bs, ps, seed=9, trials_per_cell=trials_per_cell) as s:
        s.autotune(1)
        s.do_ensemble(60)
        c = s.counters_all()
        for k in range(len(rs)):
            b, ov, tot, anyo = s.energy(k)
            assert not anyo and tot == c["bonds"][k], k
            rr, OO, bb = s.state(k)
            ob, oov, otot = oracle.energy(rr, OO, bb)
            assert np.array_equal(b, ob) and otot == tot, k
        assert (c["tries"][:, 0] > 0).all()


@pytest.mark.parametrize("ensemble,n,tile,tiles_lo,tiles_hi", [
    ("nvt", 1500, 512, 1, 64),      # 512-cell tiles, a look-back over all of them (26^3 cells, 35 tiles)
    ("nvt", 4000, 512, 65, 320),    # 512-cell tiles beyond 64 (38^3 cells, 108 tiles)
    ("nvt", 20000, 2048, 65, 1 << 30),  # 2048-cell tiles in three steps (64^3 cells, 128 tiles)
    ("npt", 1000, 512, 1, 64),      # NPT grids reserve room to grow: 24^3 -> 28^3 cells, 43 tiles
    ("npt", 3000, 512, 65, 320),    # 34^3 -> 38^3 cells, 108 tiles
    ("npt", 10000, 2048, 65, 1 << 30)])  # 52^3 -> 58^3 cells, 96 tiles of 2048
def test_large_grid_scan(pp, oracle, ensemble, n, tile, tiles_lo, tiles_hi):
    """Dilute systems (rho=0.05) whose cell layouts take each form of the
    rebuild's scan: 512-cell tiles with a look-back over at most 64 tiles or
    over more, up to the limit of 320 tiles, and 2048-cell tiles in three
    steps.  The form follows from the cells the context reserves (NPT
    replicas reserve room for their grids to grow), which the test reads
    back.  The tracked energy equals a full recompute, the cell populations
    check, and the energies of the evolved snapshot equal the oracle's bit
    for bit (NPT: through trial rebuilds and accepted volume moves)."""
    r, O, box = _start(oracle, n, 0.05, 8)
    prm = dict(T=0.11)
    if ensemble == "npt":
        prm = dict(T=0.11, ensemble=pp.NPT, P=0.02, rescale_factor_max=0.002)
    with pp.System(r, O, box, seed=5, trials_per_cell=4, **prm) as s:
        cap, got_tile, tiles = s.layout()
        assert got_tile == tile and tiles_lo <= tiles <= tiles_hi, (cap, got_tile, tiles)
        assert tiles == -(-cap // tile)
        s.do_ensemble(60)
        c = s.counters()
        b, ov, tot, anyo = s.energy()
        assert not anyo and c["bonds"] == tot
        rr, OO, bb = s.state()
        ob, oov, otot = oracle.energy(rr, OO, bb)
        assert np.array_equal(b, ob) and not oov.any()
        s.check_energy()
        if ensemble == "npt":
            assert c["accepted"][1] > 0
