"""Grand-canonical insertions and removals (Ensemble = 1, MC_add_remove,
MC.c:342-448), made cell-local for the checkerboard:

* hard spheres (patches that never bond) at low activity follow the virial
  expansion of the density in the activity, rho = z - 2 B2 z^2 +
  (8 B2^2 - 3 B3) z^3, B2 = 2 pi / 3, B3 = 5 pi^2 / 18 (closed form);
* the GPU's energies of the evolving system equal the oracle's bit for bit
  and the tracked energy equals a full recompute;
* trajectories do not depend on the launch shape, and an NVT replica in the
  same batch follows the trajectory it follows alone;
* <N/V> and <E/N> at T = 0.2, z = 0.2 match the reference program's GC runs
  (tests/golden/gc_ref.json) within 3 combined standard errors;
* too small a GC_N_max is refused, and a checkpoint resumes bit for bit."""
import math

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _mean_se(x):
    return float(np.mean(x)), float(np.std(x, ddof=1) / math.sqrt(len(x)))


def test_hard_spheres_follow_the_virial_expansion(pp, oracle):
    z, L, R = 0.005, 30.0, 64
    B2, B3 = 2 * math.pi / 3, 5 * math.pi ** 2 / 18
    rho_th = z - 2 * B2 * z * z + (8 * B2 * B2 - 3 * B3) * z ** 3
    r, O, box = oracle.generate(int(rho_th * L ** 3), rho_th, seed=3)
    box = np.full(3, L)
    prm = dict(T=1.0, ensemble=pp.GC, activity=z, gc_n_max=3000, kf_cosmax=1.0, disp_max=0.3, theta_max=0.5)
    with pp.System.batch([r] * R, [O] * R, [box] * R, [prm] * R, seed=12) as s:
        s.do_ensemble(300)
        n = []
        for _ in range(200):
            s.do_ensemble(10)
            n.append(s.counters_all()["n"])
        c = s.counters_all()
        assert (c["gc_accepted"] > 0).all()
    rho = np.array(n).mean(axis=0) / L ** 3
    m, se = _mean_se(rho)
    assert abs(m - rho_th) < 4 * se + 2e-3 * rho_th, (m, se, rho_th)


def test_gc_energies_match_the_oracle(pp, oracle):
    r, O, box = oracle.generate(300, 0.2, seed=4)
    with pp.System(r, O, box, T=0.15, ensemble=pp.GC, activity=0.1, gc_n_max=1500, seed=5) as s:
        for _ in range(4):
            s.do_ensemble(50)
            c = s.counters()
            b, ov, tot, anyo = s.energy()
            assert not anyo and tot == c["bonds"] and len(b) == c["n"]
            rr, OO, bb = s.state()
            ob, oov, otot = oracle.energy(rr, OO, bb)
            assert np.array_equal(b, ob) and not oov.any() and otot == tot
            bl = s.bond_list()
            assert [len(x) for x in bl] == list(ob)
        s.check_energy()
        assert c["n"] != 300 and c["gc_tries"][0] > 0 and c["gc_tries"][1] > 0


@pytest.mark.parametrize("tpc", [32, 64])
def test_gc_trajectory_independent_of_launch_shape(pp, oracle, tpc):
    r, O, box = oracle.generate(400, 0.3, seed=6)
    kw = dict(T=0.15, ensemble=pp.GC, activity=0.1, gc_n_max=2000, seed=8)
    with pp.System(r, O, box, threads_per_cell=16, block_size=64, **kw) as s:
        s.do_ensemble(60)
        ref = s.state(), s.counters()
    with pp.System(r, O, box, threads_per_cell=tpc, **kw) as s:
        s.do_ensemble(60)
        got = s.state(), s.counters()
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    assert ref[1] == got[1]


def test_nvt_replica_unaffected_by_a_gc_replica(pp, oracle):
    ra, Oa, ba = oracle.generate(800, 0.3, seed=9)
    rb, Ob, bb = oracle.generate(200, 0.1, seed=10)
    with pp.System(ra, Oa, ba, T=0.12, seed=4) as s:
        s.do_ensemble(40)
        alone = s.state()
    prm = [dict(T=0.12), dict(T=0.2, ensemble=pp.GC, activity=0.2, gc_n_max=1500)]
    with pp.System.batch([ra, rb], [Oa, Ob], [ba, bb], prm, seed=4) as s:
        s.do_ensemble(40)
        both = s.state(0)
        assert s.counters(1)["n"] != 200
    for a, b in zip(alone, both):
        assert np.array_equal(a, b)


def test_gc_matches_reference_runs(pp, tmp_path):
    ref = golden("gc_ref.json")
    from test_gpu_stats import _reference_start
    r, O, box = _reference_start(pp, tmp_path, ref["n0"], ref["rho0"])
    R = 32
    prm = dict(T=ref["T"], ensemble=pp.GC, activity=ref["activity"], gc_n_max=ref["n_max"])
    with pp.System.batch([r] * R, [O] * R, [box] * R, [prm] * R, seed=77) as s:
        rho, e = [], []
        V = float(np.prod(box))
        for _ in range(ref["sweeps"] // (2 * ref["print_every"])):
            c = s.counters_all()
            rho.append(c["n"] / V)
            e.append(-c["bonds"] / np.maximum(c["n"], 1))
            s.do_ensemble(ref["print_every"])
    rho, e = np.array(rho), np.array(e)
    cut = int(ref["discard"] * len(rho))
    for got, want, want_se in ((rho, ref["density"], ref["density_se"]), (e, ref["E_per_N"], ref["E_per_N_se"])):
        m, se = _mean_se(got[cut:].mean(axis=0))
        assert abs(m - want) < 3 * math.hypot(se, want_se), (m, se, want, want_se)


def test_too_small_gc_n_max_is_refused(pp, oracle):
    r, O, box = oracle.generate(300, 0.2, seed=4)
    with pytest.raises(pp.PPError) as e:
        pp.System(r, O, box, T=0.2, ensemble=pp.GC, activity=0.1, gc_n_max=200)
    assert e.value.status == 1
    # room for the particles but not for one GC phase (a cell of each colour)
    with pp.System(r, O, box, T=0.2, ensemble=pp.GC, activity=0.1, gc_n_max=310) as s:
        with pytest.raises(pp.PPError) as e:
            s.do_ensemble(1)
        assert e.value.status == 1


def test_gc_checkpoint_resumes_bit_for_bit(pp, oracle, tmp_path):
    r, O, box = oracle.generate(300, 0.2, seed=4)
    kw = dict(T=0.15, ensemble=pp.GC, activity=0.1, gc_n_max=1500, seed=13)
    path = str(tmp_path / "gc.chk")
    with pp.System(r, O, box, **kw) as s:
        s.do_ensemble(30)
        s.checkpoint_save(path)
        s.do_ensemble(30)
        ref = s.state(), s.counters()
    with pp.System(r, O, box, **kw) as s:
        s.checkpoint_load(path)
        s.do_ensemble(30)
        got = s.state(), s.counters()
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    for k in ("n", "bonds", "gc_tries", "gc_accepted", "tries", "accepted"):
        assert ref[1][k] == got[1][k], k
