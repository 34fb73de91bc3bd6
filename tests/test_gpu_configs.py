"""The batched, learned and NPT workloads at their full sizes, against the
oracle (the reference energy routine, MC.c:272-309):

* 256 uneven NVT replicas, N log-uniform in [500, 8000], T uniform in
  [0.08, 0.20], rho uniform in [0.05, 0.60], autotuned launch shape, in both
  trial schedules;
* learned-proposal NVT at N=65536, rho=0.3, T=0.09, with a network trained
  on the device and attached to 80 % of the colour passes;
* 16 NPT replicas of N=131072 at P=0.02, T=0.11, rescale_factor_max=0.01,
  from rho=0.05 (volume moves, MC.c:450-513).

After the run every replica's per-particle bonds and overlap flags equal the
oracle's on the GPU's own snapshot, bit for bit, and the incrementally
tracked energy equals the full recompute."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _check_replica(s, oracle, k, tracked):
    b, ov, tot, anyo = s.energy(k)
    assert not anyo and tot == tracked, (k, tot, tracked)
    rr, OO, bb = s.state(k)
    ob, oov, otot = oracle.energy(rr, OO, bb)
    assert np.array_equal(b, ob) and not oov.any() and otot == tot, k
    return rr, bb


@pytest.mark.parametrize("trials_per_cell", [0, 16])
def test_uneven_256_replicas_full_size(pp, oracle, trials_per_cell):
    rs, Os, bs, ps = pp.uneven_replicas(256, seed=1)
    n = np.array([len(r) for r in rs])
    T = np.array([p["T"] for p in ps])
    assert n.min() < 800 and n.max() > 6000 and T.min() < 0.09 and T.max() > 0.19
    with pp.System.batch(rs, Os, bs, ps, seed=31, trials_per_cell=trials_per_cell) as s:
        s.autotune(2)
        s.do_ensemble(200)
        c = s.counters_all()
        assert (c["tries"][:, 0] > 0).all() and (c["accepted"][:, 0] > 0).all()
        for k in range(len(rs)):
            _check_replica(s, oracle, k, c["bonds"][k])
        s.check_energy()


def test_learned_gel_full_size(pp, oracle):
    r, O, box = pp.generate(65536, 0.3, 4)
    with pp.System(r, O, box, T=0.09, seed=41) as s, pp.ProposalNet(seed=3) as net:
        s.do_ensemble(200)
        l0, l1 = s.train(net, steps=300, batch=4096, pos_noise=0.03)
        assert np.isfinite(l0) and l1 < l0
        s.attach(net, 0.8)
        s.reset_counters()
        s.do_ensemble(200)
        tried, accepted = s.learn_counters()
        c = s.counters()
        assert tried > 0 and 0 < accepted < tried
        assert c["tries"][0] > tried  # plain passes ran too
        _check_replica(s, oracle, 0, c["bonds"])
        s.check_energy()


def test_npt_16_replicas_full_size(pp, oracle):
    starts = [pp.generate(131072, 0.05, 100 + k) for k in range(16)]
    prm = [dict(T=0.11, ensemble=pp.NPT, P=0.02, rescale_factor_max=0.01)] * 16
    with pp.System.batch([x[0] for x in starts], [x[1] for x in starts], [x[2] for x in starts], prm, seed=51,
                         trials_per_cell=16) as s:
        s.do_ensemble(100)
        c = s.counters_all()
        assert (c["tries"][:, 1] == 100 * 16).all()
        assert c["accepted"][:, 1].sum() > 0
        for k in range(16):
            rr, bb = _check_replica(s, oracle, k, c["bonds"][k])
            # the box the oracle saw is the replica's current (possibly
            # anisotropic, Q7) box, and its volume is the one tracked
            assert np.isclose(np.prod(bb), c["volume"][k], rtol=1e-12)
        s.check_energy()
