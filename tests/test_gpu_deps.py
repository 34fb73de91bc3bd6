"""The dependency-driven sweep (launch form PP_LAUNCH_DEPS: the 8 colour passes
in one launch, each cell starting once its neighbour cells of earlier passes
are done) against the per-pass launches: trajectories and counters must be
identical bit for bit, for one system and for batches, in NVT and NPT, with
one trial per member and with 16 trials per cell, and at the full N=32768
workload, whose final bonds must also equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PER_PASS, DEPS = 1, 3


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _run(pp, r, O, box, sweeps, **kw):
    with pp.System(r, O, box, **kw) as s:
        s.do_ensemble(sweeps)
        return s.state(), s.counters()


def _same(a, b):
    (sa, ca), (sb, cb) = a, b
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    assert ca == cb


@pytest.mark.parametrize("tpc,block,trials", [(32, 128, 16), (32, 64, 0), (64, 256, 16), (64, 128, 0)])
def test_single_system_matches_per_pass(pp, oracle, tpc, block, trials):
    """N=3000, rho=0.4, T=0.11, 25 sweeps."""
    r, O, box = oracle.generate(3000, 0.4, seed=9)
    kw = dict(T=0.11, seed=99, trials_per_cell=trials, threads_per_cell=tpc, block_size=block)
    _same(_run(pp, r, O, box, 25, launch=PER_PASS, **kw), _run(pp, r, O, box, 25, launch=DEPS, **kw))


@pytest.mark.parametrize("trials", [0, 16])
def test_batch_matches_per_pass(pp, trials):
    """12 uneven replicas (N 500-8000, T 0.08-0.20, rho 0.05-0.60), 20 sweeps:
    the several-replica form of the kernel, each replica with its own grid and
    colour order."""
    rs, Os, bs, ps = pp.uneven_replicas(12, seed=5)
    out = []
    for launch in (PER_PASS, DEPS):
        with pp.System.batch(rs, Os, bs, ps, seed=7, trials_per_cell=trials, launch=launch) as s:
            s.do_ensemble(20)
            out.append(([s.state(k) for k in range(s.R)], s.counters_all()))
    for a, b in zip(out[0][0], out[1][0]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    for key in out[0][1]:
        assert np.array_equal(np.asarray(out[0][1][key]), np.asarray(out[1][1][key])), key


def test_npt_matches_per_pass(pp, oracle):
    """NPT, N=2000 from rho=0.2, P=0.5, T=0.15, 16 trials per cell (16 volume
    attempts per sweep), 40 sweeps: the passes between volume moves."""
    r, O, box = oracle.generate(2000, 0.2, seed=4)
    kw = dict(T=0.15, seed=5, trials_per_cell=16, ensemble=pp.NPT, P=0.5, rescale_factor_max=0.05)
    a = _run(pp, r, O, box, 40, launch=PER_PASS, **kw)
    b = _run(pp, r, O, box, 40, launch=DEPS, **kw)
    _same(a, b)
    assert a[1]["accepted"][1] > 0


def test_full_size_matches_per_pass_and_oracle(pp, oracle):
    """N=32768, rho=0.5, T=0.12, 16 trials per cell, 60 sweeps from the
    seeded start: identical to the per-pass launches, and the final
    per-particle bonds equal the oracle's."""
    r, O, box = pp.generate(32768, 0.5, 1)
    kw = dict(T=0.12, seed=1, trials_per_cell=16)
    a = _run(pp, r, O, box, 60, launch=PER_PASS, **kw)
    b = _run(pp, r, O, box, 60, launch=DEPS, **kw)
    _same(a, b)
    (r1, O1, b1), c = b
    bonds, ov, tot = oracle.energy(r1, O1, b1, 0.12, 0.92)
    assert not ov.any() and tot == c["bonds"]


def test_over_full_cell_is_reported(pp):
    """A slab of 64 particles per cell (as in the per-pass test): the teams
    of those cells refuse them and still publish their cells, so the other
    cells' waits end and the sweep reports PP_ERR_CELL_CAP."""
    g = np.stack(np.meshgrid(np.arange(8.0), np.arange(16.0), np.arange(16.0), indexing="ij"), -1).reshape(-1, 3) + 0.5
    O = np.broadcast_to(np.eye(3), (len(g), 3, 3)).copy()
    box = np.array([32.0, 16.0, 16.0])
    with pp.System(g, O, box, T=0.1, cell_occupancy=15.0, seed=3, launch=DEPS) as s:
        with pytest.raises(pp.PPError) as e:
            s.do_ensemble(1)
        assert e.value.status == 5  # PP_ERR_CELL_CAP


def test_sixteen_lanes_are_refused(pp, oracle):
    r, O, box = oracle.generate(500, 0.3, seed=1)
    with pytest.raises(pp.PPError) as e:
        pp.System(r, O, box, T=0.1, threads_per_cell=16, launch=DEPS)
    assert e.value.status == 1  # PP_ERR_ARG
