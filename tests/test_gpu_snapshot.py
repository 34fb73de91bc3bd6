"""Asynchronous snapshots (pp_snapshot_start / _finish): the state at the
start, in particle-index order exactly as pp_download returns it, while the
sweeps enqueued after the start run on; for one system, a replica of a batch
and a grand-canonical replica; and the refusals of a second start and of a
finish without a start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_snapshot_is_the_state_at_its_start(pp, oracle):
    r, O, box = oracle.generate(4000, 0.4, seed=2)
    with pp.System(r, O, box, T=0.12, seed=9, trials_per_cell=16) as s:
        s.do_ensemble(10)
        want = s.state()
        s.snapshot_start()
        s.do_ensemble(20)
        got = s.snapshot_finish()
        _same(got, want)
        later = s.state()
        assert not np.array_equal(later[0], want[0])
        # a second snapshot after the first one finished
        s.snapshot_start()
        _same(s.snapshot_finish(), later)


def test_snapshot_of_a_batch_replica_and_a_gc_replica(pp, oracle):
    rs, Os, bs, ps = pp.uneven_replicas(3, seed=3)
    r, O, box = oracle.generate(300, 0.2, seed=4)
    rs.append(r)
    Os.append(O)
    bs.append(box)
    ps.append(dict(T=0.15, ensemble=pp.GC, activity=0.1, gc_n_max=1500))
    with pp.System.batch(rs, Os, bs, ps, seed=6) as s:
        s.do_ensemble(30)
        for k in (1, 3):
            want = s.state(k)
            s.snapshot_start(k)
            s.do_ensemble(5)
            _same(s.snapshot_finish(), want)


def test_snapshot_misuse_is_refused(pp, oracle):
    r, O, box = oracle.generate(500, 0.3, seed=5)
    with pp.System(r, O, box, T=0.12, seed=1) as s:
        with pytest.raises(pp.PPError) as e:
            s.snapshot_finish()
        assert e.value.status == 1
        s.snapshot_start()
        with pytest.raises(pp.PPError) as e:
            s.snapshot_start()
        assert e.value.status == 1
        s.snapshot_finish()
        with pytest.raises(pp.PPError):
            s.snapshot_start(7)
