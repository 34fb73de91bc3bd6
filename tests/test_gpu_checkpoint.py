"""Checkpoint and resume: a run split by pp_checkpoint_save / _load into a
fresh context continues bit for bit (positions, orientations, box, tracked
volume, counters), because the state is kept exactly and the sweep index is
the Philox counter."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _states(s):
    return [s.state(k) for k in range(s.R)]


def _same(a, b):
    for (r1, O1, b1), (r2, O2, b2) in zip(a, b):
        assert np.array_equal(r1, r2) and np.array_equal(O1, O2) and np.array_equal(b1, b2)


@pytest.mark.parametrize("case", ["nvt", "npt", "batch", "learned"])
def test_resume_is_bit_exact(pp, oracle, tmp_path, case):
    tpc = 0
    if case == "batch":
        rs, Os, bs, ps = pp.uneven_replicas(6, seed=3, n_range=(200, 900))
        tpc = 4
    else:
        r, O, box = oracle.generate(600, 0.4, seed=5)
        prm = dict(T=0.15)
        if case == "npt":
            prm = dict(T=0.2, ensemble=pp.NPT, P=0.2, rescale_factor_max=0.05)
        rs, Os, bs, ps = [r], [O], [box], [prm]
    net = None
    if case == "learned":
        net = pp.ProposalNet(seed=4)

    def make():
        s = pp.System.batch(rs, Os, bs, ps, seed=77, trials_per_cell=tpc)
        if net is not None:
            s.attach(net, 0.5, 0.1)
        return s

    with make() as s:
        s.do_ensemble(150)
        full = _states(s)
        cf = s.counters_all()
    path = tmp_path / "run.ppchk"
    with make() as s:
        s.do_ensemble(70)
        s.checkpoint_save(path)
    with make() as s:
        s.checkpoint_load(path)
        assert s.counters()["sweeps"] == 70
        s.do_ensemble(80)
        _same(full, _states(s))
        c = s.counters_all()
    for key in ("tries", "accepted", "bonds", "volume", "box"):
        assert np.array_equal(np.asarray(cf[key]), np.asarray(c[key])), key
    if net is not None:
        net.close()


def test_checkpoint_refuses_a_mismatched_context(pp, oracle, tmp_path):
    r, O, box = oracle.generate(300, 0.3, seed=2)
    path = tmp_path / "a.ppchk"
    with pp.System(r, O, box, T=0.2, seed=1) as s:
        s.do_ensemble(3)
        s.checkpoint_save(path)
    with pp.System(r, O, box, T=0.2, seed=2) as s:
        with pytest.raises(pp.PPError, match="seed"):
            s.checkpoint_load(path)
    with pp.System(r, O, box, T=0.25, seed=1) as s:
        with pytest.raises(pp.PPError, match="parameters"):
            s.checkpoint_load(path)
    (tmp_path / "bad.ppchk").write_bytes(b"not a checkpoint at all")
    with pp.System(r, O, box, T=0.2, seed=1) as s:
        with pytest.raises(pp.PPError, match="not a checkpoint"):
            s.checkpoint_load(tmp_path / "bad.ppchk")
