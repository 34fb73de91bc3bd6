"""Box trials evaluated on the current cell lists (no trial rebuild while
the lists' cells, rescaled with the trial, stay at least 1 + delta wide),
with their energy taken from the sweep's close-pair list (the default: the
step's attempts enqueued ahead on that list, the host taking over from the
first attempt that needs more), from the same list with the host reading
every attempt's paths first (PP_NPT_STEPWISE=1), pair by pair from the cell
records (PP_NPT_RECORD_TRIAL=1) or per particle (PP_NPT_PARTICLE_TRIAL=1),
against the form that regroups every trial into the spare buffer
(PP_NPT_FULL_TRIAL=1): the same trajectories bit for bit --
positions,
orientations, boxes, volumes and move counters -- in NPT (MC.c:450-513) and
with box-shape moves (MC.c:515-585), for one system and for a batch that
mixes NPT and NVT replicas, whose trials alternate between the two forms as
grids change.  The final states' bonds equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _run(pp, monkeypatch, form, make, steps, lx=None):
    monkeypatch.setenv("PP_NPT_FULL_TRIAL", "1" if form == "full" else "0")
    monkeypatch.setenv("PP_NPT_PARTICLE_TRIAL", "1" if form == "particles" else "0")
    monkeypatch.setenv("PP_NPT_RECORD_TRIAL", "1" if form == "records" else "0")
    monkeypatch.setenv("PP_NPT_STEPWISE", "1" if form == "stepwise" else "0")
    with make() as s:
        if lx:
            s.set_lx_move(*lx)
        s.do_ensemble(steps)
        states = [s.state(k) for k in range(s.R)]
        c = s.counters_all()
        energies = [s.energy(k) for k in range(s.R)]
    return states, c, energies


def _same_runs(a, b):
    for (ra, Oa, ba), (rb, Ob, bb) in zip(a[0], b[0]):
        assert np.array_equal(ra, rb) and np.array_equal(Oa, Ob) and np.array_equal(ba, bb)
    for key in ("tries", "accepted", "bonds", "box", "volume"):
        assert np.array_equal(a[1][key], b[1][key]), key


NPT = dict(ensemble=5)


@pytest.mark.parametrize("case", ["npt_dense", "npt_dilute", "mixed_batch", "box_shape", "coincident_paths"])
def test_box_trials_on_current_lists_keep_the_trajectory(pp, oracle, monkeypatch, case):
    if case == "npt_dense":
        r, O, box = oracle.generate(500, 0.3, seed=21)
        make = lambda: pp.System(r, O, box, T=0.2, P=0.2, rescale_factor_max=0.05, seed=3, **NPT)
        steps, lx = 400, None
    elif case == "npt_dilute":
        r, O, box = oracle.generate(3000, 0.05, seed=22)
        make = lambda: pp.System(r, O, box, T=0.11, P=0.02, rescale_factor_max=0.01, seed=4, trials_per_cell=4,
                                 **NPT)
        steps, lx = 150, None
    elif case == "mixed_batch":
        starts = [oracle.generate(n, rho, seed=30 + k) for k, (n, rho) in
                  enumerate([(400, 0.3), (900, 0.1), (300, 0.5), (600, 0.2)])]
        prm = [dict(T=0.2, P=0.3, rescale_factor_max=0.05, **NPT), dict(T=0.15),
               dict(T=0.25, P=0.5, rescale_factor_max=0.03, **NPT), dict(T=0.12, P=0.05, rescale_factor_max=0.05, **NPT)]
        make = lambda: pp.System.batch([x[0] for x in starts], [x[1] for x in starts], [x[2] for x in starts], prm,
                                       seed=5, trials_per_cell=2)
        steps, lx = 300, None
    elif case == "coincident_paths":
        # in every attempt one replica's large steps need trial rebuilds (and
        # the per-particle energy) while the others' small ones take the
        # close-pair list or the records: each trial counted by one path only
        starts = [oracle.generate(n, rho, seed=50 + k) for k, (n, rho) in enumerate([(500, 0.4), (500, 0.4), (700, 0.3)])]
        prm = [dict(T=0.2, P=0.3, rescale_factor_max=0.6, **NPT), dict(T=0.2, P=0.3, rescale_factor_max=0.01, **NPT),
               dict(T=0.15, P=0.2, rescale_factor_max=0.02, **NPT)]
        make = lambda: pp.System.batch([x[0] for x in starts], [x[1] for x in starts], [x[2] for x in starts], prm,
                                       seed=8, trials_per_cell=4)
        steps, lx = 250, None
    else:
        r, O, box = oracle.generate(400, 0.1, seed=23)
        make = lambda: pp.System(r, O, box, T=0.15, seed=6)
        steps, lx = (400, (2.0, 0.85 * box[1], 1.15 * box[1]))
    full = _run(pp, monkeypatch, "full", make, steps, lx)
    fast = _run(pp, monkeypatch, "pairs", make, steps, lx)
    _same_runs(full, fast)
    _same_runs(full, _run(pp, monkeypatch, "records", make, steps, lx))
    _same_runs(full, _run(pp, monkeypatch, "particles", make, steps, lx))
    _same_runs(full, _run(pp, monkeypatch, "stepwise", make, steps, lx))
    c = fast[1]
    assert c["accepted"][:, 1].sum() > 0 and c["accepted"][:, 1].sum() < c["tries"][:, 1].sum()
    for k, (rr, OO, bb) in enumerate(fast[0]):
        b, ov, tot, anyo = fast[2][k]
        ob, oov, otot = oracle.energy(rr, OO, bb)
        assert np.array_equal(b, ob) and not anyo and tot == otot == c["bonds"][k], k
