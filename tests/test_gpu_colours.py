"""The checkerboard invariant, a debug check of the concurrent colour passes:
no particle lies within the bond range 1 + delta of a particle in another cell
of the same colour.  Checked live, after the rebuild and after every colour
pass of real sweeps, in every launch form, for one system, a batch and NPT;
and the count itself against a brute-force numpy count of the same definition
on grids that break the invariant (an odd number of cells per axis, cells
narrower than the bond range)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _conflicts(r, box, nside, delta=0.12):
    """Particles within the bond range of a particle in another cell of the
    same colour, with the library's cell and minimum-image arithmetic."""
    r = np.asarray(r, float)
    box = np.asarray(box, float)
    inv = 1.0 / box
    t = r * inv
    t = t - np.floor(t)
    idx = np.minimum((t * np.asarray(nside)).astype(np.int64), np.asarray(nside) - 1)
    col = ((idx[:, 0] & 1) << 2) | ((idx[:, 1] & 1) << 1) | (idx[:, 2] & 1)
    rc = 1.0 + delta
    rc2 = rc * rc
    hit = np.zeros(len(r), bool)
    for j in range(len(r)):
        d = r - r[j]
        d = d - box * np.rint(d * inv)
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        other = (col == col[j]) & np.any(idx != idx[j], axis=1)
        hit[j] = np.any(other & (d2 < rc2))
    return int(hit.sum())


@pytest.mark.parametrize("launch,trials", [(1, 16), (1, 0), (3, 16), (2, 0)])
def test_live_checks_find_no_conflict(pp, oracle, launch, trials):
    """N=4000, rho=0.5, T=0.12, 20 sweeps with checks after the rebuild and
    every pass (per-pass launches) or every sweep launch (cooperative and
    dependency-driven forms)."""
    r, O, box = oracle.generate(4000, 0.5, seed=13)
    with pp.System(r, O, box, T=0.12, seed=3, trials_per_cell=trials, launch=launch) as s:
        s.colour_check(True)
        s.do_ensemble(20)
        total, checks = s.colour_check(False)
        assert checks == 20 * (9 if launch == 1 else 2)
        assert total == 0
        assert s.colour_conflicts() == 0


def test_live_checks_batch_and_npt(pp, oracle):
    """Six uneven NVT replicas and two NPT replicas whose boxes move, 30
    sweeps of live checks."""
    rs, Os, bs, ps = pp.uneven_replicas(6, seed=8)
    for k, (n, rho) in enumerate([(700, 0.3), (1200, 0.1)]):
        r, O, box = oracle.generate(n, rho, seed=40 + k)
        rs.append(r)
        Os.append(O)
        bs.append(box)
        ps.append(dict(T=0.15, ensemble=pp.NPT, P=0.3, rescale_factor_max=0.05))
    with pp.System.batch(rs, Os, bs, ps, seed=4, trials_per_cell=4) as s:
        s.colour_check(True)
        s.do_ensemble(30)
        total, checks = s.colour_check(None)
        assert checks == 30 * 9 * s.R and total == 0
        assert s.counters_all()["accepted"][6:, 1].sum() > 0


@pytest.mark.parametrize("nside", [(5, 5, 5), (20, 20, 20), (4, 6, 3)])
def test_conflict_count_on_grids_that_break_the_invariant(pp, oracle, nside):
    """N=1500, rho=0.3 (box 17.1): 5 cells per axis (cells 0 and 4 touch
    across the boundary with the same parity), 20 per axis (cells of 0.86,
    narrower than 1.12) and a mixed grid: the device's count equals numpy's
    and is positive; the library's own grid has none."""
    r, O, box = oracle.generate(1500, 0.3, seed=17)
    with pp.System(r, O, box, T=0.12, seed=5) as s:
        s.do_ensemble(3)
        rr, _, bb = s.state()
        want = _conflicts(rr, bb, nside)
        assert want > 0
        assert s.colour_conflicts(nside=nside) == want
        assert s.colour_conflicts() == 0
        n_even = [v - v % 2 for v in s.grid()[0]]
        assert s.colour_conflicts(nside=n_even) == _conflicts(rr, bb, n_even) == 0
