"""The allocation-free energy entry point (pp_energy_snapshot_noalloc)
against pp_energy_snapshot, which allocates a context per call: the same
per-particle bonds, overlap flags and totals, bit for bit, on consecutive
calls that share the fixed device array -- a large snapshot, then a smaller
one, then the reference's snapshots, an overlapping one and a dilute box
whose cells the call must enlarge to fit its grid -- and sizes beyond its
limit refused before any device work."""
import numpy as np
import pytest

from conftest import golden, golden_files

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp


def _same(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2] and a[3] == b[3]


def test_noalloc_matches_the_allocating_call_on_consecutive_calls(pp, oracle):
    big = oracle.generate(32768, 0.5, seed=11)
    small = oracle.generate(1000, 0.1, seed=12)
    for r, O, box in (big, small):
        want = pp.energy_snapshot(r, O, box)
        got = pp.energy_snapshot_noalloc(r, O, box)
        _same(got, want)
        ob, oov, otot = oracle.energy(r, O, box)
        assert np.array_equal(got[0], ob) and got[2] == otot and not got[3]
    # the same small call again: nothing of the previous calls leaks in
    _same(pp.energy_snapshot_noalloc(*small), pp.energy_snapshot(*small))


@pytest.mark.parametrize("name", golden_files("snap_"))
def test_noalloc_on_the_reference_snapshots(pp, name):
    g = golden(name)
    args = (g["r"], g["O"], g["box"], float(g["delta"]), float(g["cosmax"]))
    b, ov, tot, anyo = pp.energy_snapshot_noalloc(*args)
    assert np.array_equal(b, g["bonds"]) and np.array_equal(ov, g["overlap"].astype(bool))
    assert anyo == bool(g["overlap"].any())
    _same((b, ov, tot, anyo), pp.energy_snapshot(*args))


def test_noalloc_enlarges_cells_of_a_dilute_box(pp, oracle):
    """rho = 0.004, N = 20000: cells of side 1.12 would number 3.5 million,
    beyond the call's 2^18; larger cells give the same energies."""
    r, O, box = oracle.generate(20000, 0.004, seed=13)
    assert np.prod(np.floor(box / 1.12)) > (1 << 18)
    got = pp.energy_snapshot_noalloc(r, O, box)
    _same(got, pp.energy_snapshot(r, O, box))
    ob, oov, otot = oracle.energy(r, O, box)
    assert np.array_equal(got[0], ob) and got[2] == otot


def test_noalloc_refuses_sizes_beyond_its_limit(pp):
    n = pp.NOALLOC_MAX_N + 1
    r = np.zeros((n, 3))
    O = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    with pytest.raises(pp.PPError) as e:
        pp.energy_snapshot_noalloc(r, O, np.full(3, 100.0))
    assert e.value.status == 1  # PP_ERR_ARG
    # a call within the limit still works afterwards
    r, O, box = pp.generate(500, 0.2, 3)
    _same(pp.energy_snapshot_noalloc(r, O, box), pp.energy_snapshot(r, O, box))
