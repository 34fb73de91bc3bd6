"""Edge cases of the C ABI on the GPU: the smallest systems, ragged batches
with an empty replica, boxes too small for a checkerboard grid, cells over
their capacity, and arguments out of range - each either runs correctly or
is refused with the status the header documents."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pp():
    import patchyparticles_amd as pp
    pp.lib()
    return pp



def test_single_particle(pp):
    r = np.array([[1.0, 2.0, 3.0]])
    O = np.eye(3)[None]
    with pp.System(r, O, np.full(3, 6.0), T=0.1, seed=1) as s:
        s.do_ensemble(100)
        c = s.counters()
        assert c["tries"][0] == 100 and c["bonds"] == 0
        assert 0 < c["accepted"][0] <= 100
        b, ov, tot, anyo = s.energy()
        assert tot == 0 and not anyo


def test_ragged_batch_with_an_empty_replica(pp, oracle):
    r1, O1, b1 = oracle.generate(300, 0.3, seed=1)
    r2, O2, b2 = oracle.generate(50, 0.1, seed=2)
    rs = [r1, np.zeros((0, 3)), r2]
    Os = [O1, np.zeros((0, 3, 3)), O2]
    bs = [b1, np.full(3, 8.0), b2]
    prm = [dict(T=0.2)] * 3
    with pp.System.batch(rs, Os, bs, prm, seed=3) as s:
        s.do_ensemble(50)
        c = s.counters_all()
        assert list(c["tries"][:, 0]) == [50 * 300, 0, 50 * 50]
        for k in (0, 2):
            rr, OO, bb = s.state(k)
            b, ov, tot, anyo = s.energy(k)
            ob, oov, otot = oracle.energy(rr, OO, bb)
            assert np.array_equal(b, ob) and tot == c["bonds"][k]


def test_box_too_small_for_the_grid(pp):
    # a checkerboard needs at least 4 even cells of side >= 1.12 per axis
    r = np.array([[0.5, 0.5, 0.5], [2.5, 2.5, 2.5]])
    O = np.array([np.eye(3)] * 2)
    with pytest.raises(pp.PPError) as e:
        pp.System(r, O, np.full(3, 4.0), T=0.1)
    assert e.value.status == 4  # PP_ERR_BOX


@pytest.mark.parametrize("kw,status", [(dict(T=0.0), 1), (dict(T=0.1, ensemble=3), 8),
                                       (dict(T=0.1, trials_per_cell=17), 1), (dict(T=0.1, threads_per_cell=48), 1),
                                       (dict(T=0.1, disp_max=-0.1), 1)])
def test_arguments_out_of_range_are_refused(pp, oracle, kw, status):
    r, O, box = oracle.generate(100, 0.1, seed=4)
    with pytest.raises(pp.PPError) as e:
        pp.System(r, O, box, **kw)
    assert e.value.status == status


def test_replica_index_out_of_range(pp, oracle):
    r, O, box = oracle.generate(100, 0.1, seed=5)
    with pp.System(r, O, box, T=0.1) as s:
        for f in (lambda: s.state(1), lambda: s.energy(-1), lambda: s.counters(3)):
            with pytest.raises(pp.PPError) as e:
                f()
            assert e.value.status == 1


def test_cell_over_capacity_is_reported(pp):
    """Cells of about 60 particles (cell_occupancy = 40 on a dense lattice):
    more than a sweep cell's 24 slots, refused with PP_ERR_CELL_CAP instead of
    being swept."""
    a = 1.05
    g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), -1).reshape(-1, 3) * a + 0.3
    O = np.broadcast_to(np.eye(3), (len(g), 3, 3)).copy()
    box = np.full(3, 16 * a)
    with pytest.raises(pp.PPError) as e:
        with pp.System(g, O, box, T=0.1, cell_occupancy=40.0) as s:
            s.do_ensemble(1)
    assert e.value.status == 5  # PP_ERR_CELL_CAP


def test_small_team_capacity_is_reported(pp):
    """An anisotropic lattice (spacings 1.5, 1, 1) in a box of 4^3 sweep
    cells of side 3 puts exactly 2 x 3 x 3 = 18 particles in every cell,
    whatever the grid shift: 32-lane teams sweep it (the tracked energy stays
    exact), 16-lane teams (one member per lane) refuse it with
    PP_ERR_CELL_CAP."""
    g = np.stack(np.meshgrid(np.arange(8) * 1.5, np.arange(12) * 1.0, np.arange(12) * 1.0, indexing="ij"),
                 -1).reshape(-1, 3) + 0.5  # exact differences: contacts at exactly 1, no overlap
    O = np.broadcast_to(np.eye(3), (len(g), 3, 3)).copy()
    box = np.full(3, 12.0)
    with pp.System(g, O, box, T=0.1, cell_occupancy=12.0, trials_per_cell=16) as s:
        assert s.grid()[0] == [4, 4, 4]
        s.do_ensemble(3)
        b, ov, tot, anyo = s.energy()
        assert s.counters()["bonds"] == tot and not anyo
    with pytest.raises(pp.PPError) as e:
        with pp.System(g, O, box, T=0.1, cell_occupancy=12.0, trials_per_cell=16, threads_per_cell=16,
                       block_size=64) as s:
            s.do_ensemble(1)
    assert e.value.status == 5  # PP_ERR_CELL_CAP


def test_sweep_grids_keep_same_colour_cells_apart(pp, oracle):
    """The checkerboard's invariant, on the grids the library picks for many
    boxes (and after NPT volume moves): an even number of cells per axis, at
    least 4, each of side >= 1 + delta, so two cells of one colour are at
    least one cell side apart across the periodic boundary too, and no two
    particles moved in one pass can interact."""
    rng = np.random.default_rng(12)
    for _ in range(20):
        n = int(rng.integers(50, 400))
        rho = float(rng.uniform(0.05, 0.6))
        r, O, box = oracle.generate(n, rho, seed=int(rng.integers(1 << 30)))
        delta = float(rng.uniform(0.05, 0.2))
        with pp.System(r, O, box, T=0.2, kf_delta=delta) as s:
            nside, side = s.grid()
            assert all(k % 2 == 0 and k >= 4 for k in nside), (nside, box)
            assert min(side) >= 1 + delta, (side, delta)
    r, O, box = oracle.generate(300, 0.3, seed=3)
    with pp.System(r, O, box, T=0.2, ensemble=pp.NPT, P=2.0, rescale_factor_max=0.1, seed=2) as s:
        for _ in range(20):
            s.do_ensemble(50)
            nside, side = s.grid()
            assert all(k % 2 == 0 and k >= 4 for k in nside) and min(side) >= 1.12
        assert s.counters()["volume"] < np.prod(box)  # the box did shrink


@pytest.mark.parametrize("learned_fraction", [0.0, 1.0])
def test_over_full_cell_is_reported_by_every_pass(pp, learned_fraction):
    """A simple-cubic slab of spacing 1 filling half of a 32 x 16 x 16 box,
    swept on 8 x 4 x 4 cells of side 4: the mean occupancy (16) passes the
    host's check, but every cell inside the slab holds 64 particles, more
    than a team of 32 lanes holds one per lane.  Plain passes and learned
    passes (learned_fraction = 1: no plain pass runs) both refuse such a
    cell with PP_ERR_CELL_CAP instead of moving a member no lane holds."""
    g = np.stack(np.meshgrid(np.arange(8.0), np.arange(16.0), np.arange(16.0), indexing="ij"), -1).reshape(-1, 3) + 0.5
    O = np.broadcast_to(np.eye(3), (len(g), 3, 3)).copy()
    box = np.array([32.0, 16.0, 16.0])
    with pp.System(g, O, box, T=0.1, cell_occupancy=15.0, seed=3) as s, pp.ProposalNet(seed=1) as net:
        assert s.grid()[0] == [8, 4, 4]
        if learned_fraction > 0:
            s.attach(net, learned_fraction)
        with pytest.raises(pp.PPError) as e:
            s.do_ensemble(1)
        assert e.value.status == 5  # PP_ERR_CELL_CAP
