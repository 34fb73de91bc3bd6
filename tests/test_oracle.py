"""The CPU oracle pinned against the reference's own outputs (golden vectors
written by tests/golden/make_golden.py from the reference code) and against
what the mathematics fixes (strict thresholds, symmetry, closed forms)."""
import hashlib
import json
import math
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, golden, golden_files


def test_pair_known_answers(oracle):
    g = golden("pair_kat.npz")
    got = np.array([oracle.pair(g["rp"][i], g["Op"][i], g["rq"][i], g["Oq"][i], g["box"][i])
                    for i in range(len(g["expected"]))])
    assert np.array_equal(got, g["expected"])
    # the set covers all three outcomes
    assert set(np.unique(g["expected"])) == {-1, 0, 1}


@pytest.mark.parametrize("name", golden_files("snap_"))
@pytest.mark.parametrize("brute", [False, True])
def test_snapshot_energies(oracle, name, brute):
    g = golden(name)
    b, ov, tot = oracle.energy(g["r"], g["O"], g["box"], float(g["delta"]), float(g["cosmax"]), brute=brute)
    assert np.array_equal(b, g["bonds"]), name
    assert np.array_equal(ov, g["overlap"]), name
    if not ov.any():
        assert tot == g["bonds"].sum() // 2


@pytest.mark.parametrize("name", ["traj_nvt.npz", "traj_npt.npz"])
def test_trajectory_bit_exact(oracle, name):
    g = golden(name)
    kw = json.loads(str(g["params"]))
    sim = oracle.Sim(g["r_in"], g["O_in"], g["box_in"], T=kw["T"], seed=kw["seed"],
                     ensemble=kw.get("ensemble", 0), pressure=kw.get("pressure", 0.0),
                     rescale_factor_max=kw.get("rescale_factor_max", 0.0))
    sim.run(int(g["steps"]), int(g["print_every"]))
    r, O, box = sim.state()
    assert np.array_equal(box, g["box"])
    assert np.array_equal(r, g["r"])
    assert np.array_equal(O, g["O"])
    assert sim.energy == float(g["energy"])


def test_generator_bytes(oracle):
    with open(os.path.join(GOLDEN, "generator.json")) as f:
        cases = json.load(f)
    for c in cases:
        r, O, box = oracle.generate(c["n"], c["rho"])
        path = os.path.join(tempfile.mkdtemp(), "generated.rrr")
        oracle.rrr_save(path, 0, box, r, O, generator_header=True)
        with open(path, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == c["sha256"]


def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32-10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert list(oracle.philox(ctr, key)) == out


def _aligned(u, patch, base):
    """rotation taking base patch `patch` onto unit vector u"""
    b = base[patch]
    v = np.cross(b, u)
    s, c = np.linalg.norm(v), float(np.dot(b, u))
    if s < 1e-14:
        return np.eye(3)
    k = v / s
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + s * K + (1 - c) * K @ K


def test_strict_thresholds(oracle):
    """d^2 < 1 overlaps and d^2 < (1+delta)^2 may bond, both strict
    (MC.c:240-241); p sits at the origin so that r_q - r_p is exact."""
    base = oracle.base_patches()
    box = np.array([10.0, 10.0, 10.0])
    u = np.array([1.0, 0.0, 0.0])
    Op = _aligned(u, 0, base)   # a patch of p along +x
    Oq = _aligned(-u, 2, base)  # a patch of q along -x
    rp = np.zeros(3)
    below_one = np.nextafter(1.0, 0.0)
    assert oracle.pair(rp, Op, u, Oq, box) == 1               # d^2 = 1: contact, bonded
    assert oracle.pair(rp, Op, below_one * u, Oq, box) == -1  # d^2 < 1: overlap
    rc2 = 1.12 * 1.12                                         # SQR(1 + delta)
    x = np.sqrt(rc2)
    while x * x >= rc2:
        x = np.nextafter(x, 0.0)
    while np.nextafter(x, 2.0) ** 2 < rc2:
        x = np.nextafter(x, 2.0)
    assert x * x < rc2 <= np.nextafter(x, 2.0) ** 2
    assert oracle.pair(rp, Op, x * u, Oq, box) == 1                       # just inside the shell
    assert oracle.pair(rp, Op, np.nextafter(x, 2.0) * u, Oq, box) == 0    # at/after the shell edge
    # q turned so that no patch faces p: no bond at any distance in the shell
    Oq_away = _aligned(u, 2, base)
    for d in (1.0, 1.05, float(x)):
        assert oracle.pair(rp, Op, d * u, Oq_away, box) == 0
    # through the periodic boundary: p near 0, q near L
    assert oracle.pair(rp, Op, np.array([-1.05, 0, 0]) + box * np.array([1, 0, 0]), Op, box) in (0, 1)
    assert oracle.pair(rp, Op, np.array([10.0 - 0.5, 0, 0]), Oq, box) == -1


def test_pair_symmetry_and_images(oracle):
    rng = np.random.default_rng(5)
    box = np.array([6.0, 7.0, 8.0])
    for _ in range(400):
        rp = rng.uniform(-20, 20, 3)
        rq = rp + rng.uniform(-1.3, 1.3, 3)
        Op = oracle.gram_schmidt(rng.normal(size=(3, 3)))
        Oq = oracle.gram_schmidt(rng.normal(size=(3, 3)))
        a = oracle.pair(rp, Op, rq, Oq, box)
        assert a == oracle.pair(rq, Oq, rp, Op, box)
        # shifting either particle by whole boxes changes nothing
        k = rng.integers(-4, 5, 3)
        assert a == oracle.pair(rp + k * box, Op, rq, Oq, box)


def test_gram_schmidt_orthonormal(oracle):
    rng = np.random.default_rng(2)
    for _ in range(100):
        O = oracle.gram_schmidt(rng.normal(size=(3, 3)))
        assert np.allclose(O @ O.T, np.eye(3), atol=1e-12)


def test_two_particle_bonding_probability(oracle):
    """Two particles in a periodic box V: P(bonded) = Vb e^{1/T} /
    (Vb e^{1/T} + V - 4pi/3 - Vb), Vb = 16 pi (d^3+3d^2+3d)(1-cosmax)^2/3
    (orientation-averaged bonding volume, the expression of avb.c:26 with four
    patches).  Checked on the serial sampler within 4 block standard errors."""
    delta, cosmax, T, L = 0.12, 0.92, 0.135, 4.0
    Vb = 16 * math.pi * (delta ** 3 + 3 * delta ** 2 + 3 * delta) * (1 - cosmax) ** 2 / 3
    V = L ** 3
    w = Vb * math.exp(1 / T)
    p_exact = w / (w + V - 4 * math.pi / 3 - Vb)
    r = np.array([[1.0, 1.0, 1.0], [3.0, 3.0, 3.0]])
    O = np.array([np.eye(3), np.eye(3)])
    sim = oracle.Sim(r, O, np.array([L, L, L]), T=T, disp_max=0.8, theta_max=1.5, seed=11)
    series = []
    for _ in range(4000):
        sim.sweeps(50)
        series.append(-sim.energy)
    x = np.array(series)
    blocks = x.reshape(40, -1).mean(axis=1)
    m, se = blocks.mean(), blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(m - p_exact) < 4 * se + 1e-3, (m, se, p_exact)
